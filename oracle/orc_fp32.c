/*
 * ORACLE -- test infrastructure only (timing).
 *
 * FP32 build of whhd_oracle.c (the oracle's line transforms, Gram tables,
 * projection and recovery): every `double` of that source -- arrays, scalars,
 * lifting tables -- becomes `float`, and the Makefile compiles this unit with
 * -fsingle-precision-constant, so literals are single precision too.  It is
 * the like-for-like FP32 CPU baseline that bench.py times next to the FP64
 * oracle; the FP64 build (liboracle.so) stays the parity checker.
 *
 * The system headers come first, so the macro below reaches only the
 * oracle's own source and headers.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define double float
#include "whhd_oracle.c"
