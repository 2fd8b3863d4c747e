// Per-pair lifting steps of the registered biorthogonal CDF wavelets, written for
// 64-lane wavefronts.  One lane holds one (approximation, detail) pair; neighbour
// pairs come from cross-lane moves, and a lane whose neighbour lies outside the
// wavefront reads its own value.  Windows of 64 pairs overlap by the wavelet's
// halo so every written pair only depends on pairs inside its window.
//
// The statement sequence of each update, the index/sign reflections and the
// (type, bc) pairings follow WaveletPotentials/cuWaveletCompute.cuh:180-1551 and
// the dispatch table WaveletPotentials/cuWHHDForward.cu:744-1101.
#pragma once

// WHHD_HOST_EMULATION compiles the lane-independent parts (index functions,
// lifting statements on pair arrays) as plain host C++, for CPU checks of the
// line-parallel algorithm; the cross-lane moves exist only on the device.
#ifdef WHHD_HOST_EMULATION
#define WHHD_DEV inline
#else
#include <hip/hip_runtime.h>
#define WHHD_DEV __device__ __forceinline__
#endif

#include "wt_tables.h"

namespace whhd {

#ifndef WHHD_HOST_EMULATION
// Whole-wavefront shifts by one lane on the VALU (DPP wave_shr:1 / wave_shl:1);
// a lane without a source keeps its own value.  Shifts by m lanes chain m of
// them, so edge lanes may hold values that differ from a direct m-lane
// shuffle; they are halo lanes, whose results are never written.
WHHD_DEV int dpp_shr1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
WHHD_DEV int dpp_shl1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x130, 0xf, 0xf, false); }

WHHD_DEV float shr1(float v) { return __int_as_float(dpp_shr1(__float_as_int(v))); }
WHHD_DEV float shl1(float v) { return __int_as_float(dpp_shl1(__float_as_int(v))); }
WHHD_DEV double shr1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_shr1((int)(b & 0xffffffffLL)), hi = dpp_shr1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
WHHD_DEV double shl1(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_shl1((int)(b & 0xffffffffLL)), hi = dpp_shl1((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// lane i receives lane i-m (lane_up) / lane i+m (lane_dn)
template <class T>
WHHD_DEV T lane_up(T v, int m) {
    for (int i = 0; i < m; ++i) v = shr1(v);
    return v;
}
template <class T>
WHHD_DEV T lane_dn(T v, int m) {
    for (int i = 0; i < m; ++i) v = shl1(v);
    return v;
}
// value of lane `src` (wave-uniform, compile-time) broadcast to every lane
WHHD_DEV float lane_get(float v, int src) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), src));
}
WHHD_DEV double lane_get(double v, int src) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
#endif  // WHHD_HOST_EMULATION

// ---------------------------------------------------------------------------
// Index / sign reflections.  Forward forms receive a = 2k, d = 2k+1; inverse
// forms a = d = k; L is the level dimension.
// ---------------------------------------------------------------------------
struct IxHalfN {  // half-sample symmetric (cuWT_CDF3_5_n / _n_i)
    template <class T>
    WHHD_DEV static void fwd(int& a, int& d, T& sa, T& sd, int L) {
        if (a < 0) a = -1 - a; else if (a >= L) a = 2 * L - 1 - a;
        if (d < 0) d = -1 - d; else if (d >= L) d = 2 * L - 1 - d;
        sa = T(1); sd = T(1);
    }
    template <class T>
    WHHD_DEV static void inv(int& a, int& d, T& sa, T& sd, int L) {
        sa = T(1); sd = T(1);
        if (a < 0) { a = -1 - a; sd = T(-1); } else if (a >= L) { a = 2 * L - 1 - a; sd = T(-1); }
        d = a;
    }
};
struct IxHalfD {  // half-sample antisymmetric (cuWT_CDF3_5_d / _d_i)
    template <class T>
    WHHD_DEV static void fwd(int& a, int& d, T& sa, T& sd, int L) {
        sa = T(1); sd = T(1);
        if (d < 0) { d = -1 - d; sd = T(-1); } else if (d >= L) { d = 2 * L - 1 - d; sd = T(-1); }
        if (a < 0) { a = -1 - a; sa = T(-1); } else if (a >= L) { a = 2 * L - 1 - a; sa = T(-1); }
    }
    template <class T>
    WHHD_DEV static void inv(int& a, int& d, T& sa, T& sd, int L) {
        sa = T(1); sd = T(1);
        if (a < 0) { a = -1 - a; sa = T(-1); } else if (a >= L) { a = 2 * L - 1 - a; sa = T(-1); }
        d = a;
    }
};
struct IxWholeN {  // whole-sample symmetric (cuWT_CDF2_6_n, cuWT_CDF4_4_n and _i forms)
    template <class T>
    WHHD_DEV static void fwd(int& a, int& d, T& sa, T& sd, int L) {
        if (a < 0) { a = -a; d = -d; } else if (a >= L) { a = 2 * L - a; d = 2 * L - d; }
        sa = T(1); sd = T(1);
    }
    template <class T>
    WHHD_DEV static void inv(int& a, int& d, T& sa, T& sd, int L) {
        sa = T(1); sd = T(1);
        if (a < 0) { a = -a; d = -d - 1; } else if (a >= L) { a = 2 * L - a; d = 2 * L - d - 1; }
    }
};
struct IxWholeD26 {  // cuWT_CDF2_6_d / _d_i
    template <class T>
    WHHD_DEV static void fwd(int& a, int& d, T& sa, T& sd, int L) {
        if (a < 0) a = -a; else if (a > L) a = 2 * L - a;
        sa = T(1); sd = T(1);
        if (d < 0) { d = -d; sd = T(-1); } else if (d >= L) { d = 2 * L - d; sd = T(-1); }
    }
    template <class T>
    WHHD_DEV static void inv(int& a, int& d, T& sa, T& sd, int L) {
        sa = T(1); sd = T(1);
        if (a < 0) { a = -a; d = -d - 1; sd = T(-1); }
        else if (a >= L) { a = 2 * L - a; d = 2 * L - d - 1; sd = T(-1); }
    }
};
struct IxWholeZ26 {  // cuWT_CDF2_6_z / _z_i
    template <class T>
    WHHD_DEV static void fwd(int& a, int& d, T& sa, T& sd, int L) {
        if (a < 0) { a = -a; d = -d; } else if (a >= L) { a = 2 * L - a; d = 2 * L - d; }
        sa = (a == 0 || a == L) ? T(2) : T(1);
        sd = T(1);
    }
    template <class T>
    WHHD_DEV static void inv(int& a, int& d, T& sa, T& sd, int L) {
        sa = (a == 0 || a == L) ? T(2) : T(1);
        sd = T(1);
        if (a < 0) { a = -a; d = -d - 1; } else if (a >= L) { a = 2 * L - a; d = 2 * L - d - 1; }
    }
};
struct IxWholeD44 {  // cuWT_CDF4_4_d / _d_i
    template <class T>
    WHHD_DEV static void fwd(int& a, int& d, T& sa, T& sd, int L) {
        sa = T(1); sd = T(1);
        if (a < 0) { a = -a; d = -d; sd = T(-1); sa = T(-1); }
        else if (a >= L) { a = 2 * L - a; d = 2 * L - d; sd = T(-1); sa = T(-1); }
        if (a == L) sa = T(1);
    }
    template <class T>
    WHHD_DEV static void inv(int& a, int& d, T& sa, T& sd, int L) {
        sa = T(1); sd = T(1);
        if (a < 0) { a = -a; d = -d - 1; sd = T(-1); sa = T(-1); }
        else if (a >= L) { d = 2 * L - d - 1; sd = T(-1); }
        if (a > L) { a = 2 * L - a; sa = T(-1); }
    }
};

// ---------------------------------------------------------------------------
// Lifting updates (forward `fwd`, inverse `inv`).
// ---------------------------------------------------------------------------
#define C(x) T(tab::x)

struct UpCdf35 {
    static constexpr int halo = 3;
    template <class T>
    WHHD_DEV static void fwd(T& a, T& d, T sa, T sd) {
        d *= sd;
        a = sa * a + C(Q5_fu_1[0]) * lane_up(d, 1);
        d += C(Q5_fp[0]) * lane_dn(a, 1) + C(Q5_fp[1]) * a;
        a += C(Q5_fu_2[4]) * lane_up(d, 2);
        a += C(Q5_fu_2[3]) * lane_up(d, 1);
        a += C(Q5_fu_2[2]) * d;
        a += C(Q5_fu_2[1]) * lane_dn(d, 1);
        a += C(Q5_fu_2[0]) * lane_dn(d, 2);
        a *= C(Q5_scl[0]);
        d *= C(Q5_scl[1]);
    }
    template <class T>
    WHHD_DEV static void inv(T& a, T& d, T sa, T sd) {
        a = sa * a * C(Q5_scl[1]);
        d = sd * d * C(Q5_scl[0]);
        a -= C(Q5_fu_2[4]) * lane_up(d, 2);
        a -= C(Q5_fu_2[3]) * lane_up(d, 1);
        a -= C(Q5_fu_2[2]) * d;
        a -= C(Q5_fu_2[1]) * lane_dn(d, 1);
        a -= C(Q5_fu_2[0]) * lane_dn(d, 2);
        d -= C(Q5_fp[1]) * a + C(Q5_fp[0]) * lane_dn(a, 1);
        a -= C(Q5_fu_1[0]) * lane_up(d, 1);
    }
};

struct UpCdf37 {
    static constexpr int halo = 4;
    template <class T>
    WHHD_DEV static void fwd(T& a, T& d, T sa, T sd) {
        d *= sd;
        a = sa * a + C(Q5_fu_1[0]) * lane_up(d, 1);
        d += C(Q5_fp[0]) * lane_dn(a, 1) + C(Q5_fp[1]) * a;
        a += C(Q7_fu_2[6]) * lane_up(d, 3);
        a += C(Q7_fu_2[5]) * lane_up(d, 2);
        a += C(Q7_fu_2[4]) * lane_up(d, 1);
        a += C(Q7_fu_2[3]) * d;
        a += C(Q7_fu_2[2]) * lane_dn(d, 1);
        a += C(Q7_fu_2[1]) * lane_dn(d, 2);
        a += C(Q7_fu_2[0]) * lane_dn(d, 3);
        a *= C(Q5_scl[0]);
        d *= C(Q5_scl[1]);
    }
    template <class T>
    WHHD_DEV static void inv(T& a, T& d, T sa, T sd) {
        a = sa * a * C(Q5_scl[1]);
        d = sd * d * C(Q5_scl[0]);
        a -= C(Q7_fu_2[6]) * lane_up(d, 3);
        a -= C(Q7_fu_2[5]) * lane_up(d, 2);
        a -= C(Q7_fu_2[4]) * lane_up(d, 1);
        a -= C(Q7_fu_2[3]) * d;
        a -= C(Q7_fu_2[2]) * lane_dn(d, 1);
        a -= C(Q7_fu_2[1]) * lane_dn(d, 2);
        a -= C(Q7_fu_2[0]) * lane_dn(d, 3);
        d -= C(Q5_fp[1]) * a + C(Q5_fp[0]) * lane_dn(a, 1);
        a -= C(Q5_fu_1[0]) * lane_up(d, 1);
    }
};

struct UpCdf26 {
    static constexpr int halo = 3;
    template <class T>
    WHHD_DEV static void fwd(T& a, T& d, T sa, T sd) {
        a *= sa;
        d += C(L6_fp[0]) * a + C(L6_fp[1]) * lane_dn(a, 1);
        d *= sd;
        a += C(L6_fu[5]) * lane_up(d, 3);
        a += C(L6_fu[4]) * lane_up(d, 2);
        a += C(L6_fu[3]) * lane_up(d, 1);
        a += C(L6_fu[2]) * d;
        a += C(L6_fu[1]) * lane_dn(d, 1);
        a += C(L6_fu[0]) * lane_dn(d, 2);
        a *= C(L6_scl[0]);
        d *= C(L6_scl[1]);
        a /= sa;
    }
    template <class T>
    WHHD_DEV static void inv(T& a, T& d, T sa, T sd) {
        d = sd * d * C(L6_scl[0]);
        a = sa * a * C(L6_scl[1]);
        a -= C(L6_fu[5]) * lane_up(d, 3);
        a -= C(L6_fu[4]) * lane_up(d, 2);
        a -= C(L6_fu[3]) * lane_up(d, 1);
        a -= C(L6_fu[2]) * d;
        a -= C(L6_fu[1]) * lane_dn(d, 1);
        a -= C(L6_fu[0]) * lane_dn(d, 2);
        d -= C(L6_fp[1]) * a + C(L6_fp[0]) * lane_dn(a, 1);
        a /= sa;
    }
};

struct UpCdf28 {
    static constexpr int halo = 4;
    template <class T>
    WHHD_DEV static void fwd(T& a, T& d, T sa, T sd) {
        a *= sa;
        d += C(L6_fp[0]) * a + C(L6_fp[1]) * lane_dn(a, 1);
        d *= sd;
        a += C(L8_fu[7]) * lane_up(d, 4);
        a += C(L8_fu[6]) * lane_up(d, 3);
        a += C(L8_fu[5]) * lane_up(d, 2);
        a += C(L8_fu[4]) * lane_up(d, 1);
        a += C(L8_fu[3]) * d;
        a += C(L8_fu[2]) * lane_dn(d, 1);
        a += C(L8_fu[1]) * lane_dn(d, 2);
        a += C(L8_fu[0]) * lane_dn(d, 3);
        a *= C(L6_scl[0]);
        d *= C(L6_scl[1]);
        a /= sa;
    }
    template <class T>
    WHHD_DEV static void inv(T& a, T& d, T sa, T sd) {
        d = sd * d * C(L6_scl[0]);
        a = sa * a * C(L6_scl[1]);
        a -= C(L8_fu[7]) * lane_up(d, 4);
        a -= C(L8_fu[6]) * lane_up(d, 3);
        a -= C(L8_fu[5]) * lane_up(d, 2);
        a -= C(L8_fu[4]) * lane_up(d, 1);
        a -= C(L8_fu[3]) * d;
        a -= C(L8_fu[2]) * lane_dn(d, 1);
        a -= C(L8_fu[1]) * lane_dn(d, 2);
        a -= C(L8_fu[0]) * lane_dn(d, 3);
        d -= C(L6_fp[1]) * a + C(L6_fp[0]) * lane_dn(a, 1);
        a /= sa;
    }
};

struct UpCdf44 {
    static constexpr int halo = 3;
    template <class T>
    WHHD_DEV static void fwd(T& a, T& d, T sa, T sd) {
        a *= sa;
        d *= sd;
        a += C(C4_fu_1[1]) * lane_up(d, 1) + C(C4_fu_1[0]) * d;
        d += C(C4_fp[1]) * a + C(C4_fp[0]) * lane_dn(a, 1);
        a += C(C4_fu_2[3]) * lane_up(d, 2);
        a += C(C4_fu_2[2]) * lane_up(d, 1);
        a += C(C4_fu_2[1]) * d;
        a += C(C4_fu_2[0]) * lane_dn(d, 1);
        a *= C(C4_scl[0]);
        d *= C(C4_scl[1]);
    }
    template <class T>
    WHHD_DEV static void inv(T& a, T& d, T sa, T sd) {
        d = sd * d * C(C4_scl_inv[1]);
        a = sa * a * C(C4_scl_inv[0]);
        a -= C(C4_fu_2[3]) * lane_up(d, 2);
        a -= C(C4_fu_2[2]) * lane_up(d, 1);
        a -= C(C4_fu_2[1]) * d;
        a -= C(C4_fu_2[0]) * lane_dn(d, 1);
        a *= sa;
        d *= sd;
        d -= C(C4_fp[1]) * a + C(C4_fp[0]) * lane_dn(a, 1);
        d *= sd;
        a -= C(C4_fu_1[1]) * lane_up(d, 1) + C(C4_fu_1[0]) * d;
    }
};

struct UpCdf46 {
    static constexpr int halo = 4;
    template <class T>
    WHHD_DEV static void fwd(T& a, T& d, T sa, T sd) {
        a *= sa;
        d *= sd;
        a += C(C4_fu_1[1]) * lane_up(d, 1) + C(C4_fu_1[0]) * d;
        d += C(C4_fp[1]) * a + C(C4_fp[0]) * lane_dn(a, 1);
        a += C(C6_fu_2[5]) * lane_up(d, 3);
        a += C(C6_fu_2[4]) * lane_up(d, 2);
        a += C(C6_fu_2[3]) * lane_up(d, 1);
        a += C(C6_fu_2[2]) * d;
        a += C(C6_fu_2[1]) * lane_dn(d, 1);
        a += C(C6_fu_2[0]) * lane_dn(d, 2);
        a *= C(C4_scl[0]);
        d *= C(C4_scl[1]);
    }
    template <class T>
    WHHD_DEV static void inv(T& a, T& d, T sa, T sd) {
        d = sd * d * C(C4_scl_inv[1]);
        a = sa * a * C(C4_scl_inv[0]);
        a -= C(C6_fu_2[5]) * lane_up(d, 3);
        a -= C(C6_fu_2[4]) * lane_up(d, 2);
        a -= C(C6_fu_2[3]) * lane_up(d, 1);
        a -= C(C6_fu_2[2]) * d;
        a -= C(C6_fu_2[1]) * lane_dn(d, 1);
        a -= C(C6_fu_2[0]) * lane_dn(d, 2);
        a *= sa;
        d *= sd;
        d -= C(C4_fp[1]) * a + C(C4_fp[0]) * lane_dn(a, 1);
        d *= sd;
        a -= C(C4_fu_1[1]) * lane_up(d, 1) + C(C4_fu_1[0]) * d;
    }
};

struct UpCdf20 {
    static constexpr int halo = 1;
    template <class T>
    WHHD_DEV static void fwd(T& a, T& d, T, T) {
        d += C(L6_fp[0]) * a + C(L6_fp[1]) * lane_dn(a, 1);
        a *= C(L6_scl[0]);
        d *= -C(L6_scl[1]);
    }
    template <class T>
    WHHD_DEV static void inv(T& a, T& d, T sa, T sd) {
        d = -sd * d * C(L6_scl[0]);
        a = sa * a * C(L6_scl[1]);
        d -= C(L6_fp[1]) * a + C(L6_fp[0]) * lane_dn(a, 1);
    }
};

struct UpCdf11 {
    static constexpr int halo = 0;
    template <class T>
    WHHD_DEV static void fwd(T& a, T& d, T sa, T sd) {
        a *= sa;
        d = sd * d + C(T1_fp[0]) * a;
        a += C(T1_fu[0]) * d;
        a *= C(T1_scl[0]);
        d *= C(T1_scl[1]);
    }
    template <class T>
    WHHD_DEV static void inv(T& a, T& d, T, T) {
        d = d * C(T1_scl[0]);
        a = a * C(T1_scl[1]);
        a -= C(T1_fu[0]) * d;
        d -= C(T1_fp[0]) * a;
    }
};

struct UpCdf17 {
    static constexpr int halo = 3;
    template <class T>
    WHHD_DEV static void fwd(T& a, T& d, T sa, T sd) {
        a *= sa;
        d = sd * d + C(T1_fp[0]) * a;
        a += C(T7_fu[6]) * lane_up(d, 3);
        a += C(T7_fu[5]) * lane_up(d, 2);
        a += C(T7_fu[4]) * lane_up(d, 1);
        a += C(T7_fu[3]) * d;
        a += C(T7_fu[2]) * lane_dn(d, 1);
        a += C(T7_fu[1]) * lane_dn(d, 2);
        a += C(T7_fu[0]) * lane_dn(d, 3);
        a *= C(T7_scl[0]);
        d *= C(T7_scl[1]);
    }
    template <class T>
    WHHD_DEV static void inv(T& a, T& d, T sa, T sd) {
        d = sd * d / C(T7_scl[1]);
        a = sa * a / C(T7_scl[0]);
        a -= C(T7_fu[6]) * lane_up(d, 3);
        a -= C(T7_fu[5]) * lane_up(d, 2);
        a -= C(T7_fu[4]) * lane_up(d, 1);
        a -= C(T7_fu[3]) * d;
        a -= C(T7_fu[2]) * lane_dn(d, 1);
        a -= C(T7_fu[1]) * lane_dn(d, 2);
        a -= C(T7_fu[0]) * lane_dn(d, 3);
        d -= C(T1_fp[0]) * a;
    }
};
#undef C

// ---------------------------------------------------------------------------
// A registered (wavelet, bc) combination: reflection, lifting, and the dense
// coarsest-level matrices (M x M, row-major) for both directions.
// ---------------------------------------------------------------------------
template <class Ix, class Up, const double* L10, const double* IL10, int M>
struct Wavelet {
    using ix = Ix;
    using up = Up;
    static constexpr int halo = Up::halo;
    static constexpr int m10 = M;
    static constexpr const double* l10 = L10;
    static constexpr const double* il10 = IL10;
};

// y = M x over lanes 0..M-1 (lane r gets row r); other lanes return row 0.
template <int M, const double* Mat, class T>
WHHD_DEV T apply_l10(T v, int lane) {
    const int r = lane < M ? lane : 0;
    T acc = T(0);
#pragma unroll
    for (int c = 0; c < M; ++c) {
        T xc = lane_get(v, c);
        T coef = T(0);
#pragma unroll
        for (int rr = 0; rr < M; ++rr) coef = (r == rr) ? T(Mat[rr * M + c]) : coef;
        acc += xc * coef;
    }
    return acc;
}

using W_Cdf35N = Wavelet<IxHalfN, UpCdf35, tab::QN5_L10, tab::iQN5_L10, 4>;
using W_Cdf35D = Wavelet<IxHalfD, UpCdf35, tab::QD5_L10, tab::iQD5_L10, 4>;
using W_Cdf37N = Wavelet<IxHalfN, UpCdf37, tab::QN7_L10, tab::iQN7_L10, 4>;
using W_Cdf37D = Wavelet<IxHalfD, UpCdf37, tab::QD7_L10, tab::iQD7_L10, 4>;
using W_Cdf26N = Wavelet<IxWholeN, UpCdf26, tab::LN6_L10, tab::iLN6_L10, 5>;
using W_Cdf26D = Wavelet<IxWholeD26, UpCdf26, tab::LD6_L10, tab::iLD6_L10, 5>;
using W_Cdf26Z = Wavelet<IxWholeZ26, UpCdf26, tab::LZ6_L10, tab::iLZ6_L10, 5>;
using W_Cdf28N = Wavelet<IxWholeN, UpCdf28, tab::LN8_L10, tab::iLN8_L10, 5>;
using W_Cdf28D = Wavelet<IxWholeD26, UpCdf28, tab::LD8_L10, tab::iLD8_L10, 5>;
using W_Cdf28Z = Wavelet<IxWholeZ26, UpCdf28, tab::LZ8_L10, tab::iLZ8_L10, 5>;
using W_Cdf44N = Wavelet<IxWholeN, UpCdf44, tab::CN4_L10, tab::iCN4_L10, 5>;
using W_Cdf44D = Wavelet<IxWholeD44, UpCdf44, tab::CD4_L10, tab::iCD4_L10, 5>;
using W_Cdf46N = Wavelet<IxWholeN, UpCdf46, tab::CN6_L10, tab::iCN6_L10, 5>;
using W_Cdf46D = Wavelet<IxWholeD44, UpCdf46, tab::CD6_L10, tab::iCD6_L10, 5>;
using W_Cdf20N = Wavelet<IxWholeN, UpCdf20, tab::_20_L10, tab::i20_L10, 5>;
using W_Cdf20D = Wavelet<IxWholeD26, UpCdf20, tab::_20_L10, tab::i20_L10, 5>;
using W_Cdf11N = Wavelet<IxHalfN, UpCdf11, tab::_11_L10, tab::i11_L10, 4>;
using W_Cdf11D = Wavelet<IxHalfD, UpCdf11, tab::_11_L10, tab::i11_L10, 4>;
using W_Cdf17N = Wavelet<IxHalfN, UpCdf17, tab::TN7_L10, tab::iTN7_L10, 4>;
using W_Cdf17D = Wavelet<IxHalfD, UpCdf17, tab::TD7_L10, tab::iTD7_L10, 4>;

}  // namespace whhd
