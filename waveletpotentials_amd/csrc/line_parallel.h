// Line-parallel multi-level transforms: one lane owns one whole line, a
// wavefront advances 64 lines in lockstep.  Every sample position is a
// compile-time constant (the level loop, the pair blocks and the reflections
// are unrolled for a fixed line length), so a level's approximations live in
// VGPRs, the boundary reflections and signs fold into constants, and the
// lifting statements become straight FMA chains.  Global traffic is one read
// and one write per sample: strided lines (x and z axes) give each load a
// full row of 64 consecutive floats, contiguous lines (y axis) give each lane
// a run of consecutive samples of its own line.
//
// The lifting statements themselves are the same templates as the windowed
// kernel (lifting.h), instantiated on `Arr`, a static array of pairs whose
// neighbour shift reads the pair's own value at the array ends.  Pairs are
// processed in blocks of B with `halo` extra pairs on each side, so every
// written pair only depends on pairs inside its block (the same cone argument
// as the 64-lane windows).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "lifting.h"

namespace whhd {
namespace lp {

template <class S, int NP>
struct Arr {
    S v[NP];
    WHHD_DEV Arr() {}
    WHHD_DEV Arr(double c) {
#pragma unroll
        for (int i = 0; i < NP; ++i) v[i] = S(c);
    }
    WHHD_DEV S& operator[](int i) { return v[i]; }
    WHHD_DEV const S& operator[](int i) const { return v[i]; }
#define WHHD_ARR_OP(op)                                                       \
    WHHD_DEV Arr& operator op##=(const Arr& o) {                              \
        _Pragma("unroll") for (int i = 0; i < NP; ++i) v[i] op## = o.v[i];    \
        return *this;                                                         \
    }                                                                         \
    WHHD_DEV friend Arr operator op(const Arr& x, const Arr& y) {             \
        Arr r;                                                                \
        _Pragma("unroll") for (int i = 0; i < NP; ++i) r.v[i] = x.v[i] op y.v[i]; \
        return r;                                                             \
    }
    WHHD_ARR_OP(+)
    WHHD_ARR_OP(-)
    WHHD_ARR_OP(*)
    WHHD_ARR_OP(/)
#undef WHHD_ARR_OP
    WHHD_DEV friend Arr operator-(const Arr& x) {
        Arr r;
#pragma unroll
        for (int i = 0; i < NP; ++i) r.v[i] = -x.v[i];
        return r;
    }
};

// pair i receives pair i-m / i+m; pairs without a source keep their own value
// (found by argument-dependent lookup from the lifting templates)
template <class S, int NP>
WHHD_DEV Arr<S, NP> lane_up(const Arr<S, NP>& x, int m) {
    Arr<S, NP> r;
#pragma unroll
    for (int i = 0; i < NP; ++i) r.v[i] = (i - m >= 0) ? x.v[i - m] : x.v[i];
    return r;
}
template <class S, int NP>
WHHD_DEV Arr<S, NP> lane_dn(const Arr<S, NP>& x, int m) {
    Arr<S, NP> r;
#pragma unroll
    for (int i = 0; i < NP; ++i) r.v[i] = (i + m < NP) ? x.v[i + m] : x.v[i];
    return r;
}

template <int I>
using ic = std::integral_constant<int, I>;

template <int B, int E, class F>
WHHD_DEV void static_for(F&& f) {
    if constexpr (B < E) {
        f(ic<B>{});
        static_for<B + 1, E>(f);
    }
}

constexpr int kBlock = 16;  // pairs written per block
// Lines per wavefront.  Fewer than 64 leaves lanes idle but doubles the
// number of wavefronts, which the memory system needs to overlap latency:
// a 256^3 pass has only ~65k lines, i.e. one 64-line wavefront per SIMD.
constexpr int kLinesPerWave = 64;

// ---- I/O of one lane's line ----------------------------------------------
// Strided lines (x and z axes): sample i of the lane's line is at
// base + i * step; a load of sample i across the wavefront reads 64
// consecutive floats.  Accesses go through buffer resources: the lane's line
// start is the per-lane byte offset and i * step the scalar offset, so no
// vector instructions are spent on addresses.  Lanes that must not store get
// an offset past the end of the buffer, which the hardware discards.
constexpr uint32_t kBufWord3 = 0x00020000;  // raw dword buffer, gfx9 family
constexpr uint32_t kOutOfRange = 0xF0000000u;

struct StridedIO {
    __amdgpu_buffer_rsrc_t rs;
    __amdgpu_buffer_rsrc_t rd;
    uint32_t voff_ld;  // byte offset of the lane's line start (loads)
    uint32_t voff_st;  // same, or out of range for lanes that do not store
    int step_bytes;
    int n;
    WHHD_DEV float load(int i) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff_ld, i * step_bytes, 0));
    }
    WHHD_DEV void store(int i, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rd, voff_st,
                                              i * step_bytes, 0);
    }
    // samples [START, START+CNT) of the line; each store is already one row
    template <int START, int CNT>
    WHHD_DEV void store_run(const float* vals) const {
#pragma unroll
        for (int v = 0; v < CNT; ++v) store(START + v, vals[v]);
    }
};

// Contiguous lines (y axis): the lane's line is n consecutive floats.  Loads
// fetch 16-byte runs of four samples (dword alignment suffices on gfx950) and
// the compiler folds the four element reads of a run into one load; the last
// run of a line is read sample by sample so nothing past the array is touched.
struct __attribute__((packed, aligned(4))) Quad {
    float v[4];
};
constexpr int kStageRun = 32;       // longest run written at once
constexpr int kStageRow = 64 + 1;   // LDS row: one value of each of the 64 lines

struct ContigIO {
    const float* src;
    float* dst;
    int64_t base;
    int64_t step;
    bool valid;
    int n;
    float* stage;          // this wavefront's LDS staging area, kStageRun rows
    int64_t first_line;    // line of lane 0
    int64_t total_lines;
    int lane;
    WHHD_DEV float load(int i) const {
        const int q = i & ~3;
        if (q + 3 < n) {
            const Quad r = *reinterpret_cast<const Quad*>(src + base + q);
            return r.v[i & 3];
        }
        return src[base + i];
    }
    // Samples [START, START+CNT) of every lane's line.  The run is transposed
    // through LDS so that consecutive lanes store consecutive 16-byte pieces
    // of one line (full 64-128 byte segments per line) instead of 64 lines
    // touched 16 bytes at a time.
    template <int START, int CNT>
    WHHD_DEV void store_run(const float* vals) const {
        static_assert(CNT <= kStageRun, "run too long");
        constexpr int NQ = (CNT + 3) / 4;          // 16-byte pieces per line
        constexpr int LPI = 64 / NQ;               // lines per pass (NQ divides 64 here)
        static_assert(64 % NQ == 0, "pieces per line must divide 64");
#pragma unroll
        for (int v = 0; v < CNT; ++v) stage[v * kStageRow + lane] = vals[v];
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        const int q = lane % NQ;
#pragma unroll
        for (int r = 0; r < NQ; ++r) {
            const int l = lane / NQ + r * LPI;
            const int64_t line = first_line + l;
            const bool mine = l < kLinesPerWave;
            float w[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int idx = 4 * q + e;
                w[e] = idx < CNT ? stage[idx * kStageRow + l] : 0.f;
            }
            if (mine && line < total_lines) {
                float* out = dst + line * n + START + 4 * q;
                if (4 * q + 3 < CNT) {
                    Quad qv;
#pragma unroll
                    for (int e = 0; e < 4; ++e) qv.v[e] = w[e];
                    *reinterpret_cast<Quad*>(out) = qv;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                        if (4 * q + e < CNT) out[e] = w[e];
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
    }
};

// ---- forward --------------------------------------------------------------
// One forward level: the 2^(l+1)+EXT samples come from `in` (or the line when
// TOP), approximations go to `out`, details to the line at [W, 2W-EXT).
template <class W, int l, int EXT, bool TOP, class IO>
WHHD_DEV void fwd_level(const IO& io, const float* in, float* out) {
    constexpr int H = W::halo;
    constexpr int L = 1 << (l + 1);
    constexpr int Wd = L / 2 + EXT;
    constexpr int NP = kBlock + 2 * H;
    constexpr int NB = (Wd + kBlock - 1) / kBlock;
    static_for<0, NB>([&](auto bi) {
        constexpr int b0 = decltype(bi)::value * kBlock;
        Arr<float, NP> A, D, SA, SD;
        static_for<0, NP>([&](auto ji) {
            constexpr int j = decltype(ji)::value;
            constexpr int k = b0 - H + j;
            int ai = 2 * k, di = 2 * k + 1;
            float sa, sd;
            W::ix::fwd(ai, di, sa, sd, L);
            di = di < 0 ? 0 : di;
            ai = ai < 0 ? 0 : (ai > L + EXT - 1 ? L + EXT - 1 : ai);
            di = di > L + EXT - 1 ? L + EXT - 1 : di;
            SA[j] = sa;
            SD[j] = sd;
            if constexpr (k < Wd + H) {
                if constexpr (TOP) {
                    A[j] = io.load(ai);
                    D[j] = io.load(di);
                } else {
                    A[j] = in[ai];
                    D[j] = in[di];
                }
            } else {
                A[j] = 0.f;
                D[j] = 0.f;
            }
        });
        W::up::fwd(A, D, SA, SD);
        constexpr int nd = (Wd - EXT - b0) < kBlock ? (Wd - EXT - b0) : kBlock;  // details here
        float run[kBlock];
        static_for<H, H + kBlock>([&](auto ji) {
            constexpr int j = decltype(ji)::value;
            constexpr int k = b0 - H + j;
            if constexpr (k < Wd) {
                out[k] = A[j];
                if constexpr (k < Wd - EXT) run[k - b0] = D[j];
            }
        });
        if constexpr (nd > 0) io.template store_run<Wd + b0, nd>(run);
    });
}

template <class W, int l, int EXT, class IO>
WHHD_DEV void fwd_levels_below(const IO& io, const float* in) {
    // `in` holds the 2^(l+1)+EXT approximations produced by level l+1
    if constexpr (l >= 2) {
        float out[(1 << l) + EXT];
        fwd_level<W, l, EXT, false>(io, in, out);
        fwd_levels_below<W, l - 1, EXT>(io, out);
    } else {
        constexpr int M = W::m10;
        float run[M];
        static_for<0, M>([&](auto ri) {
            constexpr int r = decltype(ri)::value;
            float acc = 0.f;
            static_for<0, M>([&](auto ci) {
                constexpr int c = decltype(ci)::value;
                acc += in[c] * float(W::l10[r * M + c]);
            });
            run[r] = acc;
        });
        io.template store_run<0, M>(run);
    }
}

template <class W, int LEV, int EXT, class IO>
WHHD_DEV void forward_line_lp(const IO& io) {
    if constexpr (LEV > 2) {
        float out[(1 << (LEV - 1)) + EXT];
        fwd_level<W, LEV - 1, EXT, true>(io, nullptr, out);
        fwd_levels_below<W, LEV - 2, EXT>(io, out);
    } else {
        float in[(1 << LEV) + EXT];
        static_for<0, (1 << LEV) + EXT>([&](auto i) { in[decltype(i)::value] = io.load(decltype(i)::value); });
        fwd_levels_below<W, 1, EXT>(io, in);
    }
}

// ---- inverse --------------------------------------------------------------
// One inverse level l: approximations [0, 2^l+EXT) from `in`, details from the
// line at [2^l+EXT, 2^(l+1)+EXT); reconstructed samples to `out`, or to the
// line when LAST.
template <class W, int l, int EXT, bool LAST, class IO>
WHHD_DEV void inv_level(const IO& io, const float* in, float* out) {
    constexpr int H = W::halo;
    constexpr int L = 1 << l;
    constexpr int Wd = L + EXT;
    constexpr int NP = kBlock + 2 * H;
    constexpr int NB = (Wd + kBlock - 1) / kBlock;
    static_for<0, NB>([&](auto bi) {
        constexpr int b0 = decltype(bi)::value * kBlock;
        Arr<float, NP> A, D, SA, SD;
        static_for<0, NP>([&](auto ji) {
            constexpr int j = decltype(ji)::value;
            constexpr int k = b0 - H + j;
            int ai = k, di = k;
            float sa, sd;
            W::ix::inv(ai, di, sa, sd, L);
            ai = ai < 0 ? 0 : (ai > Wd - 1 ? Wd - 1 : ai);
            di = di < 0 ? 0 : (di > L - 1 ? L - 1 : di);
            SA[j] = sa;
            SD[j] = sd;
            if constexpr (k < L + H) {
                A[j] = in[ai];
                D[j] = io.load(L + EXT + di);
            } else {
                A[j] = 0.f;
                D[j] = 0.f;
            }
        });
        W::up::inv(A, D, SA, SD);
        // samples [2*b0, 2*b0 + ns) are produced by this block
        constexpr int ns = (2 * L + EXT - 2 * b0) < 2 * kBlock ? (2 * L + EXT - 2 * b0) : 2 * kBlock;
        float run[2 * kBlock];
        static_for<H, H + kBlock>([&](auto ji) {
            constexpr int j = decltype(ji)::value;
            constexpr int k = b0 - H + j;
            if constexpr (k < L + EXT) {
                if constexpr (LAST) run[2 * (k - b0)] = A[j];
                else out[2 * k] = A[j];
            }
            if constexpr (k < L) {
                if constexpr (LAST) run[2 * (k - b0) + 1] = D[j];
                else out[2 * k + 1] = D[j];
            }
        });
        if constexpr (LAST && ns > 0) io.template store_run<2 * b0, ns>(run);
    });
}

template <class W, int l, int LEV, int EXT, class IO>
WHHD_DEV void inv_levels_from(const IO& io, const float* in) {
    // `in` holds the 2^l+EXT approximations of level l
    if constexpr (l == LEV - 1) {
        inv_level<W, l, EXT, true>(io, in, nullptr);
    } else {
        float out[(1 << (l + 1)) + EXT];
        inv_level<W, l, EXT, false>(io, in, out);
        inv_levels_from<W, l + 1, LEV, EXT>(io, out);
    }
}

template <class W, int LEV, int EXT, class IO>
WHHD_DEV void inverse_line_lp(const IO& io) {
    constexpr int M = W::m10;
    float c[M], a[M];
    static_for<0, M>([&](auto i) { c[decltype(i)::value] = io.load(decltype(i)::value); });
    static_for<0, M>([&](auto ri) {
        constexpr int r = decltype(ri)::value;
        float acc = 0.f;
        static_for<0, M>([&](auto ci) {
            constexpr int cc = decltype(ci)::value;
            acc += c[cc] * float(W::il10[r * M + cc]);
        });
        a[r] = acc;
    });
    if constexpr (LEV > 2) {
        inv_levels_from<W, 2, LEV, EXT>(io, a);
    } else {
        io.template store_run<0, M>(a);
    }
}

// ---- kernel -----------------------------------------------------------------
// Lines are numbered li = outer * lines_inner + inner; line li starts at
// inner * inner_stride + outer * outer_stride and its samples are `step` apart.
struct LpGeom {
    int64_t total_lines;
    int64_t total_elems;  // elements in the array (buffer range)
    int lines_inner;
    int64_t inner_stride;
    int64_t outer_stride;
    int64_t step;
};

template <class W, int LEV, int EXT, bool FWD, bool CONTIG>
__global__ void __launch_bounds__(256) lp_kernel(float* __restrict__ dst, const float* __restrict__ src,
                                                 LpGeom g) {
    const int lane_ = threadIdx.x & 63;
    const int64_t wave_first = ((int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kLinesPerWave;
    const int64_t li = wave_first + (lane_ % kLinesPerWave);
    const bool in_range = li < g.total_lines;
    const bool valid = in_range && lane_ < kLinesPerWave;  // lanes past kLinesPerWave mirror, never store
    const int64_t l = in_range ? li : 0;
    const int64_t outer = l / g.lines_inner;
    const int64_t inner = l - outer * g.lines_inner;
    constexpr int n = (1 << LEV) + EXT;
    if constexpr (CONTIG) {
        __shared__ float stage[4][kStageRun * kStageRow];
        const int lane = threadIdx.x & 63;
        ContigIO io{src, dst, inner * g.inner_stride + outer * g.outer_stride, g.step, valid, n,
                    stage[threadIdx.x >> 6], wave_first, g.total_lines, lane};
        if (FWD) forward_line_lp<W, LEV, EXT>(io);
        else inverse_line_lp<W, LEV, EXT>(io);
    } else {
        const uint32_t bytes = (uint32_t)(g.total_elems * 4);
        const uint32_t voff = (uint32_t)((inner * g.inner_stride + outer * g.outer_stride) * 4);
        StridedIO io{__builtin_amdgcn_make_buffer_rsrc((void*)src, 0, bytes, kBufWord3),
                     __builtin_amdgcn_make_buffer_rsrc((void*)dst, 0, bytes, kBufWord3),
                     voff, valid ? voff : kOutOfRange, (int)(g.step * 4), n};
        if (FWD) forward_line_lp<W, LEV, EXT>(io);
        else inverse_line_lp<W, LEV, EXT>(io);
    }
}

// Launch the line-parallel kernel for `levels` (5..8); false if none exists.
template <class W, int EXT, bool FWD, bool CONTIG>
bool lp_launch_one(int levels, float* dst, const float* src, const LpGeom& g, hipStream_t s) {
    const dim3 block(256);
    const int64_t lines_per_block = 4 * kLinesPerWave;
    const dim3 grid((unsigned)((g.total_lines + lines_per_block - 1) / lines_per_block));
    switch (levels) {
    case 5: hipLaunchKernelGGL((lp_kernel<W, 5, EXT, FWD, CONTIG>), grid, block, 0, s, dst, src, g); return true;
    case 6: hipLaunchKernelGGL((lp_kernel<W, 6, EXT, FWD, CONTIG>), grid, block, 0, s, dst, src, g); return true;
    case 7: hipLaunchKernelGGL((lp_kernel<W, 7, EXT, FWD, CONTIG>), grid, block, 0, s, dst, src, g); return true;
    case 8: hipLaunchKernelGGL((lp_kernel<W, 8, EXT, FWD, CONTIG>), grid, block, 0, s, dst, src, g); return true;
    default: return false;
    }
}

}  // namespace lp

template <class W, int EXT>
bool lp_launch(int levels, bool fwd, bool contig, float* dst, const float* src,
               const lp::LpGeom& g, hipStream_t s) {
    if (contig)
        return fwd ? lp::lp_launch_one<W, EXT, true, true>(levels, dst, src, g, s)
                   : lp::lp_launch_one<W, EXT, false, true>(levels, dst, src, g, s);
    return fwd ? lp::lp_launch_one<W, EXT, true, false>(levels, dst, src, g, s)
               : lp::lp_launch_one<W, EXT, false, false>(levels, dst, src, g, s);
}

}  // namespace whhd
