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

#ifndef WHHD_HOST_EMULATION
#include <hip/hip_runtime.h>
#define WHHD_SCHED_BARRIER() __builtin_amdgcn_sched_barrier(0)
#else
#define WHHD_SCHED_BARRIER() ((void)0)
#endif

#include <cstdint>
#include <type_traits>
#include <utility>

#include "lifting.h"

#ifndef WHHD_LOOK_S
#define WHHD_LOOK_S 32
#endif
#ifndef WHHD_LOOK_C
#define WHHD_LOOK_C 32
#endif

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

#ifndef WHHD_KBLOCK
#define WHHD_KBLOCK 16
#endif
constexpr int kBlock = WHHD_KBLOCK;  // pairs written per block

// ---- aligned final stores of the inverse (contiguous lines) ----------------
// Lines of N = 2^LEV + 1 samples (N = 1 mod 32), 64 lines per wavefront
// from a multiple of 64 lines: line l starts l mod 32 floats past a
// 128-byte boundary, so its 128-byte lines begin at the samples i = lead(l)
// mod 32.  The inverse's last level produces the run [32B, 32B + 32) at
// step B < K = (N - 1) / 32 and the single sample N - 1 at step K.  Every
// run goes to LDS; step B >= 1 stores, for every line, the whole 128-byte
// line [32(B-1) + lead, 32B + lead) from the previous run (rows lead..31)
// and this one (rows 0..lead-1).  Step 0's run stays in area 2 to the end:
// its first lead samples share a 128-byte line with the previous line's
// last samples.  Step K stores [32(K-1) + lead, N) of line l together with
// [0, lead - 1) of line l + 1; a line with lead 0 stores its run [32(K-1),
// 32K) and, in a ninth round, sample N - 1 with [0, 31) of line l + 1.  So
// no 128-byte line is written in two parts at different times.  Lines 63
// and 0 of two wavefronts never share a 128-byte line (64 N = 0 mod 32); in
// a partial last wavefront the last line's own lane stores its final
// samples one by one, since its last 128-byte line would reach into a line
// that does not exist.  The runs alternate between areas 0 and 1, and the
// unpack of step B uses area B mod 2, which holds no run still needed.
// Each step stores 16-byte pieces: lane -> piece lane mod 8 of line
// line(lane, r) in round r.
namespace tail {
constexpr int kRow = 64 + 1;           // LDS row: one value of each of the 64 lines
constexpr int kArea = 32 * kRow;       // one run of 32 samples
constexpr int kAreas = 3;
template <int N>
constexpr bool applies() { return N % 32 == 1 && N >= 33; }
template <int N>
constexpr int steps() { return (N - 1) / 32; }
constexpr int lead(int l) { return (32 - (l & 31)) & 31; }
constexpr int cur_area(int B) { return B == 0 ? 2 : B % 2; }
constexpr int prv_area(int B) { return B == 1 ? 2 : (B + 1) % 2; }
constexpr int rounds(bool fin) { return fin ? 9 : 8; }
// round 8 (final step): lines 0 and 32 by lanes 0-15; the other lanes repeat them and do not store
constexpr int line(int lane, int r) { return r < 8 ? (lane >> 3) + 8 * r : ((lane >> 3) & 1) * 32; }
constexpr int j0(int l, int r) { return r < 8 ? lead(l) : 32; }
// sample of line l that element j of its 128-byte line at step B is (j >= 33
// at the final step: sample j - 33 of line l + 1)
constexpr int first(int B, int j0, int q) { return 32 * (B - 1) + j0 + 4 * q; }
// LDS word (from the wavefront's staging base) holding element j of line l's
// 128-byte line at step B
// (written as j * kRow + l plus a constant of the part j falls in)
constexpr int word(bool fin, int B, int j, int l) {
    return j * kRow + l +
           (j < 32 ? prv_area(B) * kArea
                   : (!fin || j == 32) ? cur_area(B) * kArea - kArea : 2 * kArea - 33 * kRow + 1);
}
// line whose lane stores its final samples itself (64: none)
constexpr int lone(bool fin, int lines_here) { return fin && lines_here < 64 ? lines_here - 1 : 64; }
constexpr bool stores(bool fin, int r, int lane, int l, int lines_here, int lone) {
    return fin ? (r < 8 || lane < 16) && l < lines_here && l != lone : l < lines_here;
}
}  // namespace tail

#ifndef WHHD_HOST_EMULATION
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
    static constexpr int kLook = WHHD_LOOK_S;  // samples loaded ahead of use
    static constexpr bool kTwoPhase = false;   // loads land in the line registers directly
    struct Pend {};
    __amdgpu_buffer_rsrc_t rs;
    __amdgpu_buffer_rsrc_t rd;
    uint32_t voff_ld;  // byte offset of the lane's line start (loads)
    uint32_t voff_st;  // same, or out of range for lanes that do not store
    int step_bytes;
    int n;
    WHHD_DEV float load(int i) const {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, voff_ld, i * step_bytes, 0));
    }
    // The stores' scalar offsets come from step_st, an opaque copy of
    // step_bytes in the forward: otherwise the compiler keeps every load's
    // offset i * step for the store of sample i much later, and the forward
    // spills ~70 of them to VGPR lanes (a write and a read VALU instruction
    // each) instead of multiplying again.
    int step_st;
    WHHD_DEV void store(int i, float v) const {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rd, voff_st,
                                              i * step_st, 0);
    }
    // samples [START, START+CNT) of the line; each store is already one row
    template <int START, int CNT>
    WHHD_DEV void store_run(const float* vals) const {
#pragma unroll
        for (int v = 0; v < CNT; ++v) store(START + v, vals[v]);
    }
};

// Contiguous lines (y axis): the lane's line is n consecutive floats.  Loads
// fetch 16-byte runs of four samples (dword alignment suffices on gfx950:
// odd-length lines start at any dword) and the four element reads of a run
// fold into one load; the last run of a line is read sample by sample so
// nothing past the line is requested.  Stores go through a buffer resource.
using U4 = __attribute__((ext_vector_type(4))) uint32_t;
struct __attribute__((packed, aligned(4))) Quad {
    float v[4];
};
constexpr int kStageRun = 32;       // longest run written at once
constexpr int kStageRow = 64 + 1;   // LDS row: one value of each of the 64 lines
static_assert(kStageRow == tail::kRow && kStageRun * kStageRow == tail::kArea, "staging layout");
static_assert(64 * 32 <= kStageRun * kStageRow && (kStageRun * kStageRow) % 4 == 0,
              "a staging area holds 64 rows of 32 floats, 16-byte aligned");

struct ContigIO {
    static constexpr int kLook = 0;            // (two-phase: one block ahead)
    static constexpr bool kTwoPhase = true;
    const float* line;    // the lane's own line (tail loads)
    __amdgpu_buffer_rsrc_t rs;
    __amdgpu_buffer_rsrc_t rd;
    uint32_t voff_st;     // byte offset of the lane's line, or out of range for lanes that do not store
    uint32_t first_off;   // byte offset of the line of lane 0
    int lines_here;       // lines of this wavefront that exist (<= 64)
    int n;
    float* stage;         // this wavefront's LDS staging area, kStageRun rows
    int lane;
    // Two-phase loads.  fetch<A, B> issues the loads of samples [A, B) of the
    // wavefront's 64 lines as coalesced 16-byte pieces: in each instruction
    // consecutive lanes read consecutive pieces of one line (8 lanes cover 128
    // bytes of a line), instead of every lane reading 16 bytes of its own line,
    // which touches 64 lines per instruction and streams poorly from HBM.
    // unpack<A, B> later passes the pieces through LDS so that each lane ends
    // up with its own line's samples; between the two the wavefront computes
    // the previous block, which hides the load latency.  A is a multiple of 4;
    // B is too, or the line length (a one-sample tail of an odd line, read by
    // each lane for its own line).  All four elements of every 16-byte raw
    // buffer load are used (LLVM narrows such a load whose leading elements
    // are unused without moving its offset, so single samples are never
    // extracted from one).
    static constexpr int kMaxQ = 16;  // 16-byte pieces per lane in flight
    struct Pend {
        Quad q[kMaxQ];
        float tail;
    };
    template <int NQ, int K>
    static constexpr int chunk() { return NQ - K >= 8 ? 8 : NQ - K >= 4 ? 4 : NQ - K >= 2 ? 2 : 1; }
    template <int Q0, int NQ, int K>
    WHHD_DEV void fetch_chunks(Pend& p) const {
        if constexpr (K < NQ) {
            constexpr int c = chunk<NQ, K>();
            // one per-lane offset per chunk shape; rounds differ by constants
            // (scalar offsets).  Lines past the end read past the array and
            // get zeros from the buffer range check.
            const uint32_t voff = first_off + (uint32_t)(((lane / c) * n + 4 * (lane % c)) * 4);
#pragma unroll
            for (int r = 0; r < c; ++r) {
                const U4 v = __builtin_amdgcn_raw_buffer_load_b128(
                    rs, voff, (r * (64 / c) * n + 4 * (Q0 + K)) * 4, 0);
                p.q[K + r] = __builtin_bit_cast(Quad, v);
            }
            fetch_chunks<Q0, NQ, K + c>(p);
        }
    }
    template <int A, int B>
    WHHD_DEV void fetch(Pend& p) const {
        if constexpr (A < B) {
            static_assert((A & 3) == 0, "range start must be a multiple of 4");
            constexpr int QA = A / 4, QB = (B & ~3) / 4, NQ = QB - QA;
            static_assert(NQ <= kMaxQ, "range too long");
            fetch_chunks<QA, NQ, 0>(p);
            if constexpr ((B & 3) != 0) {
                static_assert((B & 3) == 1, "only a one-sample tail");
                p.tail = line[B - 1];
            }
            // keep the loads ahead of the block computed next
            WHHD_SCHED_BARRIER();
        }
    }
    // Pieces pass through LDS as rows of 32 floats, one row per line, with
    // the 16-byte piece j of line l at quad j ^ (l mod 8): a piece is one
    // 16-byte LDS access, and the 8 lanes of a group (8 lines read by their
    // own lanes, or 8 pieces of one line written) reach 8 different quads,
    // so no bank is hit twice.  A staging area (32 x 65 floats) holds the 64 rows.
    static WHHD_DEV int quad(int l, int j) { return 32 * l + 4 * (j ^ (l & 7)); }
    template <int A, int NQ, int K>
    WHHD_DEV void unpack_chunks(const Pend& p, float* x, float* st) const {
        if constexpr (K < NQ) {
            constexpr int c = chunk<NQ, K>();
            // piece lane mod c of line lane / c + r * (64 / c); lines 64 / c
            // apart share l mod 8, so one address serves every r
            const int w0 = quad(lane / c, lane % c);
#pragma unroll
            for (int r = 0; r < c; ++r)
                *(U4*)(st + w0 + 32 * r * (64 / c)) = __builtin_bit_cast(U4, p.q[K + r]);
            __builtin_amdgcn_wave_barrier();
            asm volatile("" ::: "memory");
#pragma unroll
            for (int j = 0; j < c; ++j) {
                const Quad v = __builtin_bit_cast(Quad, *(const U4*)(st + quad(lane, j)));
#pragma unroll
                for (int e = 0; e < 4; ++e) x[A + 4 * K + 4 * j + e] = v.v[e];
            }
            __builtin_amdgcn_wave_barrier();
            asm volatile("" ::: "memory");
            unpack_chunks<A, NQ, K + c>(p, x, st);
        }
    }
    // AREA: which staging area passes the pieces (0 unless the aligned final
    // stores below keep a run in area 0 or 1)
    template <int A, int B, int AREA = 0>
    WHHD_DEV void unpack(const Pend& p, float* x) const {
        if constexpr (A < B) {
            constexpr int NQ = (B & ~3) / 4 - A / 4;
            unpack_chunks<A, NQ, 0>(p, x, area<AREA>());
            if constexpr ((B & 3) != 0) x[B - 1] = p.tail;
        }
    }
    template <int K>
    WHHD_DEV float* area() const { return stage + K * tail::kArea; }

    // Aligned final stores of the inverse (lp::tail): the run of step B goes
    // to LDS and whole 128-byte lines are stored from there.
    template <int N, int B>
    WHHD_DEV void store_tail(const float* vals) const {
        constexpr bool FIN = B == tail::steps<N>();
        float* cur = stage + tail::cur_area(B) * tail::kArea;
        if constexpr (FIN) {
            cur[lane] = vals[0];
        } else {
#pragma unroll
            for (int v = 0; v < 32; ++v) cur[v * kStageRow + lane] = vals[v];
        }
        __builtin_amdgcn_wave_barrier();
        asm volatile("" ::: "memory");
        if constexpr (B > 0) {
            const int q = lane & 7;
            const int lone = tail::lone(FIN, lines_here);
#pragma unroll
            for (int r = 0; r < tail::rounds(FIN); ++r) {
                const int l = tail::line(lane, r), j0 = tail::j0(l, r);
                U4 w;
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    w[e] = __builtin_bit_cast(uint32_t, stage[tail::word(FIN, B, j0 + 4 * q + e, l)]);
                const uint32_t off = tail::stores(FIN, r, lane, l, lines_here, lone)
                                         ? first_off + (uint32_t)((l * N + tail::first(B, j0, q)) * 4)
                                         : kOutOfRange;
                __builtin_amdgcn_raw_buffer_store_b128(w, rd, off, 0, 0);
            }
            if constexpr (FIN) {
                if (lane == lone) {
                    for (int j = tail::lead(lane); j <= 32; ++j)
                        __builtin_amdgcn_raw_buffer_store_b32(
                            __builtin_bit_cast(uint32_t, stage[tail::word(true, B, j, lane)]), rd, voff_st,
                            tail::first(B, j, 0) * 4, 0);
                }
            }
            __builtin_amdgcn_wave_barrier();
            asm volatile("" ::: "memory");
        }
    }
    // Samples [START, START+CNT) of every lane's line.  Whole 16-byte pieces
    // are transposed through LDS so that consecutive lanes store consecutive
    // pieces of one line (64-128 byte segments per line) instead of 64 lines
    // touched 16 bytes at a time; a remainder of CNT % 4 samples is stored by
    // each lane for its own line.  Lines past the end get an out-of-range
    // offset, which the hardware discards, so there is no branch.
    template <int START, int CNT>
    WHHD_DEV void store_run(const float* vals) const {
        constexpr int CQ = CNT & ~3;
        if constexpr (CQ > 0) {
            static_assert(CQ <= kStageRun, "run too long");
            constexpr int NQ = CQ / 4;                 // 16-byte pieces per line
            constexpr int LPI = 64 / NQ;               // lines per pass
            static_assert(64 % NQ == 0, "pieces per line must divide 64");
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                Quad v;
#pragma unroll
                for (int e = 0; e < 4; ++e) v.v[e] = vals[4 * j + e];
                *(U4*)(stage + quad(lane, j)) = __builtin_bit_cast(U4, v);
            }
            __builtin_amdgcn_wave_barrier();
            asm volatile("" ::: "memory");
            const int q = lane % NQ;
            // lines LPI apart share l mod 8: one LDS address for every r
            const int w0 = quad(lane / NQ, q);
#pragma unroll
            for (int r = 0; r < NQ; ++r) {
                const int l = lane / NQ + r * LPI;
                const U4 w = *(const U4*)(stage + w0 + 32 * r * LPI);
                const uint32_t off = l < lines_here ? first_off + (uint32_t)((l * n + START + 4 * q) * 4)
                                                    : kOutOfRange;
                __builtin_amdgcn_raw_buffer_store_b128(w, rd, off, 0, 0);
            }
            __builtin_amdgcn_wave_barrier();
            asm volatile("" ::: "memory");
        }
#pragma unroll
        for (int v = CQ; v < CNT; ++v)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, vals[v]), rd, voff_st,
                                                  (START + v) * 4, 0);
    }
};

#endif  // WHHD_HOST_EMULATION

// ---- streamed loads ------------------------------------------------------
// Both directions consume a line's samples in increasing index order (the
// forward's top level block by block; the inverse's coarse block, then the
// details of level 2, 3, ... block by block).  The loads are issued ahead of
// their use: before a block is computed, the samples up to its need plus a
// look-ahead (IO::kLook samples) are requested, and a scheduling barrier keeps the compiler
// from sinking them back to their uses.  Without it the loads of a block are
// issued right before the block and each block pays the full memory latency.
constexpr int cmin(int a, int b) { return a < b ? a : b; }
constexpr int round4(int v) { return (v + 3) & ~3; }

// Issue loads of samples [FROM, TO) into x (FROM >= TO: nothing).
template <int FROM, int TO, class IO>
WHHD_DEV void issue(const IO& io, float* x) {
    if constexpr (FROM < TO) {
        static_for<FROM, TO>([&](auto i) { x[decltype(i)::value] = io.load(decltype(i)::value); });
        WHHD_SCHED_BARRIER();
    }
}

// ---- forward --------------------------------------------------------------
// Samples of the line a forward top-level block b (pairs [b*kBlock - H,
// (b+1)*kBlock + H)) can read; reflections at the line end stay below N.
template <int H, int N>
constexpr int fwd_need(int b) { return cmin(N, 2 * (b * kBlock + kBlock + H) + 2); }
template <int H, int N, int LA>
constexpr int fwd_front(int b) { return b < 0 ? 0 : cmin(N, round4(fwd_need<H, N>(b) + LA)); }

// One forward level: the 2^(l+1)+EXT samples come from `in`, approximations
// go to `out`, details to the line at [W, 2W-EXT).  For the top level (TOP)
// `in` is filled here from the line by streamed loads.
template <class W, int l, int EXT, bool TOP, class IO>
WHHD_DEV void fwd_level(const IO& io, float* in, float* out) {
    constexpr int H = W::halo;
    constexpr int L = 1 << (l + 1);
    constexpr int N = L + EXT;
    constexpr int Wd = L / 2 + EXT;
    constexpr int NP = kBlock + 2 * H;
    constexpr int NB = (Wd + kBlock - 1) / kBlock;
    [[maybe_unused]] typename IO::Pend pend;
    // contiguous lines store the details of two blocks at once (128 bytes from
    // a 128-byte boundary when the line is aligned), not 64-byte halves written
    // a block apart
    [[maybe_unused]] float held[kBlock];
    if constexpr (TOP && IO::kTwoPhase) io.template fetch<0, fwd_front<H, N, 0>(0)>(pend);
    static_for<0, NB>([&](auto bi) {
        constexpr int b = decltype(bi)::value;
        constexpr int b0 = b * kBlock;
        if constexpr (TOP && IO::kTwoPhase) {
            // block b's samples were fetched one block earlier; fetch block b+1's now
            io.template unpack<fwd_front<H, N, 0>(b - 1), fwd_front<H, N, 0>(b)>(pend, in);
            io.template fetch<fwd_front<H, N, 0>(b), fwd_front<H, N, 0>(b + 1)>(pend);
        } else if constexpr (TOP) {
            issue<fwd_front<H, N, IO::kLook>(b - 1), fwd_front<H, N, IO::kLook>(b)>(io, in);
        }
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
                A[j] = in[ai];
                D[j] = in[di];
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
        // details of the next block (a block is held only when that one stores)
        constexpr int nd_next = (Wd - EXT - b0 - kBlock) < kBlock ? (Wd - EXT - b0 - kBlock) : kBlock;
        if constexpr (IO::kTwoPhase && nd > 0) {
            if constexpr (b % 2 == 0 && nd == kBlock && b + 1 < NB && nd_next > 0) {
#pragma unroll
                for (int v = 0; v < kBlock; ++v) held[v] = run[v];
            } else if constexpr (b % 2 == 1) {
                float pair[2 * kBlock];
#pragma unroll
                for (int v = 0; v < kBlock; ++v) pair[v] = held[v];
#pragma unroll
                for (int v = 0; v < nd; ++v) pair[kBlock + v] = run[v];
                io.template store_run<Wd + b0 - kBlock, kBlock + nd>(pair);
            } else {
                io.template store_run<Wd + b0, nd>(run);
            }
        } else if constexpr (nd > 0) {
            io.template store_run<Wd + b0, nd>(run);
        }
    });
}

template <class W, int l, int EXT, class IO>
WHHD_DEV void fwd_levels_below(const IO& io, float* in) {
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
    float x[(1 << LEV) + EXT];
    if constexpr (LEV > 2) {
        float out[(1 << (LEV - 1)) + EXT];
        fwd_level<W, LEV - 1, EXT, true>(io, x, out);
        fwd_levels_below<W, LEV - 2, EXT>(io, out);
    } else {
        if constexpr (IO::kTwoPhase) {
            typename IO::Pend pend;
            io.template fetch<0, (1 << LEV) + EXT>(pend);
            io.template unpack<0, (1 << LEV) + EXT>(pend, x);
        } else {
            issue<0, (1 << LEV) + EXT>(io, x);
        }
        fwd_levels_below<W, 1, EXT>(io, x);
    }
}

// ---- inverse --------------------------------------------------------------
// Highest line index (exclusive) block b of inverse level l reads: its
// details [L+EXT, 2L+EXT) around pairs [b*kBlock - H, (b+1)*kBlock + H).
template <int H, int EXT, int N>
constexpr int inv_need(int l, int b) {
    return cmin(N, (1 << l) + EXT + cmin(1 << l, b * kBlock + kBlock + H + 1));
}
template <int H, int EXT, int N, int M, int LA>
constexpr int inv_front(int l, int b) {
    return cmin(N, round4((l < 2 ? M : inv_need<H, EXT, N>(l, b)) + LA));
}
template <int H, int EXT, int N, int M, int LA>
constexpr int inv_front_before(int l, int b) {
    // front of the block consumed before (l, b)
    return b > 0 ? inv_front<H, EXT, N, M, LA>(l, b - 1)
                 : (l == 2 ? inv_front<H, EXT, N, M, LA>(1, 0)
                           : inv_front<H, EXT, N, M, LA>(l - 1, (((1 << (l - 1)) + EXT + kBlock - 1) / kBlock) - 1));
}

// Two-phase fronts: the samples step (l, b) needs, rounded to whole 16-byte
// pieces (l = 1 is the coarse block); the step after (l, b) is (l, b+1), then
// (l+1, 0), and after the last block of level LEV-1 nothing.
template <int H, int EXT, int N, int M>
constexpr int inv_need2(int l, int b) {
    return cmin(N, round4(l < 2 ? M : inv_need<H, EXT, N>(l, b)));
}
template <int H, int EXT, int N, int M, int LEV>
constexpr int inv_need2_next(int l, int b) {
    return l < 2 ? (LEV > 2 ? inv_need2<H, EXT, N, M>(2, 0) : N)
                 : (b + 1 < (((1 << l) + EXT + kBlock - 1) / kBlock)
                        ? inv_need2<H, EXT, N, M>(l, b + 1)
                        : (l + 1 < LEV ? inv_need2<H, EXT, N, M>(l + 1, 0) : N));
}

// One inverse level l: approximations [0, 2^l+EXT) from `in`, details from
// the streamed line x at [2^l+EXT, 2^(l+1)+EXT); reconstructed samples to
// `out`, or to the line when LAST.
template <class W, int l, int LEV, int EXT, bool LAST, class IO>
WHHD_DEV void inv_level(const IO& io, typename IO::Pend& pend, float* x, const float* in,
                        float* out) {
    constexpr int N = (1 << LEV) + EXT;
    constexpr int H = W::halo;
    constexpr int M = W::m10;
    constexpr int L = 1 << l;
    constexpr int Wd = L + EXT;
    constexpr int NP = kBlock + 2 * H;
    constexpr int NB = (Wd + kBlock - 1) / kBlock;
    static_for<0, NB>([&](auto bi) {
        constexpr int b = decltype(bi)::value;
        constexpr int b0 = b * kBlock;
        if constexpr (IO::kTwoPhase) {
            constexpr int prev = b > 0 ? inv_need2<H, EXT, N, M>(l, b - 1)
                                       : (l == 2 ? inv_need2<H, EXT, N, M>(1, 0)
                                                 : inv_need2<H, EXT, N, M>(l - 1, (((1 << (l - 1)) + EXT + kBlock - 1) / kBlock) - 1));
            // the last level's aligned stores keep runs in areas 0 and 1
            io.template unpack<prev, inv_need2<H, EXT, N, M>(l, b), (LAST && tail::applies<N>()) ? b % 2 : 0>(pend, x);
            io.template fetch<inv_need2<H, EXT, N, M>(l, b), inv_need2_next<H, EXT, N, M, LEV>(l, b)>(pend);
        } else {
            issue<inv_front_before<H, EXT, N, M, IO::kLook>(l, b), inv_front<H, EXT, N, M, IO::kLook>(l, b)>(io, x);
        }
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
                D[j] = x[L + EXT + di];
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
        if constexpr (LAST && ns > 0) {
            if constexpr (IO::kTwoPhase && tail::applies<N>()) io.template store_tail<N, b>(run);
            else io.template store_run<2 * b0, ns>(run);
        }
    });
}

template <class W, int l, int LEV, int EXT, class IO>
WHHD_DEV void inv_levels_from(const IO& io, typename IO::Pend& pend, float* x, const float* in) {
    // `in` holds the 2^l+EXT approximations of level l
    if constexpr (l == LEV - 1) {
        inv_level<W, l, LEV, EXT, true>(io, pend, x, in, nullptr);
    } else {
        float out[(1 << (l + 1)) + EXT];
        inv_level<W, l, LEV, EXT, false>(io, pend, x, in, out);
        inv_levels_from<W, l + 1, LEV, EXT>(io, pend, x, out);
    }
}

template <class W, int LEV, int EXT, class IO>
WHHD_DEV void inverse_line_lp(const IO& io) {
    constexpr int M = W::m10;
    constexpr int N = (1 << LEV) + EXT;
    constexpr int H = W::halo;
    float x[N];
    typename IO::Pend pend;
    if constexpr (IO::kTwoPhase) {
        io.template fetch<0, inv_need2<H, EXT, N, M>(1, 0)>(pend);
        io.template unpack<0, inv_need2<H, EXT, N, M>(1, 0)>(pend, x);
        io.template fetch<inv_need2<H, EXT, N, M>(1, 0), inv_need2_next<H, EXT, N, M, LEV>(1, 0)>(pend);
    } else {
        issue<0, inv_front<H, EXT, N, M, IO::kLook>(1, 0)>(io, x);
    }
    float a[M];
    static_for<0, M>([&](auto ri) {
        constexpr int r = decltype(ri)::value;
        float acc = 0.f;
        static_for<0, M>([&](auto ci) {
            constexpr int cc = decltype(ci)::value;
            acc += x[cc] * float(W::il10[r * M + cc]);
        });
        a[r] = acc;
    });
    if constexpr (LEV > 2) {
        inv_levels_from<W, 2, LEV, EXT>(io, pend, x, a);
    } else {
        io.template store_run<0, M>(a);
    }
}

#ifndef WHHD_HOST_EMULATION
// ---- kernel -----------------------------------------------------------------
// Lines are numbered li = outer * lines_inner + inner; line li starts at
// inner * inner_stride + outer * outer_stride and its samples are `step` apart.
// Problems of a batch are arrays of equal extents `batch_stride` elements
// apart; grid row blockIdx.y transforms problem blockIdx.y.
struct LpGeom {
    int64_t total_lines;
    int64_t total_elems;  // elements in one problem's array (buffer range)
    int lines_inner;
    int64_t inner_stride;
    int64_t outer_stride;
    int64_t step;
    int64_t batch_stride;
};

template <class W, int LEV, int EXT, bool FWD, bool CONTIG>
// Two wavefronts per SIMD: a pass over 256^3 has about one wavefront of lines
// per SIMD, so the register budget goes to loads in flight, not occupancy.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
lp_kernel(float* __restrict__ dst, const float* __restrict__ src, LpGeom g) {
    src += (int64_t)blockIdx.y * g.batch_stride;
    dst += (int64_t)blockIdx.y * g.batch_stride;
    // Workgroups go to the 8 XCDs round-robin (blockIdx.x % 8), each XCD with
    // its own L2; give every XCD a contiguous run of tiles, so that the cache
    // lines two neighbouring tiles share (rows of 2^L + 1 floats start
    // unaligned) are fetched and written through one L2.
    const unsigned nwg = gridDim.x, bxr = blockIdx.x % 8u, bxq = blockIdx.x / 8u;
    const unsigned tile = bxr * (nwg / 8u) + (bxr < nwg % 8u ? bxr : nwg % 8u) + bxq;
    const int lane_ = threadIdx.x & 63;
    const int64_t wave_first = ((int64_t)tile * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64;
    const int64_t li = wave_first + lane_;
    const bool valid = li < g.total_lines;  // lanes past the last line load line 0, never store
    const int64_t l = valid ? li : 0;
    const int64_t outer = l / g.lines_inner;
    const int64_t inner = l - outer * g.lines_inner;
    constexpr int n = (1 << LEV) + EXT;
    const uint32_t bytes = (uint32_t)(g.total_elems * 4);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc((void*)src, 0, bytes, kBufWord3);
    const __amdgpu_buffer_rsrc_t rd = __builtin_amdgcn_make_buffer_rsrc((void*)dst, 0, bytes, kBufWord3);
    if constexpr (CONTIG) {
        // contiguous lines are numbered in memory order: line li starts at li * n
        // one staging area of kStageRun x kStageRow per wavefront (dynamic
        // LDS), three for the inverse's aligned stores
        extern __shared__ __attribute__((aligned(16))) float lp_stage[];
        constexpr int kWaveStage = (!FWD && tail::applies<n>() ? tail::kAreas : 1) * kStageRun * kStageRow;
        const int64_t here = g.total_lines - wave_first;
        const uint32_t voff = (uint32_t)(l * n * 4);
        ContigIO io{src + l * n, rs, rd, valid ? voff : kOutOfRange,
                    (uint32_t)(wave_first * n * 4), (int)(here < 64 ? here : 64), n,
                    lp_stage + (threadIdx.x >> 6) * kWaveStage, lane_};
        if (FWD) forward_line_lp<W, LEV, EXT>(io);
        else inverse_line_lp<W, LEV, EXT>(io);
    } else {
        const uint32_t voff = (uint32_t)((inner * g.inner_stride + outer * g.outer_stride) * 4);
        int step_st = (int)(g.step * 4);
        if constexpr (FWD) asm volatile("s_mov_b32 %0, %0" : "+s"(step_st));
        StridedIO io{rs, rd, voff, valid ? voff : kOutOfRange, (int)(g.step * 4), n, step_st};
        if (FWD) forward_line_lp<W, LEV, EXT>(io);
        else inverse_line_lp<W, LEV, EXT>(io);
    }
}

// Launch the line-parallel kernel for `levels` (5..8) over `nb` problems with
// `waves` (1, 2 or 4) wavefronts per workgroup; false if none exists.
template <class W, int EXT, bool FWD, bool CONTIG>
bool lp_launch_one(int levels, float* dst, const float* src, const LpGeom& g, int waves, int nb,
                   hipStream_t s) {
    if (waves != 1 && waves != 2 && waves != 4) return false;
    if (nb < 1 || nb > 65535) return false;
    const dim3 block(64 * waves);
    const int64_t lines_per_block = 64 * waves;
    const dim3 grid((unsigned)((g.total_lines + lines_per_block - 1) / lines_per_block), (unsigned)nb);
    // the inverse of lines of 2^levels + 1 samples stores through three runs
    const int n = (1 << levels) + EXT;
    const int rows = !FWD && n % 32 == 1 && n >= 33 ? tail::kAreas * kStageRun : kStageRun;
    const size_t lds = CONTIG ? (size_t)waves * rows * kStageRow * sizeof(float) : 0;
    switch (levels) {
    case 5: hipLaunchKernelGGL((lp_kernel<W, 5, EXT, FWD, CONTIG>), grid, block, lds, s, dst, src, g); return true;
    case 6: hipLaunchKernelGGL((lp_kernel<W, 6, EXT, FWD, CONTIG>), grid, block, lds, s, dst, src, g); return true;
    case 7: hipLaunchKernelGGL((lp_kernel<W, 7, EXT, FWD, CONTIG>), grid, block, lds, s, dst, src, g); return true;
    case 8: hipLaunchKernelGGL((lp_kernel<W, 8, EXT, FWD, CONTIG>), grid, block, lds, s, dst, src, g); return true;
    default: return false;
    }
}

#endif  // WHHD_HOST_EMULATION

}  // namespace lp

#ifndef WHHD_HOST_EMULATION
template <class W, int EXT>
bool lp_launch(int levels, bool fwd, bool contig, float* dst, const float* src,
               const lp::LpGeom& g, int waves, int nb, hipStream_t s) {
    if (contig)
        return fwd ? lp::lp_launch_one<W, EXT, true, true>(levels, dst, src, g, waves, nb, s)
                   : lp::lp_launch_one<W, EXT, false, true>(levels, dst, src, g, waves, nb, s);
    return fwd ? lp::lp_launch_one<W, EXT, true, false>(levels, dst, src, g, waves, nb, s)
               : lp::lp_launch_one<W, EXT, false, false>(levels, dst, src, g, waves, nb, s);
}

#endif  // WHHD_HOST_EMULATION

}  // namespace whhd
