// Host build of the line-parallel transform algorithm (line_parallel.h with
// WHHD_HOST_EMULATION): one "lane" at a time, through an I/O object that
// reads and writes a line directly.  Exercises the level/block structure,
// reflections, streamed look-ahead loads and run stores of the GPU kernels on
// the CPU; the device-only parts (buffer addressing, LDS transposition) are
// not covered here.
//
//   lp_emulate(dst, src, lines, n, stride, levels, wavelet, forward, look)
//     wavelet 0: CDF3.7 'n' (n = 2^levels), 1: CDF4.6 'd' (n = 2^levels + 1);
//     line j is src[j*stride + i], i < n; look = look-ahead (0, 16 or 32),
//     or -1 for the two-phase loads of the contiguous-line kernels (returns -2
//     if their ranges are out of order, leave gaps or miss part of the line).
#define WHHD_HOST_EMULATION 1
#include "../../waveletpotentials_amd/csrc/line_parallel.h"

#include <cstdint>

namespace {

template <int LOOK>
struct HostIO {
    static constexpr int kLook = LOOK;
    static constexpr bool kTwoPhase = false;
    struct Pend {};
    const float* src;
    float* dst;
    int n;
    float load(int i) const { return src[i]; }
    template <int START, int CNT>
    void store_run(const float* vals) const {
        for (int v = 0; v < CNT; ++v) dst[START + v] = vals[v];
    }
};

// Two-phase loads as the contiguous-line kernels do them: fetch<A, B> holds
// samples [A, B) until unpack<A, B>; ranges must come in order, without gaps,
// start on multiples of 4 and end on multiples of 4 or at the line end.
struct HostIO2 {
    static constexpr int kLook = 0;
    static constexpr bool kTwoPhase = true;
    struct Pend {
        float v[65];
        int a = -1, b = -1;
    };
    const float* src;
    float* dst;
    int n;
    int* fetched_to;   // end of the last fetched range
    int* unpacked_to;  // end of the last unpacked range
    int* bad;
    template <int A, int B>
    void fetch(Pend& p) const {
        if constexpr (A < B) {
            static_assert((A & 3) == 0 && B - A <= 64, "range");
            if (A != *fetched_to || ((B & 3) && B != n)) ++*bad;
            for (int i = A; i < B; ++i) p.v[i - A] = src[i];
            p.a = A;
            p.b = B;
            *fetched_to = B;
        }
    }
    int* tail_step;    // next step of the aligned final stores
    template <int A, int B, int AREA = 0>
    void unpack(const Pend& p, float* x) const {
        if constexpr (A < B) {
            if (p.a != A || p.b != B || A != *unpacked_to) ++*bad;
            for (int i = A; i < B; ++i) x[i] = p.v[i - A];
            *unpacked_to = B;
        }
    }
    template <int START, int CNT>
    void store_run(const float* vals) const {
        for (int v = 0; v < CNT; ++v) dst[START + v] = vals[v];
    }
    // the aligned final stores take the run [32B, 32B + 32) at step B and the
    // sample N - 1 at the last step, in order (the wavefront-level storing is
    // checked by lp_tail_model)
    template <int N, int B>
    void store_tail(const float* vals) const {
        constexpr int K = whhd::lp::tail::steps<N>();
        if (B != *tail_step) ++*bad;
        *tail_step = B + 1;
        if constexpr (B == K) dst[N - 1] = vals[0];
        else
            for (int v = 0; v < 32; ++v) dst[32 * B + v] = vals[v];
    }
};

template <class W, int LEV, int EXT, int LOOK>
int run(float* dst, const float* src, int64_t lines, int64_t stride, bool fwd) {
    constexpr int n = (1 << LEV) + EXT;
    for (int64_t j = 0; j < lines; ++j) {
        if constexpr (LOOK >= 0) {
            HostIO<LOOK> io{src + j * stride, dst + j * stride, n};
            if (fwd) whhd::lp::forward_line_lp<W, LEV, EXT>(io);
            else whhd::lp::inverse_line_lp<W, LEV, EXT>(io);
        } else {
            int f = 0, u = 0, bad = 0, t = 0;
            HostIO2 io{src + j * stride, dst + j * stride, n, &f, &u, &bad, &t};
            if (fwd) whhd::lp::forward_line_lp<W, LEV, EXT>(io);
            else whhd::lp::inverse_line_lp<W, LEV, EXT>(io);
            if (bad || f != n || u != n) return -2;  // out of order, gaps, or not the whole line
            const int steps = !fwd && whhd::lp::tail::applies<n>() ? whhd::lp::tail::steps<n>() + 1 : 0;
            if (t != steps) return -2;                // aligned final stores incomplete
        }
    }
    return 0;
}

template <class W, int EXT, int LOOK>
int by_level(float* dst, const float* src, int64_t lines, int64_t stride, int levels, bool fwd) {
    switch (levels) {
    case 5: return run<W, 5, EXT, LOOK>(dst, src, lines, stride, fwd);
    case 6: return run<W, 6, EXT, LOOK>(dst, src, lines, stride, fwd);
    case 7: return run<W, 7, EXT, LOOK>(dst, src, lines, stride, fwd);
    case 8: return run<W, 8, EXT, LOOK>(dst, src, lines, stride, fwd);
    default: return -1;
    }
}

template <int LOOK>
int by_wavelet(float* dst, const float* src, int64_t lines, int64_t stride, int levels, int wt,
               bool fwd) {
    if (wt == 0) return by_level<whhd::W_Cdf37N, 0, LOOK>(dst, src, lines, stride, levels, fwd);
    if (wt == 1) return by_level<whhd::W_Cdf46D, 1, LOOK>(dst, src, lines, stride, levels, fwd);
    return -1;
}

}  // namespace

extern "C" int lp_emulate(float* dst, const float* src, long long lines, int n, long long stride,
                          int levels, int wt, int fwd, int look) {
    if (!dst || !src || dst == src || lines < 0 || stride < n) return -1;
    if (n != (1 << levels) + (wt == 1 ? 1 : 0)) return -1;
    switch (look) {
    case 0: return by_wavelet<0>(dst, src, lines, stride, levels, wt, fwd != 0);
    case 16: return by_wavelet<16>(dst, src, lines, stride, levels, wt, fwd != 0);
    case 32: return by_wavelet<32>(dst, src, lines, stride, levels, wt, fwd != 0);
    case -1: return by_wavelet<-1>(dst, src, lines, stride, levels, wt, fwd != 0);
    default: return -1;
    }
}

// Wavefront-level model of the inverse's aligned final stores (lp::tail), on
// the same index functions as the kernel: 64 lines of n samples, the first
// `lines_here` of which exist; sample i of line l is l * 1000 + i.  Each step
// first writes every lane's run to the three LDS areas, then every lane stores
// its rounds of 16-byte pieces (and, at the last step of a partial wavefront,
// the last line's lane its final samples one by one).  out (64 * n floats)
// receives the stores.  Returns 0, or the number of violations: an LDS word
// outside the areas, a store outside the existing lines, a sample stored
// twice or never, or a wrong value.
extern "C" int lp_tail_model(int n, int lines_here, float* out) {
    namespace T = whhd::lp::tail;
    if (n % 32 != 1 || n < 33 || lines_here < 1 || lines_here > 64) return -1;
    const int K = (n - 1) / 32;
    static float lds[T::kAreas * T::kArea];
    int* count = new int[64 * n]();
    int bad = 0;
    auto put = [&](int l, int i, float v) {
        if (l < 0 || l >= lines_here || i < 0 || l * n + i >= lines_here * n) { ++bad; return; }
        ++count[l * n + i];
        out[l * n + i] = v;
    };
    auto get = [&](int w) -> float {
        if (w < 0 || w >= T::kAreas * T::kArea) { ++bad; return 0.f; }
        return lds[w];
    };
    for (int B = 0; B <= K; ++B) {
        const bool fin = B == K;
        float* cur = lds + T::cur_area(B) * T::kArea;
        for (int lane = 0; lane < 64; ++lane) {
            if (fin) cur[lane] = (float)(lane * 1000 + n - 1);
            else
                for (int v = 0; v < 32; ++v) cur[v * T::kRow + lane] = (float)(lane * 1000 + 32 * B + v);
        }
        if (B == 0) continue;
        const int lone = T::lone(fin, lines_here);
        for (int lane = 0; lane < 64; ++lane) {
            const int q = lane & 7;
            for (int r = 0; r < T::rounds(fin); ++r) {
                const int l = T::line(lane, r), j0 = T::j0(l, r);
                float w[4];
                for (int e = 0; e < 4; ++e) w[e] = get(T::word(fin, B, j0 + 4 * q + e, l));
                if (!T::stores(fin, r, lane, l, lines_here, lone)) continue;
                // a piece runs on in memory from line l into line l + 1
                for (int e = 0; e < 4; ++e) {
                    const int i = T::first(B, j0, q) + e;
                    put(l + i / n, i % n, w[e]);
                }
            }
            if (fin && lane == lone)
                for (int j = T::lead(lane); j <= 32; ++j) put(lane, T::first(B, j, 0), get(T::word(true, B, j, lane)));
        }
    }
    for (int l = 0; l < lines_here; ++l)
        for (int i = 0; i < n; ++i)
            if (count[l * n + i] != 1 || out[l * n + i] != (float)(l * 1000 + i)) ++bad;
    delete[] count;
    return bad;
}
