// whhd_compat.hpp -- source-level drop-in for the reference's classes on the
// recovery path, implemented over the C ABI of libwhhd.so (include/whhd.h).
//
// A translation unit that used
//     #include "cuWaveletPotentialRecover3D.cuh"   (and friends)
// can include this header instead, add `using namespace whhd_compat;`, and
// link -lwhhd.  Class names, constructor and method signatures follow the
// reference headers:
//   CuWaveletPotentialRecover3D  WaveletPotentials/cuWaveletPotentialRecover3D.cuh:8-31
//   CuWaveletPotentialRecover2D  WaveletPotentials/cuWaveletPotentialRecover2D.cuh:8-31
//   CuWHHDForward::Solve1D       WaveletPotentials/cuWHHDForward.cuh:11-23
//   CuDivLocalProject3D          WaveletPotentials/cuDivLocalProject3D.cuh:7-21
//   CuGetInnerProduct::GetIp     WaveletPotentials/cuGetInnerProduct.cuh:7-20
//   CuSolidLevelSet3D            WaveletPotentials/cuSolidLevelSet3D.cuh:8-25
//   CuFastSweeping::Solve        WaveletPotentials/cuFastSweeping.cuh:6-13
//   CuMultigridFrac3D            WaveletPotentials/cuMultigridFrac3D.cuh:12-18
//   CuGradient::Gradient3D_Frac_P, CuConvergence::GetB_3d_frac_vel
//                                WaveletPotentials/cuGradient.cuh:42-43, cuConvergence.cuh:16-17
//   CudaMatAdd/MatMul/ArraySum/AbsSum/NormSum/FindMaxValue
//                                WaveletPotentials/cudaMath.cuh:127-139, 202
//   CuDst3D, CuParallelSweep3D   WaveletPotentials/cuDst3D.cuh:6-13, cuParallelSweep3D.cuh:9-23
//   CuDst                        WaveletPotentials/cuDst.cuh
// Differences, all deliberate:
//   - DTYPE is float (the reference builds FP64 or FP32 by macro); arrays are
//     device memory, y fastest, as in the reference.
//   - Failures throw whhd_compat::Error (the reference aborts the process).
//   - Work goes to the stream set with SetStream (default: the null stream);
//     the reference always uses the null stream.
//   - CuGetInnerProduct's tables are computed in FP64 and stored as DTYPE
//     (the reference accumulates them in float).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "whhd.h"

namespace whhd_compat {

using DTYPE = float;
using DTYPE2 = float2;
using DTYPE3 = float3;

// WaveletPotentials/WaveletTypes.h:7-28
enum WaveletType : int {
    WT_CDF3_9 = WHHD_WT_CDF3_9, WT_CDF3_11 = WHHD_WT_CDF3_11, WT_CDF3_5 = WHHD_WT_CDF3_5,
    WT_CDF2_10 = WHHD_WT_CDF2_10, WT_CDF2_12 = WHHD_WT_CDF2_12, WT_CDF2_6 = WHHD_WT_CDF2_6,
    WT_CDF4_4 = WHHD_WT_CDF4_4, WT_CDF4_6 = WHHD_WT_CDF4_6, WT_CDF4_8 = WHHD_WT_CDF4_8,
    WT_CDF4_10 = WHHD_WT_CDF4_10, WT_CDF2_8 = WHHD_WT_CDF2_8, WT_CDF3_7 = WHHD_WT_CDF3_7,
    WT_CDF26_44 = WHHD_WT_CDF26_44, WT_CDF28_46 = WHHD_WT_CDF28_46,
    WT_CDF44_26 = WHHD_WT_CDF44_26, WT_CDF46_28 = WHHD_WT_CDF46_28, WT_CDF2_0 = WHHD_WT_CDF2_0,
    WT_CDF1_1 = WHHD_WT_CDF1_1, WT_CDF1_7 = WHHD_WT_CDF1_7
};

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& where)
        : std::runtime_error(where + ": " + whhd_status_string(s) + ": " + whhd_last_error()),
          status(s) {}
};

inline void check(int status, const char* where) {
    if (status != WHHD_OK) throw Error(status, where);
}

// WaveletPotentials/WaveletTypes.h:102-143
inline char proj_bc(char bc) { return whhd_proj_bc(bc); }
inline WaveletType proj_WaveletType(WaveletType t) {
    return static_cast<WaveletType>(whhd_proj_wavelet(t));
}

// Wavelet names of the Gram-table keys (WaveletPotentials/WaveletTypes.h:67-100,
// WaveletTypes.cpp:4-67): "<bc>_cdf<p>.<q>"; an unknown name reads as CDF3.5,
// as in the reference.
inline const std::pair<WaveletType, const char*>* wavelet_names(int* count) {
    static const std::pair<WaveletType, const char*> t[] = {
        {WT_CDF3_5, "cdf3.5"},   {WT_CDF2_6, "cdf2.6"},   {WT_CDF4_4, "cdf4.4"},
        {WT_CDF4_6, "cdf4.6"},   {WT_CDF3_7, "cdf3.7"},   {WT_CDF2_8, "cdf2.8"},
        {WT_CDF3_9, "cdf3.9"},   {WT_CDF3_11, "cdf3.11"}, {WT_CDF2_10, "cdf2.10"},
        {WT_CDF2_12, "cdf2.12"}, {WT_CDF4_8, "cdf4.8"},   {WT_CDF4_10, "cdf4.10"},
        {WT_CDF2_0, "cdf2.0"},   {WT_CDF1_1, "cdf1.1"},   {WT_CDF1_7, "cdf1.7"}};
    *count = (int)(sizeof(t) / sizeof(t[0]));
    return t;
}
inline WaveletType type_from_string(const std::string& str) {
    int n = 0;
    const auto* t = wavelet_names(&n);
    const std::string tail = str.size() > 2 ? str.substr(2) : std::string();
    for (int i = 0; i < n; ++i)
        if (tail == t[i].second) return t[i].first;
    return WT_CDF3_5;
}
inline std::string string_from_type(WaveletType a, char bc) {
    int n = 0;
    const auto* t = wavelet_names(&n);
    for (int i = 0; i < n; ++i)
        if (t[i].first == a) return std::string(1, bc) + "_" + t[i].second;
    throw Error(WHHD_UNSUPPORTED_WAVELET, "string_from_type");
}

// Handles are move-only owners of a C-ABI handle.
template <class H, void (*Destroy)(H)>
class Owned {
public:
    Owned() = default;
    Owned(const Owned&) = delete;
    Owned& operator=(const Owned&) = delete;
    Owned(Owned&& o) noexcept : h_(std::exchange(o.h_, nullptr)), stream_(o.stream_) {}
    ~Owned() { if (h_) Destroy(h_); }
    void SetStream(hipStream_t s) { stream_ = s; }
    hipStream_t Stream() const { return stream_; }

protected:
    H h_ = nullptr;
    hipStream_t stream_ = nullptr;
};

class CuWaveletPotentialRecover3D
    : public Owned<whhd_recover3d_t, whhd_recover3d_destroy> {
public:
    CuWaveletPotentialRecover3D() = delete;
    // dim: cells per axis (powers of two)
    CuWaveletPotentialRecover3D(int3 dim, DTYPE3 dx, char bc, WaveletType type_w_curl,
                                WaveletType type_w_div) {
        check(whhd_recover3d_create(dim.x, dim.y, dim.z, dx.x, dx.y, dx.z, bc, type_w_curl,
                                    type_w_div, &h_),
              "CuWaveletPotentialRecover3D");
    }
    void Solve(DTYPE* qx, DTYPE* qy, DTYPE* qz, DTYPE* xv, DTYPE* yv, DTYPE* zv) {
        check(whhd_recover3d_solve(h_, qx, qy, qz, xv, yv, zv, stream_), "Solve");
    }
    void Solve(DTYPE* qx, DTYPE* qy, DTYPE* qz, DTYPE* xv_out, DTYPE* yv_out, DTYPE* zv_out,
               DTYPE* xv, DTYPE* yv, DTYPE* zv) {
        check(whhd_recover3d_solve_proj(h_, qx, qy, qz, xv_out, yv_out, zv_out, xv, yv, zv,
                                        stream_),
              "Solve");
    }
    void CutCoefficients(DTYPE* qx_out, DTYPE* qy_out, DTYPE* qz_out, DTYPE* qx, DTYPE* qy,
                         DTYPE* qz, int level) {
        check(whhd_recover3d_cut(h_, qx_out, qy_out, qz_out, qx, qy, qz, level, stream_),
              "CutCoefficients");
    }
};

class CuWaveletPotentialRecover2D
    : public Owned<whhd_recover2d_t, whhd_recover2d_destroy> {
public:
    CuWaveletPotentialRecover2D() = delete;
    CuWaveletPotentialRecover2D(int2 dim, DTYPE2 dx, char bc, WaveletType type_w_curl,
                                WaveletType type_w_div) {
        check(whhd_recover2d_create(dim.x, dim.y, dx.x, dx.y, bc, type_w_curl, type_w_div, &h_),
              "CuWaveletPotentialRecover2D");
    }
    void Solve(DTYPE* qz, DTYPE* xv, DTYPE* yv) {
        check(whhd_recover2d_solve(h_, qz, xv, yv, stream_), "Solve");
    }
};

class CuDivLocalProject3D : public Owned<whhd_project3d_t, whhd_project3d_destroy> {
public:
    // dim = cells + 1 per axis, as the reference passes it
    // (WaveletPotentials/cuWaveletPotentialRecover3D.cu:10)
    CuDivLocalProject3D(int3 dim, int3 levels, DTYPE3 dx, WaveletType type, char bc)
        : levels_(levels) {
        check(whhd_project3d_create(dim.x - 1, dim.y - 1, dim.z - 1, dx.x, dx.y, dx.z, type, bc,
                                    &h_),
              "CuDivLocalProject3D");
    }
    void ProjectLocal_q_ccc(DTYPE* q_lwt_x, DTYPE* q_lwt_y, DTYPE* q_lwt_z, DTYPE* xv_lwt,
                            DTYPE* yv_lwt, DTYPE* zv_lwt, int3 levels) {
        same_levels(levels);
        check(whhd_project3d_run(h_, q_lwt_x, q_lwt_y, q_lwt_z, nullptr, nullptr, nullptr, xv_lwt,
                                 yv_lwt, zv_lwt, stream_),
              "ProjectLocal_q_ccc");
    }
    void ProjectLocal_q_ccc(DTYPE* q_lwt_x, DTYPE* q_lwt_y, DTYPE* q_lwt_z, DTYPE* xv_lwt_out,
                            DTYPE* yv_lwt_out, DTYPE* zv_lwt_out, DTYPE* xv_lwt, DTYPE* yv_lwt,
                            DTYPE* zv_lwt, int3 levels) {
        same_levels(levels);
        check(whhd_project3d_run(h_, q_lwt_x, q_lwt_y, q_lwt_z, xv_lwt_out, yv_lwt_out,
                                 zv_lwt_out, xv_lwt, yv_lwt, zv_lwt, stream_),
              "ProjectLocal_q_ccc");
    }

private:
    void same_levels(int3 l) const {
        if (l.x != levels_.x || l.y != levels_.y || l.z != levels_.z)
            throw Error(WHHD_INVALID_ARG, "ProjectLocal_q_ccc: levels differ from the constructor's");
    }
    int3 levels_;
};

// Singleton like the reference (WaveletPotentials/cuWHHDForward.cuh:26).
class CuWHHDForward {
public:
    static CuWHHDForward* GetInstance() {
        static CuWHHDForward inst;
        return &inst;
    }
    void SetStream(hipStream_t s) { stream_ = s; }
    void Solve1D(DTYPE* dst, DTYPE* src, int3 dim, int level, char direction, char bc = 'n',
                 WaveletType rsw_type = WT_CDF3_5, bool is_forward = true) {
        check(whhd_transform1d(dst, src, dim.x, dim.y, dim.z, level, direction, bc, rsw_type,
                               is_forward ? 1 : 0, stream_),
              "Solve1D");
    }

private:
    hipStream_t stream_ = nullptr;
};

// Gram tables, cached per key string + n (WaveletPotentials/cuGetInnerProduct.cu:
// 85-126, cuGetInnerProduct.cuh:15-19).  Keys are "<bc>_cdf<p>.<q>" as built by
// string_from_type.  GenerateIp returns the cached table when the key exists
// and builds it otherwise; GetIp returns the cached table or nullptr.  A table
// is 8 rows of n DTYPE (device memory owned by the instance): the diagonal,
// the -2, -1, +1, +2 neighbours and the three parent entries.
class CuGetInnerProduct {
public:
    static CuGetInnerProduct* GetInstance() {
        static CuGetInnerProduct inst;
        return &inst;
    }
    ~CuGetInnerProduct() {
        for (auto& kv : tables_) (void)hipFree(kv.second);
    }
    DTYPE* GetIp(std::string str, int dim) {
        auto it = tables_.find(str + std::to_string(dim));
        return it == tables_.end() ? nullptr : it->second;
    }
    DTYPE* GenerateIp(std::string str, int dim) {
        const std::string key = str + std::to_string(dim);
        auto it = tables_.find(key);
        if (it != tables_.end()) return it->second;
        if (str.size() < 3) throw Error(WHHD_INVALID_ARG, "GenerateIp: key is not <bc>_cdf<p>.<q>");
        const WaveletType type = type_from_string(str);
        const size_t n = 8 * (size_t)dim;
        double* d64 = nullptr;
        if (hipMalloc(&d64, sizeof(double) * n) != hipSuccess) throw Error(WHHD_OOM, "GenerateIp");
        int st = whhd_gram_table(d64, type, str[0], dim, nullptr);
        std::vector<double> h64(n);
        if (st == WHHD_OK &&
            hipMemcpy(h64.data(), d64, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
            st = WHHD_HIP_ERROR;
        (void)hipFree(d64);
        if (st != WHHD_OK) throw Error(st, "GenerateIp");
        std::vector<DTYPE> h(h64.begin(), h64.end());
        DTYPE* d = nullptr;
        if (hipMalloc(&d, sizeof(DTYPE) * n) != hipSuccess) throw Error(WHHD_OOM, "GenerateIp");
        if (hipMemcpy(d, h.data(), sizeof(DTYPE) * n, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(d);
            throw Error(WHHD_HIP_ERROR, "GenerateIp");
        }
        tables_[key] = d;
        return d;
    }
    DTYPE* GetIp(WaveletType type, char bc, int dim) {
        return GetIp(string_from_type(type, bc), dim);
    }
    DTYPE* GenerateIp(WaveletType type, char bc, int dim) {
        return GenerateIp(string_from_type(type, bc), dim);
    }

private:
    std::map<std::string, DTYPE*> tables_;
};

// ---- solid boundary, fractional multigrid and field operators -------------

// Owns the cell-centred level set ls_ (device), as the reference does.
class CuSolidLevelSet3D {
public:
    CuSolidLevelSet3D() = delete;
    CuSolidLevelSet3D(int3 dim, DTYPE3 dx) : dim_(dim), dx_(dx) {
        if (hipMalloc(&ls_, sizeof(DTYPE) * (size_t)dim.x * dim.y * dim.z) != hipSuccess)
            throw Error(WHHD_OOM, "CuSolidLevelSet3D");
    }
    CuSolidLevelSet3D(const CuSolidLevelSet3D&) = delete;
    CuSolidLevelSet3D& operator=(const CuSolidLevelSet3D&) = delete;
    ~CuSolidLevelSet3D() { (void)hipFree(ls_); }
    void InitLsSphere(DTYPE3 center, DTYPE radius) { InitLsSphere(ls_, center, radius, dim_, dx_); }
    // the member overload: the two-point formula (cuGetAx/Ay/Az,
    // WaveletPotentials/cuSolidLevelSet3D.cu:36-90, 348-353)
    void GetFrac(DTYPE* ax, DTYPE* ay, DTYPE* az) {
        check(whhd_frac_from_ls(ax, ay, az, ls_, dim_.x, dim_.y, dim_.z, dx_.x, dx_.y, dx_.z,
                                nullptr),
              "GetFrac");
    }
    static void InitLsSphere(DTYPE* ls, DTYPE3 center, DTYPE radius, int3 dim, DTYPE3 dx) {
        const double c[3] = {center.x, center.y, center.z}, r = radius;
        check(whhd_levelset_spheres(ls, dim.x, dim.y, dim.z, dx.x, dx.y, dx.z, c, &r, 1, nullptr),
              "InitLsSphere");
    }
    // the static overload: supersampled fractions (cuGetAx/Ay/Az_divide,
    // WaveletPotentials/cuSolidLevelSet3D.cu:164-317, 359-372)
    static void GetFrac(DTYPE* ax, DTYPE* ay, DTYPE* az, DTYPE* ls, int3 dim, DTYPE3 dx) {
        check(whhd_frac_divide(ax, ay, az, ls, dim.x, dim.y, dim.z, dx.x, dx.y, dx.z, nullptr),
              "GetFrac");
    }
    DTYPE* ls_ = nullptr;

private:
    int3 dim_;
    DTYPE3 dx_;
};

// Eikonal relaxation; its scratch array is kept between calls.
class CuFastSweeping {
public:
    static CuFastSweeping* GetInstance() {
        static CuFastSweeping inst;
        return &inst;
    }
    ~CuFastSweeping() { (void)hipFree(scratch_); }
    static void Solve(DTYPE* ls, int3 dim, DTYPE dx, int n_iters = 5) {
        CuFastSweeping* s = GetInstance();
        const size_t n = (size_t)dim.x * dim.y * dim.z;
        if (n > s->size_) {
            (void)hipFree(s->scratch_);
            s->scratch_ = nullptr;
            s->size_ = 0;
            if (hipMalloc(&s->scratch_, n * sizeof(DTYPE)) != hipSuccess)
                throw Error(WHHD_OOM, "CuFastSweeping::Solve");
            s->size_ = n;
        }
        check(whhd_eikonal_sweep(ls, s->scratch_, dim.x, dim.y, dim.z, dx, n_iters, nullptr),
              "CuFastSweeping::Solve");
    }

private:
    DTYPE* scratch_ = nullptr;
    size_t size_ = 0;
};

class CuMultigridFrac3D : public Owned<whhd_mgfrac3d_t, whhd_mgfrac3d_destroy> {
public:
    CuMultigridFrac3D(DTYPE* Ax, DTYPE* Ay, DTYPE* Az, int3 dim, DTYPE3 dx, char bc = 'n') {
        check(whhd_mgfrac3d_create(Ax, Ay, Az, dim.x, dim.y, dim.z, dx.x, dx.y, dx.z, bc, &h_),
              "CuMultigridFrac3D");
    }
    void Solve(DTYPE* dst, DTYPE* pre_v, DTYPE* f, int n_iters = 10, int nu1 = 2, int nu2 = 2,
               DTYPE weight = 0.8f) {
        check(whhd_mgfrac3d_solve(h_, dst, pre_v, f, n_iters, nu1, nu2, weight, stream_), "Solve");
    }
    DTYPE GetResdualError(DTYPE* dst, DTYPE* f) {
        double rel = 0.0;
        check(whhd_mgfrac3d_residual(h_, dst, f, &rel, stream_), "GetResdualError");
        return (DTYPE)rel;
    }
};

class CuGradient {
public:
    static CuGradient* GetInstance() {
        static CuGradient inst;
        return &inst;
    }
    void SetStream(hipStream_t s) { stream_ = s; }
    void Gradient3D_Frac_P(DTYPE* xv, DTYPE* yv, DTYPE* zv, DTYPE* p, DTYPE* ax, DTYPE* ay,
                           DTYPE* az, int3 dim, DTYPE3 dx, char bc = 'n') {
        check(whhd_gradient3d_frac(xv, yv, zv, p, ax, ay, az, dim.x, dim.y, dim.z, dx.x, dx.y, dx.z,
                                   bc, stream_),
              "Gradient3D_Frac_P");
    }

private:
    hipStream_t stream_ = nullptr;
};

class CuConvergence {
public:
    static CuConvergence* GetInstance() {
        static CuConvergence inst;
        return &inst;
    }
    void SetStream(hipStream_t s) { stream_ = s; }
    // b = -div(a u); bc only selects the reference's overload (the walls are closed)
    void GetB_3d_frac_vel(DTYPE* b, DTYPE* xv, DTYPE* yv, DTYPE* zv, DTYPE* ax, DTYPE* ay,
                          DTYPE* az, int3 dim, DTYPE3 dx, char bc = 'n') {
        (void)bc;
        check(whhd_divergence3d_frac(b, xv, yv, zv, ax, ay, az, dim.x, dim.y, dim.z, dx.x, dx.y,
                                     dx.z, stream_),
              "GetB_3d_frac_vel");
    }

private:
    hipStream_t stream_ = nullptr;
};

inline void CudaMatAdd(DTYPE* dst, DTYPE* src, int size, DTYPE scale = 1.f) {
    check(whhd_axpy(dst, src, size, scale, nullptr), "CudaMatAdd");
}
inline void CudaMatMul(DTYPE* dst, DTYPE* lsrc, DTYPE* rsrc, int size) {
    check(whhd_mul(dst, lsrc, rsrc, size, nullptr), "CudaMatMul");
}
namespace detail {
// device scratch of the reductions, kept for the process (the reference's
// CuMemoryManager keeps its buffers too)
inline DTYPE reduce(const DTYPE* x, int size, int op, const char* what) {
    static double* scratch = nullptr;
    if (!scratch && hipMalloc(&scratch, sizeof(double) * whhd_reduce_scratch_doubles()) != hipSuccess)
        throw Error(WHHD_OOM, what);
    double out = 0.0;
    check(whhd_reduce(x, size, op, scratch, &out, nullptr), what);
    return (DTYPE)out;
}
}  // namespace detail
inline DTYPE CudaArraySum(DTYPE* src, int size) {
    return detail::reduce(src, size, WHHD_REDUCE_SUM, "CudaArraySum");
}
inline DTYPE CudaArrayAbsSum(DTYPE* src, int size) {
    return detail::reduce(src, size, WHHD_REDUCE_SUM_ABS, "CudaArrayAbsSum");
}
inline DTYPE CudaArrayNormSum(DTYPE* src, int size) {
    return detail::reduce(src, size, WHHD_REDUCE_SUM_SQ, "CudaArrayNormSum");
}
inline DTYPE CudaFindMaxValue(DTYPE* src, int size) {
    return detail::reduce(src, size, WHHD_REDUCE_MAX_ABS, "CudaFindMaxValue");
}

class CuDst3D : public Owned<whhd_dst3d_t, whhd_dst3d_destroy> {
public:
    CuDst3D() = delete;
    CuDst3D(int3 dim, DTYPE3 dx) {
        check(whhd_dst3d_create(dim.x, dim.y, dim.z, dx.x, dx.y, dx.z, &h_), "CuDst3D");
    }
    void Solve(DTYPE* dst, DTYPE* f) { check(whhd_dst3d_solve(h_, dst, f, stream_), "Solve"); }
};

class CuDst : public Owned<whhd_dst3d_t, whhd_dst3d_destroy> {
public:
    CuDst() = delete;
    CuDst(int2 dim, DTYPE2 dx) { check(whhd_dst2d_create(dim.x, dim.y, dx.x, dx.y, &h_), "CuDst"); }
    void Solve(DTYPE* dst, DTYPE* f) { check(whhd_dst3d_solve(h_, dst, f, stream_), "Solve"); }
};

// bc 'n', 'd' or 'z'; the DST solve is direct, so tol is not used.
class CuParallelSweep3D : public Owned<whhd_sweep3d_t, whhd_sweep3d_destroy> {
public:
    CuParallelSweep3D(int3 dim, DTYPE3 dx, char bc) {
        check(whhd_sweep3d_create(dim.x, dim.y, dim.z, dx.x, dx.y, dx.z, bc, &h_),
              "CuParallelSweep3D");
    }
    CuParallelSweep3D(int2 dim, DTYPE2 dx, char bc) {
        check(whhd_sweep2d_create(dim.x, dim.y, dx.x, dx.y, bc, &h_), "CuParallelSweep3D");
    }
    void Solve(DTYPE* qz, DTYPE* xv, DTYPE* yv) {
        check(whhd_sweep2d_solve(h_, qz, xv, yv, stream_), "Solve");
    }
    void Solve(DTYPE* qx, DTYPE* qy, DTYPE* qz, DTYPE* xv, DTYPE* yv, DTYPE* zv) {
        check(whhd_sweep3d_solve(h_, qx, qy, qz, xv, yv, zv, stream_), "Solve");
    }
    void ProjectZero(DTYPE* qx, DTYPE* qy, DTYPE* qz, DTYPE tol = 1e-6f) {
        (void)tol;
        check(whhd_sweep3d_project(h_, qx, qy, qz, stream_), "ProjectZero");
    }
    void Project(DTYPE* qx, DTYPE* qy, DTYPE* qz, DTYPE tol = 1e-6f) { ProjectZero(qx, qy, qz, tol); }
};

// Host verification helpers of the reference's tests (host arrays).
inline void SetSliceZero(DTYPE* val, int3 dim, char ns) {
    check(whhd_set_slice_zero(val, dim.x, dim.y, dim.z, ns), "SetSliceZero");
}

class StreamInitial {
public:
    static StreamInitial* GetInstance() {
        static StreamInitial inst;
        return &inst;
    }
    void GradientStream(DTYPE* xv, DTYPE* yv, DTYPE* zv, DTYPE* qx, DTYPE* qy, DTYPE* qz,
                        int3 dim, DTYPE3 dx) {
        check(whhd_gradient_stream3d(xv, yv, zv, qx, qy, qz, dim.x, dim.y, dim.z, dx.x, dx.y, dx.z),
              "GradientStream");
    }
    void GradientStream(DTYPE* xv, DTYPE* yv, DTYPE* qz, int2 dim, DTYPE2 dx) {
        check(whhd_gradient_stream2d(xv, yv, qz, dim.x, dim.y, dx.x, dx.y), "GradientStream");
    }
};

inline DTYPE CalculateErrorL2(const DTYPE* value1, const DTYPE* value2, int3 size3) {
    const float* b[1] = {value1};
    const float* c[1] = {value2};
    const long long n[1] = {(long long)size3.x * size3.y * size3.z};
    double out = 0.0;
    check(whhd_error_l2(b, c, n, 1, &out), "CalculateErrorL2");
    return (DTYPE)out;
}
inline DTYPE CalculateErrorL2(const DTYPE* base1, const DTYPE* check1, int3 size1,
                              const DTYPE* base2, const DTYPE* check2, int3 size2,
                              const DTYPE* base3, const DTYPE* check3, int3 size3) {
    const float* b[3] = {base1, base2, base3};
    const float* c[3] = {check1, check2, check3};
    const long long n[3] = {(long long)size1.x * size1.y * size1.z,
                            (long long)size2.x * size2.y * size2.z,
                            (long long)size3.x * size3.y * size3.z};
    double out = 0.0;
    check(whhd_error_l2(b, c, n, 3, &out), "CalculateErrorL2");
    return (DTYPE)out;
}

}  // namespace whhd_compat
