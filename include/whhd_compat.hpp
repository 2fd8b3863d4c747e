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
// Differences, all deliberate:
//   - DTYPE is float (the reference builds FP64 or FP32 by macro); arrays are
//     device memory, y fastest, as in the reference.
//   - Failures throw whhd_compat::Error (the reference aborts the process).
//   - Work goes to the stream set with SetStream (default: the null stream);
//     the reference always uses the null stream.
//   - CuGetInnerProduct::GetIp returns a device array of 8*n doubles owned by
//     the instance (the reference returns DTYPE* cached by string key).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <stdexcept>
#include <string>
#include <utility>

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

// Gram tables, cached per (type, bc, n) (WaveletPotentials/cuGetInnerProduct.cu:92-126).
class CuGetInnerProduct {
public:
    static CuGetInnerProduct* GetInstance() {
        static CuGetInnerProduct inst;
        return &inst;
    }
    ~CuGetInnerProduct() {
        for (auto& kv : tables_) (void)hipFree(kv.second);
    }
    const double* GetIp(WaveletType type, char bc, int n) {
        const auto key = std::make_pair(std::string(1, bc) + std::to_string((int)type), n);
        auto it = tables_.find(key);
        if (it != tables_.end()) return it->second;
        return GenerateIp(type, bc, n);
    }
    const double* GenerateIp(WaveletType type, char bc, int n) {
        double* d = nullptr;
        if (hipMalloc(&d, sizeof(double) * 8 * (size_t)n) != hipSuccess)
            throw Error(WHHD_OOM, "GenerateIp");
        const int st = whhd_gram_table(d, type, bc, n, nullptr);
        if (st != WHHD_OK) {
            (void)hipFree(d);
            throw Error(st, "GenerateIp");
        }
        const auto key = std::make_pair(std::string(1, bc) + std::to_string((int)type), n);
        auto it = tables_.find(key);
        if (it != tables_.end()) (void)hipFree(it->second);
        tables_[key] = d;
        return d;
    }

private:
    std::map<std::pair<std::string, int>, double*> tables_;
};

}  // namespace whhd_compat
