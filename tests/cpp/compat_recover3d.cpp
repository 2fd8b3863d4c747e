// Drives the recovery through the reference-style C++ classes of
// include/whhd_compat.hpp, as a translation unit written against the
// reference headers would after switching the include.
//
//   compat_recover3d N DIR   reads DIR/{xv,yv,zv}.f32 (raw FP32, y fastest),
//                            writes DIR/{qx,qy,qz}.f32 and DIR/solve1d.f32
//                            (forward CDF3.7 'n' transform of xv along z)
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "whhd_compat.hpp"

using namespace whhd_compat;

static std::vector<float> read_f32(const std::string& path, size_t n) {
    std::vector<float> v(n);
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f || std::fread(v.data(), sizeof(float), n, f) != n) {
        std::fprintf(stderr, "cannot read %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

static void write_f32(const std::string& path, const std::vector<float>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(v.data(), sizeof(float), v.size(), f) != v.size()) {
        std::fprintf(stderr, "cannot write %s\n", path.c_str());
        std::exit(2);
    }
    std::fclose(f);
}

static float* to_dev(const std::vector<float>& h) {
    float* d = nullptr;
    if (hipMalloc(&d, h.size() * sizeof(float)) != hipSuccess) std::exit(3);
    if (hipMemcpy(d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        std::exit(3);
    return d;
}

static std::vector<float> to_host(const float* d, size_t n) {
    std::vector<float> h(n);
    if (hipMemcpy(h.data(), d, n * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess) std::exit(3);
    return h;
}

int main(int argc, char** argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s N DIR\n", argv[0]);
        return 2;
    }
    const int N = std::atoi(argv[1]);
    const std::string dir = argv[2];
    if (N < 4 || N > 512) return 2;
    const size_t nv = (size_t)(N + 1) * N * N, nq = (size_t)(N + 1) * (N + 1) * N;
    float* xv = to_dev(read_f32(dir + "/xv.f32", nv));
    float* yv = to_dev(read_f32(dir + "/yv.f32", nv));
    float* zv = to_dev(read_f32(dir + "/zv.f32", nv));
    float *qx, *qy, *qz, *t;
    if (hipMalloc(&qx, nq * 4) || hipMalloc(&qy, nq * 4) || hipMalloc(&qz, nq * 4) ||
        hipMalloc(&t, nv * 4))
        return 3;
    try {
        const float h = 1.0f / N;
        CuWaveletPotentialRecover3D rec(make_int3(N, N, N), make_float3(h, h, h), 'n', WT_CDF3_7,
                                        WT_CDF4_6);
        rec.Solve(qx, qy, qz, xv, yv, zv);
        // xv has extents (N+1, N, N): a z line has N samples
        int lz = 0;
        while ((2 << lz) <= N) ++lz;
        CuWHHDForward::GetInstance()->Solve1D(t, xv, make_int3(N + 1, N, N), lz, 'z', 'n',
                                              WT_CDF3_7, true);
        bool threw = false;
        try {
            CuWaveletPotentialRecover3D bad(make_int3(N + 3, N, N), make_float3(h, h, h), 'n',
                                            WT_CDF3_7, WT_CDF4_6);
        } catch (const Error& e) {
            threw = e.status == WHHD_NOT_POW2;
        }
        if (!threw) {
            std::fprintf(stderr, "a non power-of-two grid was accepted\n");
            return 4;
        }
    } catch (const Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    if (hipDeviceSynchronize() != hipSuccess) return 3;
    write_f32(dir + "/qx.f32", to_host(qx, nq));
    write_f32(dir + "/qy.f32", to_host(qy, nq));
    write_f32(dir + "/qz.f32", to_host(qz, nq));
    write_f32(dir + "/solve1d.f32", to_host(t, nv));
    for (float* p : {xv, yv, zv, qx, qy, qz, t}) (void)hipFree(p);
    std::printf("ok\n");
    return 0;
}
