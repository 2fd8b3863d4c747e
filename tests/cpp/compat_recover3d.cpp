// Drives the recovery through the reference-style C++ classes of
// include/whhd_compat.hpp, as a translation unit written against the
// reference headers would after switching the include.
//
//   compat_recover3d N DIR   reads DIR/{xv,yv,zv}.f32 (raw FP32, y fastest),
//                            writes DIR/{qx,qy,qz}.f32 and DIR/solve1d.f32
//                            (forward CDF3.7 'n' transform of xv along z),
//                            and the Gram tables DIR/gram_d.f32 ("d_cdf4.6",
//                            N+1) and DIR/gram_n.f32 ("n_cdf3.7", N) after
//                            checking the cache semantics of
//                            CuGetInnerProduct (GetIp is nullptr before
//                            GenerateIp; a second GenerateIp of a key returns
//                            the first pointer)
//   compat_recover3d solids N DIR cx cy cz r
//                            the solid-obstacle steps of the reference's test
//                            driver on DIR/{xv,yv,zv}.f32 (sphere at c, radius
//                            r, unit cube): level set, fractions, f, 10
//                            V-cycles, gradient correction; writes DIR/ls.f32,
//                            DIR/f.f32, DIR/p.f32 and prints the residual and
//                            max |f|; then the sweep + DST comparator on the
//                            input velocity, writing DIR/{sx,sy,sz}.f32; and
//                            the static GetFrac overload (supersampled
//                            fractions) of the level set, DIR/{dax,day,daz}.f32
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

static int solids(int argc, char** argv) {
    if (argc != 8) return 2;
    const int N = std::atoi(argv[2]);
    const std::string dir = argv[3];
    if (N < 4 || N > 256) return 2;
    const float3 c = make_float3(std::atof(argv[4]), std::atof(argv[5]), std::atof(argv[6]));
    const float r = std::atof(argv[7]);
    const size_t nv = (size_t)(N + 1) * N * N, nc = (size_t)N * N * N,
                 nq = (size_t)(N + 1) * (N + 1) * N;
    std::vector<float> hx = read_f32(dir + "/xv.f32", nv), hy = read_f32(dir + "/yv.f32", nv),
                       hz = read_f32(dir + "/zv.f32", nv);
    float *xv = to_dev(hx), *yv = to_dev(hy), *zv = to_dev(hz);
    float *ax, *ay, *az, *f, *p, *gx, *gy, *gz, *qx, *qy, *qz;
    for (float** a : {&ax, &ay, &az, &gx, &gy, &gz})
        if (hipMalloc(a, nv * 4)) return 3;
    for (float** a : {&f, &p})
        if (hipMalloc(a, nc * 4)) return 3;
    for (float** a : {&qx, &qy, &qz})
        if (hipMalloc(a, nq * 4)) return 3;
    std::vector<float> ls;
    float res = 0.f, fmax0 = 0.f, fmax1 = 0.f;
    try {
        const int3 dim = make_int3(N, N, N);
        const float h = 1.0f / N;
        const float3 dx = make_float3(h, h, h);
        CuSolidLevelSet3D sls(dim, dx);
        sls.InitLsSphere(c, r);
        sls.GetFrac(ax, ay, az);
        ls = to_host(sls.ls_, nc);
        // the static overload: supersampled fractions of the same level set
        float *dax, *day, *daz;
        for (float** a : {&dax, &day, &daz})
            if (hipMalloc(a, nv * 4)) return 3;
        CuSolidLevelSet3D::GetFrac(dax, day, daz, sls.ls_, dim, dx);
        write_f32(dir + "/dax.f32", to_host(dax, nv));
        write_f32(dir + "/day.f32", to_host(day, nv));
        write_f32(dir + "/daz.f32", to_host(daz, nv));
        for (float* a : {dax, day, daz}) (void)hipFree(a);
        CuConvergence::GetInstance()->GetB_3d_frac_vel(f, xv, yv, zv, ax, ay, az, dim, dx, 'n');
        fmax0 = CudaFindMaxValue(f, (int)nc);
        CuMultigridFrac3D mg(ax, ay, az, dim, dx, 'n');
        mg.Solve(p, nullptr, f, 10, 2, 2, 0.8f);
        res = mg.GetResdualError(p, f);
        // u <- a (u - grad p) as the reference's driver does it, then the
        // divergence of u - grad p again
        CuGradient::GetInstance()->Gradient3D_Frac_P(gx, gy, gz, p, ax, ay, az, dim, dx, 'n');
        CudaMatAdd(xv, gx, (int)nv, -1.f);
        CudaMatAdd(yv, gy, (int)nv, -1.f);
        CudaMatAdd(zv, gz, (int)nv, -1.f);
        CuConvergence::GetInstance()->GetB_3d_frac_vel(gx, xv, yv, zv, ax, ay, az, dim, dx, 'n');
        fmax1 = CudaFindMaxValue(gx, (int)nc);
        CudaMatMul(xv, xv, ax, (int)nv);
        // the sweep + DST comparator on the original velocity
        float *ox = to_dev(hx), *oy = to_dev(hy), *oz = to_dev(hz);
        CuParallelSweep3D sweep(dim, dx, 'n');
        sweep.Solve(qx, qy, qz, ox, oy, oz);
        sweep.ProjectZero(qx, qy, qz);
        if (hipDeviceSynchronize() != hipSuccess) return 3;
        for (float* a : {ox, oy, oz}) (void)hipFree(a);
    } catch (const Error& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    write_f32(dir + "/ls.f32", ls);
    write_f32(dir + "/f.f32", to_host(f, nc));
    write_f32(dir + "/p.f32", to_host(p, nc));
    write_f32(dir + "/sx.f32", to_host(qx, (size_t)N * (N + 1) * (N + 1)));
    write_f32(dir + "/sy.f32", to_host(qy, (size_t)(N + 1) * N * (N + 1)));
    write_f32(dir + "/sz.f32", to_host(qz, nq));
    for (float* a : {xv, yv, zv, ax, ay, az, f, p, gx, gy, gz, qx, qy, qz}) (void)hipFree(a);
    std::printf("%.9e %.9e %.9e\n", res, fmax0, fmax1);
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 1 && std::string(argv[1]) == "solids") return solids(argc, argv);
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
        // Gram-table cache (WaveletPotentials/cuGetInnerProduct.cu:85-126)
        CuGetInnerProduct* ip = CuGetInnerProduct::GetInstance();
        if (ip->GetIp("d_cdf4.6", N + 1) != nullptr || ip->GetIp(WT_CDF4_6, 'd', N + 1) != nullptr) {
            std::fprintf(stderr, "GetIp returned a table that was never generated\n");
            return 5;
        }
        DTYPE* gd = ip->GenerateIp("d_cdf4.6", N + 1);
        DTYPE* gd2 = ip->GenerateIp(WT_CDF4_6, 'd', N + 1);
        DTYPE* gn = ip->GenerateIp(WT_CDF3_7, 'n', N);
        DTYPE* gn2 = ip->GenerateIp("n_cdf3.7", N);
        if (gd != gd2 || gn != gn2 || ip->GetIp("d_cdf4.6", N + 1) != gd ||
            ip->GetIp(WT_CDF3_7, 'n', N) != gn || gd == gn) {
            std::fprintf(stderr, "GenerateIp did not return the cached table\n");
            return 5;
        }
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
    {
        CuGetInnerProduct* ip = CuGetInnerProduct::GetInstance();
        write_f32(dir + "/gram_d.f32", to_host(ip->GetIp("d_cdf4.6", N + 1), 8 * (size_t)(N + 1)));
        write_f32(dir + "/gram_n.f32", to_host(ip->GetIp("n_cdf3.7", N), 8 * (size_t)N));
    }
    for (float* p : {xv, yv, zv, qx, qy, qz, t}) (void)hipFree(p);
    std::printf("ok\n");
    return 0;
}
