// Field I/O on host arrays: the reference's dump formats
// (WaveletPotentials/Utils.cpp:812-990) and a PLY point cloud for exchanging
// fields with other tools.  Arrays are y fastest, element (x, y, z) of an
// (nx, ny, nz) array at y + x*ny + z*nx*ny; a text row holds the ny values of
// one (x, z), rows in memory order.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "whhd_internal.h"

namespace {

using whhd::fail;

int64_t count(int nx, int ny, int nz) { return (int64_t)nx * ny * nz; }

bool dims_ok(int nx, int ny, int nz) { return nx > 0 && ny > 0 && nz > 0; }

struct File {
    FILE* f;
    explicit File(const char* path, const char* mode) : f(path ? std::fopen(path, mode) : nullptr) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

}  // namespace

extern "C" {

// WriteDataTxt / WriteDataETxt (Utils.cpp:830-855, 938-958): `digits` after the
// point, fixed (scientific = 0) or scientific notation.
WHHD_API int whhd_write_txt(const char* path, const float* data, int nx, int ny, int nz,
                            int digits, int scientific) {
    if (!data || !dims_ok(nx, ny, nz) || digits < 0 || digits > 17)
        return fail(WHHD_INVALID_ARG, "write_txt: bad argument");
    File fp(path, "w");
    if (!fp.f) return fail(WHHD_INVALID_ARG, std::string("write_txt: cannot open ") + (path ? path : ""));
    const char* fmt = scientific ? "%.*e" : "%.*f";
    const int64_t rows = (int64_t)nx * nz;
    for (int64_t j = 0; j < rows; ++j)
        for (int i = 0; i < ny; ++i) {
            std::fprintf(fp.f, fmt, digits, (double)data[j * ny + i]);
            std::fputc(i < ny - 1 ? ' ' : '\n', fp.f);
        }
    return std::ferror(fp.f) ? fail(WHHD_INVALID_ARG, "write_txt: write failed") : WHHD_OK;
}

// ReadsDataTxt (Utils.cpp:960-971): whitespace-separated values in memory order.
WHHD_API int whhd_read_txt(const char* path, float* data, int nx, int ny, int nz) {
    if (!data || !dims_ok(nx, ny, nz)) return fail(WHHD_INVALID_ARG, "read_txt: bad argument");
    File fp(path, "r");
    if (!fp.f) return fail(WHHD_INVALID_ARG, std::string("read_txt: cannot open ") + (path ? path : ""));
    const int64_t n = count(nx, ny, nz);
    for (int64_t k = 0; k < n; ++k) {
        double v;
        if (std::fscanf(fp.f, "%lf", &v) != 1)
            return fail(WHHD_INVALID_ARG, "read_txt: fewer values than the array holds");
        data[k] = (float)v;
    }
    return WHHD_OK;
}

// WriteDataTxt_sparse (Utils.cpp:857-882): "x,y,z,value" for |value| >= threshold,
// z outermost then x then y, values in scientific notation with 5 digits.
WHHD_API int whhd_write_txt_sparse(const char* path, const float* data, int nx, int ny, int nz,
                                   double threshold) {
    if (!data || !dims_ok(nx, ny, nz)) return fail(WHHD_INVALID_ARG, "write_txt_sparse: bad argument");
    File fp(path, "w");
    if (!fp.f) return fail(WHHD_INVALID_ARG, "write_txt_sparse: cannot open the file");
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const float v = data[y + (int64_t)x * ny + (int64_t)z * nx * ny];
                if (std::fabs(v) >= threshold) std::fprintf(fp.f, "%d,%d,%d,%.5e\n", x, y, z, (double)v);
            }
    return std::ferror(fp.f) ? fail(WHHD_INVALID_ARG, "write_txt_sparse: write failed") : WHHD_OK;
}

// WriteVectorTxt_sparse (Utils.cpp:884-908) for three cell arrays:
// "x,y,z,vx,vy,vz" where the vector's length is >= threshold.
WHHD_API int whhd_write_vector_txt_sparse(const char* path, const float* vx, const float* vy,
                                          const float* vz, int nx, int ny, int nz,
                                          double threshold) {
    if (!vx || !vy || !vz || !dims_ok(nx, ny, nz))
        return fail(WHHD_INVALID_ARG, "write_vector_txt_sparse: bad argument");
    File fp(path, "w");
    if (!fp.f) return fail(WHHD_INVALID_ARG, "write_vector_txt_sparse: cannot open the file");
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const int64_t i = y + (int64_t)x * ny + (int64_t)z * nx * ny;
                const double a = vx[i], b = vy[i], c = vz[i];
                if (std::sqrt(a * a + b * b + c * c) >= threshold)
                    std::fprintf(fp.f, "%d,%d,%d,%.5e,%.5e,%.5e\n", x, y, z, a, b, c);
            }
    return std::ferror(fp.f) ? fail(WHHD_INVALID_ARG, "write_vector_txt_sparse: write failed")
                             : WHHD_OK;
}

// WriteDataBinary (Utils.cpp:910-915): the raw FP32 values; and its reader.
WHHD_API int whhd_write_binary(const char* path, const float* data, long long n) {
    if (!data || n < 0) return fail(WHHD_INVALID_ARG, "write_binary: bad argument");
    File fp(path, "wb");
    if (!fp.f) return fail(WHHD_INVALID_ARG, "write_binary: cannot open the file");
    if (std::fwrite(data, sizeof(float), (size_t)n, fp.f) != (size_t)n)
        return fail(WHHD_INVALID_ARG, "write_binary: write failed");
    return WHHD_OK;
}

WHHD_API int whhd_read_binary(const char* path, float* data, long long n) {
    if (!data || n < 0) return fail(WHHD_INVALID_ARG, "read_binary: bad argument");
    File fp(path, "rb");
    if (!fp.f) return fail(WHHD_INVALID_ARG, "read_binary: cannot open the file");
    if (std::fread(data, sizeof(float), (size_t)n, fp.f) != (size_t)n)
        return fail(WHHD_INVALID_ARG, "read_binary: fewer values than requested");
    return WHHD_OK;
}

// PLY point cloud of a cell field: one vertex per cell with |value| >= threshold
// at the cell centre ((x + 1/2) dx, (y + 1/2) dy, (z + 1/2) dz) and a float
// property "value"; ASCII, or binary little endian.  Returns the vertex count
// through *written when it is not NULL.
WHHD_API int whhd_write_ply(const char* path, const float* data, int nx, int ny, int nz, double dx,
                            double dy, double dz, double threshold, int binary,
                            long long* written) {
    if (!data || !dims_ok(nx, ny, nz) || !(dx > 0 && dy > 0 && dz > 0))
        return fail(WHHD_INVALID_ARG, "write_ply: bad argument");
    int64_t nv = 0;
    const int64_t n = count(nx, ny, nz);
    for (int64_t i = 0; i < n; ++i)
        if (std::fabs(data[i]) >= threshold) ++nv;
    File fp(path, binary ? "wb" : "w");
    if (!fp.f) return fail(WHHD_INVALID_ARG, "write_ply: cannot open the file");
    std::fprintf(fp.f,
                 "ply\nformat %s 1.0\ncomment cell field, %d x %d x %d cells\n"
                 "element vertex %lld\nproperty float x\nproperty float y\nproperty float z\n"
                 "property float value\nend_header\n",
                 binary ? "binary_little_endian" : "ascii", nx, ny, nz, (long long)nv);
    for (int z = 0; z < nz; ++z)
        for (int x = 0; x < nx; ++x)
            for (int y = 0; y < ny; ++y) {
                const float v = data[y + (int64_t)x * ny + (int64_t)z * nx * ny];
                if (!(std::fabs(v) >= threshold)) continue;
                const float p[4] = {(float)((x + 0.5) * dx), (float)((y + 0.5) * dy),
                                    (float)((z + 0.5) * dz), v};
                if (binary) {
                    uint8_t b[16];
                    for (int k = 0; k < 4; ++k) {
                        uint32_t u;
                        std::memcpy(&u, &p[k], 4);
                        for (int e = 0; e < 4; ++e) b[4 * k + e] = (uint8_t)(u >> (8 * e));
                    }
                    std::fwrite(b, 1, 16, fp.f);
                } else {
                    std::fprintf(fp.f, "%.9g %.9g %.9g %.9g\n", p[0], p[1], p[2], p[3]);
                }
            }
    if (std::ferror(fp.f)) return fail(WHHD_INVALID_ARG, "write_ply: write failed");
    if (written) *written = nv;
    return WHHD_OK;
}

}  // extern "C"
