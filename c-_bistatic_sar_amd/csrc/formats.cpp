// formats.cpp -- the reference's stage-boundary file formats.
//
// Binary matrix (binary_reader.cpp:19-88, data_export.cpp:87-125, 207-256),
// little endian:  u32 ndims | u32 isComplex | u32 typeLen | char[16] type,
// NUL padded | u64 dims[ndims] | float64 payload, (re, im) interleaved when
// complex.  CSV vectors are one comma-separated line of "%.6f" values
// (data_export.cpp:161-178).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bsar.h"

namespace {

struct File {
    FILE* f = nullptr;
    ~File() { if (f) std::fclose(f); }
    // flushes and closes a file that was written: BSAR_ERR_IO when the last
    // buffered block cannot be written (disk full) or an earlier write failed
    int close_written()
    {
        if (!f) return BSAR_ERR_IO;
        const bool bad = std::fflush(f) != 0 || std::ferror(f) != 0;
        const bool closed = std::fclose(f) == 0;
        f = nullptr;
        return (bad || !closed) ? BSAR_ERR_IO : BSAR_OK;
    }
};

int write_header(FILE* f, uint32_t ndims, const uint64_t* dims)
{
    const uint32_t isComplex = 1, typeLen = 6;
    char type[16] = {0};
    std::memcpy(type, "double", 6);
    if (std::fwrite(&ndims, 4, 1, f) != 1 || std::fwrite(&isComplex, 4, 1, f) != 1 ||
        std::fwrite(&typeLen, 4, 1, f) != 1 || std::fwrite(type, 1, 16, f) != 16 ||
        std::fwrite(dims, 8, ndims, f) != ndims)
        return BSAR_ERR_IO;
    return BSAR_OK;
}

// exportVector, data_export.cpp:161-178: std::fixed << setprecision(6).
int write_csv_vector(const std::string& path, const double* v, int n)
{
    File fh;
    fh.f = std::fopen(path.c_str(), "w");
    if (!fh.f) return BSAR_ERR_IO;
    for (int i = 0; i < n; ++i) {
        if (i > 0) std::fputc(',', fh.f);
        std::fprintf(fh.f, "%.6f", v[i]);
    }
    std::fputc('\n', fh.f);
    return fh.close_written();
}

}  // namespace

extern "C" {

static int read_matrix_impl(const char* path, uint64_t* dims, int max_dims, int* ndims_out, int* is_complex,
                            double* buf, uint64_t* count);

int bsar_read_matrix(const char* path, uint64_t* dims, int max_dims, int* ndims_out, int* is_complex,
                     double* buf, uint64_t* count)
{
    try {
        return read_matrix_impl(path, dims, max_dims, ndims_out, is_complex, buf, count);
    } catch (const std::bad_alloc&) {
        return BSAR_ERR_NOMEM;
    }
}

static int read_matrix_impl(const char* path, uint64_t* dims, int max_dims, int* ndims_out, int* is_complex,
                            double* buf, uint64_t* count)
{
    if (!path) return BSAR_ERR_INVALID;
    File fh;
    fh.f = std::fopen(path, "rb");
    if (!fh.f) return BSAR_ERR_IO;
    uint32_t nd = 0, cflag = 0, tlen = 0;
    char type[16];
    if (std::fread(&nd, 4, 1, fh.f) != 1) return BSAR_ERR_FORMAT;
    if (std::fread(&cflag, 4, 1, fh.f) != 1) return BSAR_ERR_FORMAT;
    if (std::fread(&tlen, 4, 1, fh.f) != 1) return BSAR_ERR_FORMAT;
    if (std::fread(type, 1, 16, fh.f) != 16) return BSAR_ERR_FORMAT;
    if (nd > 64) return BSAR_ERR_FORMAT;
    std::vector<uint64_t> d(nd);
    if (nd && std::fread(d.data(), 8, nd, fh.f) != nd) return BSAR_ERR_FORMAT;
    // element count; a header whose dims overflow 2^60 elements is malformed
    uint64_t total = 1;
    for (uint64_t x : d) {
        if (x && total > ((uint64_t)1 << 60) / x) return BSAR_ERR_FORMAT;
        total *= x;
    }
    const bool cplx = (cflag == 1);
    if (ndims_out) *ndims_out = (int)nd;
    if (is_complex) *is_complex = cplx ? 1 : 0;
    if (count) *count = total;
    if (dims)
        for (int i = 0; i < max_dims && i < (int)nd; ++i) dims[i] = d[i];
    if (!buf) return BSAR_OK;
    const uint64_t nvals = total * (cplx ? 2 : 1);
    if (cplx) {
        if (std::fread(buf, 8, nvals, fh.f) != nvals) return BSAR_ERR_FORMAT;
    } else {
        // real payload -> (v, 0) pairs (binary_reader.cpp:83-88)
        std::vector<double> tmp(nvals);
        if (nvals && std::fread(tmp.data(), 8, nvals, fh.f) != nvals) return BSAR_ERR_FORMAT;
        for (uint64_t i = 0; i < nvals; ++i) {
            buf[2 * i] = tmp[i];
            buf[2 * i + 1] = 0.0;
        }
    }
    return BSAR_OK;
}

int bsar_write_matrix2d(const char* path, const double* data, uint64_t rows, uint64_t cols)
{
    if (!path || (rows * cols && !data)) return BSAR_ERR_INVALID;
    File fh;
    fh.f = std::fopen(path, "wb");
    if (!fh.f) return BSAR_ERR_IO;
    const uint64_t dims[2] = {rows, cols};
    if (write_header(fh.f, 2, dims)) return BSAR_ERR_IO;
    const uint64_t n = rows * cols * 2;
    if (n && std::fwrite(data, 8, n, fh.f) != n) return BSAR_ERR_IO;
    return fh.close_written();
}

int bsar_write_image(const char* path, const double* img, int nx, int ny, int nz, const double* gx,
                     const double* gy, const double* gz)
{
    if (!path || nx < 0 || ny < 0 || nz < 0) return BSAR_ERR_INVALID;
    const uint64_t total = (uint64_t)nx * ny * nz;
    if (total && !img) return BSAR_ERR_INVALID;
    {
        File fh;
        fh.f = std::fopen(path, "wb");
        if (!fh.f) return BSAR_ERR_IO;
        const uint64_t dims[3] = {(uint64_t)nx, (uint64_t)ny, (uint64_t)nz};
        if (write_header(fh.f, 3, dims)) return BSAR_ERR_IO;
        // data_export.cpp:241-256: z outermost, x fastest -- the layout of img.
        if (total && std::fwrite(img, 16, total, fh.f) != total) return BSAR_ERR_IO;
        if (int st = fh.close_written()) return st;
    }
    // data_export.cpp:292-300: grid axes beside the image when all are given.
    if (gx && gy && gz && nx > 0 && ny > 0 && nz > 0) {
        std::string base;
        try {
            base = path;
            base = base.substr(0, base.find_last_of('.'));
        } catch (const std::bad_alloc&) {
            return BSAR_ERR_NOMEM;
        }
        int st = write_csv_vector(base + "_grid_x.csv", gx, nx);
        if (!st) st = write_csv_vector(base + "_grid_y.csv", gy, ny);
        if (!st) st = write_csv_vector(base + "_grid_z.csv", gz, nz);
        if (st) return st;
    }
    return BSAR_OK;
}

}  // extern "C"
