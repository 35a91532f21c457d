// geotiff.cpp -- single-band GeoTIFF elevation rasters read into the DEM
// layout the back-projection uses (ascending X_vec / Y_vec, Z[y][x], metres).
//
// The reference ships its DEMs as GeoTIFF (assets/*.tif: 1 arc-second WGS84
// rasters, int16, LZW, one 256 x 256 tile) but its loader without GDAL never
// decodes them (load_dem_simple.cpp:16-72 substitutes a flat 32 x 32 DEM);
// the GDAL loader it documents converts the geographic coordinates to
// metres centred at the origin (load_dem.h:9-17).  This reader does that
// without GDAL:
//
//   * TIFF 6.0 (classic, either byte order), strips or tiles, one sample per
//     pixel, 8/16/32/64-bit integer or 32/64-bit float samples;
//   * compression none, LZW (TIFF 6.0 section 13, MSB-first codes, code width
//     growing one code early), Deflate (zlib), PackBits;
//   * predictor none, horizontal differencing (2) and the floating-point
//     predictor (3, Adobe Photoshop TIFF technical note 3) on LZW and
//     Deflate data;
//   * georeferencing from ModelPixelScale + ModelTiepoint (or a
//     ModelTransformation without rotation), GeoKeys GTModelType,
//     GTRasterType (pixel is area / point), the ellipsoid keys, linear units;
//   * GDAL_NODATA pixels are replaced by the mean of the valid pixels (the
//     count is reported).
//
// Geographic rasters (degrees) become local east/north metres about the
// raster centre (lat_c, lon_c) on the ellipsoid:
//     x_i = N(lat_c) cos(lat_c) (lon_i - lon_c)      distance along the parallel
//     y_j = M(lat_j) - M(lat_c)                      meridian arc
// (N the prime-vertical radius, M the meridian arc), which keeps the axes
// separable with true distances along the centre row and column; the
// centre's UTM zone, easting and northing are reported for georeferencing.
// Projected rasters keep their metres, shifted to the raster centre.  Rows
// are reordered so that Y_vec ascends.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "bsar.h"
#include "bsar_files.h"

namespace bsar {

namespace {

constexpr double kDeg = M_PI / 180.0;

struct Tiff {
    const uint8_t* p = nullptr;
    size_t n = 0;
    bool be = false;

    bool has(uint64_t off, uint64_t len) const { return off <= n && len <= n - off; }
    uint16_t u16(size_t o) const { return be ? (uint16_t)(p[o] << 8 | p[o + 1]) : (uint16_t)(p[o] | p[o + 1] << 8); }
    uint32_t u32(size_t o) const
    {
        return be ? ((uint32_t)p[o] << 24 | (uint32_t)p[o + 1] << 16 | (uint32_t)p[o + 2] << 8 | p[o + 3])
                  : ((uint32_t)p[o] | (uint32_t)p[o + 1] << 8 | (uint32_t)p[o + 2] << 16 | (uint32_t)p[o + 3] << 24);
    }
    uint64_t u64(size_t o) const
    {
        const uint64_t a = u32(o), b = u32(o + 4);
        return be ? (a << 32 | b) : (b << 32 | a);
    }
    double f64(size_t o) const
    {
        const uint64_t u = u64(o);
        double d;
        std::memcpy(&d, &u, 8);
        return d;
    }
    float f32(size_t o) const
    {
        const uint32_t u = u32(o);
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    }
};

struct Entry {
    uint16_t tag = 0, type = 0;
    uint32_t count = 0;
    size_t off = 0;  // file offset of the value bytes
};

int type_size(int type)
{
    switch (type) {
    case 1: case 2: case 6: case 7: return 1;
    case 3: case 8: return 2;
    case 4: case 9: case 11: return 4;
    case 5: case 10: case 12: return 8;
    default: return 0;
    }
}

bool entry_values(const Tiff& t, const Entry& e, std::vector<double>& v)
{
    const int sz = type_size(e.type);
    if (!sz || !t.has(e.off, (uint64_t)sz * e.count)) return false;
    v.resize(e.count);
    for (uint32_t i = 0; i < e.count; ++i) {
        const size_t o = e.off + (size_t)i * sz;
        switch (e.type) {
        case 1: case 7: v[i] = t.p[o]; break;
        case 6: v[i] = (int8_t)t.p[o]; break;
        case 3: v[i] = t.u16(o); break;
        case 8: v[i] = (int16_t)t.u16(o); break;
        case 4: v[i] = t.u32(o); break;
        case 9: v[i] = (int32_t)t.u32(o); break;
        case 5: v[i] = t.u32(o + 4) ? (double)t.u32(o) / t.u32(o + 4) : 0.0; break;
        case 10: v[i] = (int32_t)t.u32(o + 4) ? (double)(int32_t)t.u32(o) / (int32_t)t.u32(o + 4) : 0.0; break;
        case 11: v[i] = t.f32(o); break;
        case 12: v[i] = t.f64(o); break;
        default: return false;
        }
    }
    return true;
}

// --- decompression -------------------------------------------------------

// TIFF LZW: 9..12-bit codes packed MSB first; 256 clears the table, 257 ends
// the stream; the code width grows when the next free code reaches
// 2^width - 1 (one code early).  Returns the bytes written (<= cap) or -1 on
// a corrupt stream.
int64_t lzw_decode(const uint8_t* in, size_t n, uint8_t* out, size_t cap)
{
    static thread_local uint16_t prefix[4096];
    static thread_local uint8_t suffix[4096], first[4096];
    static thread_local uint16_t len[4096];
    for (int i = 0; i < 256; ++i) {
        prefix[i] = 0xFFFF;
        suffix[i] = first[i] = (uint8_t)i;
        len[i] = 1;
    }
    const uint64_t nbits = (uint64_t)n * 8;
    uint64_t bit = 0;
    int width = 9, next = 258, old = -1;
    size_t o = 0;
    auto emit = [&](int code) {
        const size_t L = len[code];
        size_t pos = o + L;
        for (int c = code; c != 0xFFFF; c = prefix[c]) {
            --pos;
            if (pos < cap) out[pos] = suffix[c];
        }
        o += L;
    };
    while (o < cap) {
        if (bit + width > nbits) break;  // streams may end without the end code
        int code = 0;
        for (int k = 0; k < width; ++k, ++bit) code = code << 1 | ((in[bit >> 3] >> (7 - (bit & 7))) & 1);
        if (code == 257) break;
        if (code == 256) {
            width = 9;
            next = 258;
            old = -1;
            continue;
        }
        if (old < 0) {
            if (code > 255) return -1;
            emit(code);
            old = code;
            continue;
        }
        if (code > next) return -1;
        const uint8_t head = code < next ? first[code] : first[old];
        if (next < 4096) {
            prefix[next] = (uint16_t)old;
            suffix[next] = head;
            first[next] = first[old];
            len[next] = (uint16_t)(len[old] + 1);
            ++next;
        } else if (code == next) {
            return -1;
        }
        emit(code);
        old = code;
        if (next >= (1 << width) - 1 && width < 12) ++width;
    }
    return (int64_t)std::min(o, cap);
}

int64_t packbits_decode(const uint8_t* in, size_t n, uint8_t* out, size_t cap)
{
    size_t i = 0, o = 0;
    while (i < n && o < cap) {
        const int h = (int8_t)in[i++];
        if (h >= 0) {
            for (int k = 0; k <= h && i < n && o < cap; ++k) out[o++] = in[i++];
        } else if (h != -128) {
            if (i >= n) break;
            const uint8_t b = in[i++];
            for (int k = 0; k < 1 - h && o < cap; ++k) out[o++] = b;
        }
    }
    return (int64_t)o;
}

int64_t inflate_decode(const uint8_t* in, size_t n, uint8_t* out, size_t cap)
{
    z_stream z{};
    if (inflateInit(&z) != Z_OK) return -1;
    z.next_in = const_cast<uint8_t*>(in);
    z.avail_in = (uInt)n;
    z.next_out = out;
    z.avail_out = (uInt)cap;
    const int r = inflate(&z, Z_FINISH);
    const int64_t got = (int64_t)(cap - z.avail_out);
    inflateEnd(&z);
    return (r == Z_STREAM_END || r == Z_OK || r == Z_BUF_ERROR) ? got : -1;
}

// --- predictors (one row of w samples of `bytes` bytes) -------------------

template <typename U>
void undiff_row(uint8_t* row, int w)
{
    U prev;
    std::memcpy(&prev, row, sizeof(U));
    for (int i = 1; i < w; ++i) {
        U v;
        std::memcpy(&v, row + (size_t)i * sizeof(U), sizeof(U));
        v = (U)(v + prev);
        std::memcpy(row + (size_t)i * sizeof(U), &v, sizeof(U));
        prev = v;
    }
}

void swap_bytes(uint8_t* p, size_t count, int bytes)
{
    if (bytes == 1) return;
    for (size_t i = 0; i < count; ++i) std::reverse(p + i * bytes, p + (i + 1) * bytes);
}

// Floating-point predictor: byte-wise differencing over the row, the bytes
// of each sample stored as planes, most significant first.  Leaves native
// (little-endian) samples.
void fp_predictor_row(uint8_t* row, int w, int bytes, std::vector<uint8_t>& tmp)
{
    const size_t nb = (size_t)w * bytes;
    for (size_t i = 1; i < nb; ++i) row[i] = (uint8_t)(row[i] + row[i - 1]);
    tmp.assign(row, row + nb);
    for (int i = 0; i < w; ++i)
        for (int b = 0; b < bytes; ++b) row[(size_t)i * bytes + b] = tmp[(size_t)(bytes - 1 - b) * w + i];
}

// --- the projection -------------------------------------------------------

struct Kruger {
    double A, alpha[5], e;
};

Kruger kruger(double a, double f)
{
    const double n = f / (2 - f), n2 = n * n, n3 = n2 * n, n4 = n3 * n;
    Kruger k;
    k.A = a / (1 + n) * (1 + n2 / 4 + n4 / 64);
    k.alpha[1] = n / 2 - 2 * n2 / 3 + 5 * n3 / 16 + 41 * n4 / 180;
    k.alpha[2] = 13 * n2 / 48 - 3 * n3 / 5 + 557 * n4 / 1440;
    k.alpha[3] = 61 * n3 / 240 - 103 * n4 / 140;
    k.alpha[4] = 49561 * n4 / 161280;
    k.e = std::sqrt(f * (2 - f));
    return k;
}

}  // namespace

void utm_forward(double a, double f, double lat, double lon, double lon0, double* easting, double* northing)
{
    const Kruger k = kruger(a, f);
    const double phi = lat * kDeg, dl = (lon - lon0) * kDeg;
    const double s = std::sin(phi);
    // conformal latitude through t = tan(chi)
    const double t = std::sinh(std::atanh(s) - k.e * std::atanh(k.e * s));
    const double xi = std::atan2(t, std::cos(dl));
    const double eta = std::atanh(std::sin(dl) / std::sqrt(1 + t * t));
    double X = xi, Y = eta;
    for (int j = 1; j <= 4; ++j) {
        X += k.alpha[j] * std::sin(2 * j * xi) * std::cosh(2 * j * eta);
        Y += k.alpha[j] * std::cos(2 * j * xi) * std::sinh(2 * j * eta);
    }
    const double k0 = 0.9996;
    *easting = 500000.0 + k0 * k.A * Y;
    *northing = (lat < 0 ? 10000000.0 : 0.0) + k0 * k.A * X;
}

double meridian_arc(double a, double f, double lat)
{
    const Kruger k = kruger(a, f);
    const double s = std::sin(lat * kDeg);
    const double chi = std::atan(std::sinh(std::atanh(s) - k.e * std::atanh(k.e * s)));
    double X = chi;
    for (int j = 1; j <= 4; ++j) X += k.alpha[j] * std::sin(2 * j * chi);
    return k.A * X;
}

static int geotiff_load_impl(const char* path, bsar_dem_grid* out);

int geotiff_load(const char* path, bsar_dem_grid* out)
{
    try {
        return geotiff_load_impl(path, out);
    } catch (const std::bad_alloc&) {
        return file_fail(BSAR_ERR_NOMEM, std::string("GeoTIFF ") + path + ": host allocation failed");
    } catch (const std::exception& e) {
        return file_fail(BSAR_ERR_FORMAT, std::string("GeoTIFF ") + path + ": " + e.what());
    }
}

static int geotiff_load_impl(const char* path, bsar_dem_grid* out)
{
    const std::string where = std::string("GeoTIFF ") + path + ": ";
    FILE* fp = std::fopen(path, "rb");
    if (!fp) return file_fail(BSAR_ERR_IO, where + "cannot open");
    std::vector<uint8_t> buf;
    {
        uint8_t chunk[1 << 16];
        size_t r;
        while ((r = std::fread(chunk, 1, sizeof chunk, fp)) > 0) buf.insert(buf.end(), chunk, chunk + r);
        std::fclose(fp);
    }
    Tiff t;
    t.p = buf.data();
    t.n = buf.size();
    if (t.n < 8) return file_fail(BSAR_ERR_FORMAT, where + "too short");
    if (buf[0] == 'I' && buf[1] == 'I') t.be = false;
    else if (buf[0] == 'M' && buf[1] == 'M') t.be = true;
    else return file_fail(BSAR_ERR_FORMAT, where + "not a TIFF file");
    const int magic = t.u16(2);
    if (magic == 43) return file_fail(BSAR_ERR_FORMAT, where + "BigTIFF is not supported");
    if (magic != 42) return file_fail(BSAR_ERR_FORMAT, where + "bad TIFF magic");
    const uint32_t ifd = t.u32(4);
    if (!t.has(ifd, 2)) return file_fail(BSAR_ERR_FORMAT, where + "IFD offset out of range");
    const int nent = t.u16(ifd);
    if (!t.has(ifd + 2, (uint64_t)nent * 12)) return file_fail(BSAR_ERR_FORMAT, where + "IFD truncated");
    std::vector<Entry> ents(nent);
    for (int i = 0; i < nent; ++i) {
        const size_t o = ifd + 2 + (size_t)i * 12;
        Entry& e = ents[i];
        e.tag = t.u16(o);
        e.type = t.u16(o + 2);
        e.count = t.u32(o + 4);
        const uint64_t bytes = (uint64_t)type_size(e.type) * e.count;
        e.off = bytes <= 4 ? o + 8 : t.u32(o + 8);
    }
    auto find = [&](int tag) -> const Entry* {
        for (const Entry& e : ents)
            if (e.tag == tag) return &e;
        return nullptr;
    };
    auto values = [&](int tag, std::vector<double>& v) -> bool {
        const Entry* e = find(tag);
        return e && entry_values(t, *e, v);
    };
    auto scalar = [&](int tag, double dflt) -> double {
        std::vector<double> v;
        return values(tag, v) && !v.empty() ? v[0] : dflt;
    };

    // integral tag values in [lo, hi] as int64, else -1 (no out-of-range
    // double -> integer conversion on crafted tags)
    auto count_in = [](double v, double lo, double hi) -> int64_t {
        return (v >= lo && v <= hi && v == std::floor(v)) ? (int64_t)v : -1;
    };
    const int64_t W = count_in(scalar(256, 0), 1, (double)(1 << 28));
    const int64_t H = count_in(scalar(257, 0), 1, (double)(1 << 28));
    const int bps = (int)scalar(258, 1);
    const int comp = (int)scalar(259, 1);
    const int spp = (int)scalar(277, 1);
    const int pred = (int)scalar(317, 1);
    const int fmt = (int)scalar(339, 1);
    if (W <= 0 || H <= 0 || W * H > ((int64_t)1 << 28)) return file_fail(BSAR_ERR_FORMAT, where + "bad image size");
    if (spp != 1) return file_fail(BSAR_ERR_FORMAT, where + "only single-band rasters are supported");
    if (!(bps == 8 || bps == 16 || bps == 32 || bps == 64)) return file_fail(BSAR_ERR_FORMAT, where + "bits per sample");
    if (fmt == 3 && !(bps == 32 || bps == 64)) return file_fail(BSAR_ERR_FORMAT, where + "float sample size");
    if (fmt < 1 || fmt > 3) return file_fail(BSAR_ERR_FORMAT, where + "sample format");
    if (!(comp == 1 || comp == 5 || comp == 8 || comp == 32946 || comp == 32773))
        return file_fail(BSAR_ERR_FORMAT, where + "compression " + std::to_string(comp) + " is not supported");
    if (!(pred == 1 || pred == 2 || pred == 3) || (pred == 3 && fmt != 3))
        return file_fail(BSAR_ERR_FORMAT, where + "predictor " + std::to_string(pred));
    const int bytes = bps / 8;
    // the predictor applies to LZW and Deflate data only (libtiff and GDAL
    // ignore the tag on uncompressed and PackBits data)
    const int pred_used = (comp == 5 || comp == 8 || comp == 32946) ? pred : 1;

    // chunks: tiles or strips
    std::vector<double> offs, cnts;
    int64_t cw, ch;
    const bool tiled = find(322) != nullptr;
    if (tiled) {
        cw = count_in(scalar(322, 0), 1, (double)(1 << 20));
        ch = count_in(scalar(323, 0), 1, (double)(1 << 20));
        if (!values(324, offs) || !values(325, cnts)) return file_fail(BSAR_ERR_FORMAT, where + "tile offsets");
    } else {
        cw = W;
        const double rps = scalar(278, (double)H);      // RowsPerStrip, often 2^32 - 1 for "one strip"
        ch = rps >= (double)H ? H : count_in(rps, 1, (double)H);
        if (!values(273, offs) || !values(279, cnts)) return file_fail(BSAR_ERR_FORMAT, where + "strip offsets");
    }
    // a chunk holds at most 2^28 samples (the image bound), so its bytes fit
    // in 2^31 and every row offset below stays inside `raw`
    if (cw <= 0 || ch <= 0 || cw * ch > ((int64_t)1 << 28))
        return file_fail(BSAR_ERR_FORMAT, where + "bad tile/strip size");
    const int64_t across = (W + cw - 1) / cw, down = (H + ch - 1) / ch;
    if ((int64_t)offs.size() < across * down || cnts.size() < offs.size())
        return file_fail(BSAR_ERR_FORMAT, where + "missing tiles/strips");

    std::vector<double> img((size_t)(W * H));
    std::vector<uint8_t> raw, tmp;
    for (int64_t cy = 0; cy < down; ++cy)
        for (int64_t cx = 0; cx < across; ++cx) {
            const size_t ci = (size_t)(cy * across + cx);
            if (!(offs[ci] >= 0 && offs[ci] < 9e15 && cnts[ci] >= 0 && cnts[ci] < 9e15))
                return file_fail(BSAR_ERR_FORMAT, where + "chunk outside the file");
            const uint64_t off = (uint64_t)offs[ci], cnt = (uint64_t)cnts[ci];
            if (!t.has(off, cnt)) return file_fail(BSAR_ERR_FORMAT, where + "chunk outside the file");
            const int64_t rows = tiled ? ch : std::min(ch, H - cy * ch);
            const size_t want = (size_t)((uint64_t)cw * (uint64_t)rows * (uint64_t)bytes);
            raw.assign(want, 0);
            int64_t got;
            switch (comp) {
            case 1: got = (int64_t)std::min<uint64_t>(cnt, want); std::memcpy(raw.data(), t.p + off, (size_t)got); break;
            case 5: got = lzw_decode(t.p + off, cnt, raw.data(), want); break;
            case 8: case 32946: got = inflate_decode(t.p + off, cnt, raw.data(), want); break;
            default: got = packbits_decode(t.p + off, cnt, raw.data(), want); break;
            }
            if (got < 0) return file_fail(BSAR_ERR_FORMAT, where + "corrupt compressed data");
            for (int64_t r = 0; r < rows; ++r) {
                uint8_t* row = raw.data() + (size_t)(r * cw * bytes);
                if (pred_used == 3) {
                    fp_predictor_row(row, (int)cw, bytes, tmp);
                } else {
                    if (t.be) swap_bytes(row, (size_t)cw, bytes);
                    if (pred_used == 2) {
                        if (bytes == 1) undiff_row<uint8_t>(row, (int)cw);
                        else if (bytes == 2) undiff_row<uint16_t>(row, (int)cw);
                        else if (bytes == 4) undiff_row<uint32_t>(row, (int)cw);
                        else undiff_row<uint64_t>(row, (int)cw);
                    }
                }
                const int64_t y = cy * ch + r;
                if (y >= H) break;
                for (int64_t c = 0; c < cw; ++c) {
                    const int64_t x = cx * cw + c;
                    if (x >= W) break;
                    const uint8_t* s = row + c * bytes;
                    double v;
                    if (fmt == 3) {
                        if (bytes == 4) { float f; std::memcpy(&f, s, 4); v = f; }
                        else std::memcpy(&v, s, 8);
                    } else if (fmt == 2) {
                        if (bytes == 1) v = (int8_t)s[0];
                        else if (bytes == 2) { int16_t q; std::memcpy(&q, s, 2); v = q; }
                        else if (bytes == 4) { int32_t q; std::memcpy(&q, s, 4); v = q; }
                        else { int64_t q; std::memcpy(&q, s, 8); v = (double)q; }
                    } else {
                        if (bytes == 1) v = s[0];
                        else if (bytes == 2) { uint16_t q; std::memcpy(&q, s, 2); v = q; }
                        else if (bytes == 4) { uint32_t q; std::memcpy(&q, s, 4); v = q; }
                        else { uint64_t q; std::memcpy(&q, s, 8); v = (double)q; }
                    }
                    img[(size_t)(y * W + x)] = v;
                }
            }
        }

    // nodata
    int nodata_count = 0;
    if (const Entry* e = find(42113)) {
        if (t.has(e->off, e->count)) {
            const std::string s((const char*)t.p + e->off, strnlen((const char*)t.p + e->off, e->count));
            char* end = nullptr;
            const double nd = std::strtod(s.c_str(), &end);
            if (end != s.c_str()) {
                double sum = 0.0;
                int64_t valid = 0;
                for (double v : img)
                    if (v != nd && std::isfinite(v)) { sum += v; ++valid; }
                if (!valid) return file_fail(BSAR_ERR_FORMAT, where + "every pixel is nodata");
                const double fill = sum / (double)valid;
                for (double& v : img)
                    if (v == nd || !std::isfinite(v)) { v = fill; ++nodata_count; }
            }
        }
    }

    // georeferencing: pixel (i, j) -> model (X, Y) = (X0 + i sx, Y0 - j sy)
    std::vector<double> scale, tie, xf;
    double X0, Y0, sx, sy;
    if (values(33550, scale) && scale.size() >= 2 && values(33922, tie) && tie.size() >= 6) {
        sx = scale[0];
        sy = scale[1];
        X0 = tie[3] - tie[0] * sx;
        Y0 = tie[4] + tie[1] * sy;
    } else if (values(34264, xf) && xf.size() >= 16 && xf[1] == 0.0 && xf[4] == 0.0) {
        sx = xf[0];
        sy = -xf[5];
        X0 = xf[3];
        Y0 = xf[7];
    } else {
        return file_fail(BSAR_ERR_FORMAT, where + "no georeferencing (ModelPixelScale + ModelTiepoint)");
    }
    if (!(sx != 0.0 && sy != 0.0 && std::isfinite(sx) && std::isfinite(sy)))
        return file_fail(BSAR_ERR_FORMAT, where + "bad pixel scale");

    // GeoKeys
    int model = 0, raster = 1, units = 9001;
    double a = 6378137.0, f = 1.0 / 298.257223563;
    std::vector<double> keys, dparams;
    values(34736, dparams);
    if (values(34735, keys) && keys.size() >= 4) {
        const size_t nk = (size_t)keys[3];
        for (size_t k = 0; k < nk && 4 + 4 * k + 3 < keys.size(); ++k) {
            const int id = (int)keys[4 + 4 * k], loc = (int)keys[4 + 4 * k + 1];
            const double val = keys[4 + 4 * k + 3];
            auto dval = [&](double& dst) {
                if (loc == 34736 && (size_t)val < dparams.size()) dst = dparams[(size_t)val];
            };
            if (id == 1024 && loc == 0) model = (int)val;
            else if (id == 1025 && loc == 0) raster = (int)val;
            else if (id == 3076 && loc == 0) units = (int)val;
            else if (id == 2057) dval(a);
            else if (id == 2059) { double inv = 0.0; dval(inv); if (inv > 0) f = 1.0 / inv; }
        }
    }
    if (model != 1 && model != 2) return file_fail(BSAR_ERR_FORMAT, where + "GTModelType must be projected or geographic");
    const double half = raster == 2 ? 0.0 : 0.5;  // pixel is point: the tie point is the sample

    std::vector<double> mx(W), my(H);
    for (int64_t i = 0; i < W; ++i) mx[i] = X0 + ((double)i + half) * sx;
    for (int64_t j = 0; j < H; ++j) my[j] = Y0 - ((double)j + half) * sy;
    const double xc = 0.5 * (mx.front() + mx.back()), yc = 0.5 * (my.front() + my.back());
    std::vector<double> ax(W), ay(H);
    if (model == 2) {
        if (std::fabs(yc) > 89.0) return file_fail(BSAR_ERR_FORMAT, where + "raster too close to a pole");
        const double s = std::sin(yc * kDeg), e2 = f * (2 - f);
        const double Nc = a / std::sqrt(1 - e2 * s * s);
        const double kx = Nc * std::cos(yc * kDeg) * kDeg;
        const double Mc = meridian_arc(a, f, yc);
        for (int64_t i = 0; i < W; ++i) ax[i] = kx * (mx[i] - xc);
        for (int64_t j = 0; j < H; ++j) ay[j] = meridian_arc(a, f, my[j]) - Mc;
        int zone = (int)std::floor((xc + 180.0) / 6.0) + 1;
        zone = std::min(60, std::max(1, zone));
        const double lon0 = zone * 6.0 - 183.0;
        utm_forward(a, f, yc, xc, lon0, &out->e0, &out->n0);
        out->utm_zone = yc < 0 ? -zone : zone;
        out->lon0 = xc;
        out->lat0 = yc;
    } else {
        const double u = units == 9002 ? 0.3048 : units == 9003 ? 1200.0 / 3937.0 : 1.0;
        for (int64_t i = 0; i < W; ++i) ax[i] = (mx[i] - xc) * u;
        for (int64_t j = 0; j < H; ++j) ay[j] = (my[j] - yc) * u;
        out->e0 = xc * u;
        out->n0 = yc * u;
    }
    // ascending axes
    const bool flip_x = W > 1 && ax[1] < ax[0];
    const bool flip_y = H > 1 && ay[1] < ay[0];
    if (flip_x) std::reverse(ax.begin(), ax.end());
    if (flip_y) std::reverse(ay.begin(), ay.end());
    std::vector<double> z((size_t)(W * H));
    for (int64_t j = 0; j < H; ++j)
        for (int64_t i = 0; i < W; ++i)
            z[(size_t)(j * W + i)] = img[(size_t)((flip_y ? H - 1 - j : j) * W + (flip_x ? W - 1 - i : i))];
    out->nodata_count = nodata_count;
    out->source = BSAR_DEM_SOURCE_GEOTIFF;
    return dem_grid_fill(out, ax, ay, z, (int)H, (int)W);
}

}  // namespace bsar

extern "C" {

BSAR_API int bsar_geo_to_utm(double lat, double lon, int zone, double* easting, double* northing, int* zone_out)
{
    if (!easting || !northing || !(std::fabs(lat) <= 90.0) || !std::isfinite(lon) || zone < 0 || zone > 60)
        return bsar::file_fail(BSAR_ERR_INVALID, "bsar_geo_to_utm: bad latitude, longitude or zone");
    if (zone == 0) zone = std::min(60, std::max(1, (int)std::floor((lon + 180.0) / 6.0) + 1));
    bsar::utm_forward(6378137.0, 1.0 / 298.257223563, lat, lon, zone * 6.0 - 183.0, easting, northing);
    if (zone_out) *zone_out = lat < 0 ? -zone : zone;
    return BSAR_OK;
}

BSAR_API double bsar_meridian_arc(double lat)
{
    return bsar::meridian_arc(6378137.0, 1.0 / 298.257223563, lat);
}

}  // extern "C"
