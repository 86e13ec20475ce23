// File formats of the 3DMatch toolbox (host side): binary PLY, keypoint and
// descriptor .bin files, tsdf.bin/tdf.bin, patch volumes, feat.bin, text
// matrices, 16-bit PNG depth, .marvin tensors; plus the fp32 4x4 pose
// arithmetic the fusion demo runs on the host.
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <initializer_list>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"
#include "pipeline.hpp"
#include "marvin_io.hpp"

#pragma clang fp contract(off)

namespace {

struct File {
    FILE *f = nullptr;
    File(const char *path, const char *mode) : f(std::fopen(path, mode)) {}
    ~File() {
        if (f) std::fclose(f);
    }
    File(const File &) = delete;
    File &operator=(const File &) = delete;
};

int write_all(FILE *f, const void *p, size_t bytes, const char *path) {
    if (bytes && std::fwrite(p, 1, bytes, f) != bytes)
        return tdm::fail(TDM_ERR_IO, "short write to %s", path);
    return TDM_OK;
}

int read_all(FILE *f, void *p, size_t bytes, const char *path) {
    if (bytes && std::fread(p, 1, bytes, f) != bytes)
        return tdm::fail(TDM_ERR_FORMAT, "%s is truncated", path);
    return TDM_OK;
}

// Float count header as the reference writes it ((float)N).
int read_count(FILE *f, const char *path, int *n) {
    float v = 0;
    TDM_TRY(read_all(f, &v, sizeof v, path));
    TDM_CHECK(std::isfinite(v) && v >= 0 && v <= 2.0e9f && v == std::floor(v), TDM_ERR_FORMAT,
              "%s: bad count header %g", path, (double)v);
    *n = (int)v;
    return TDM_OK;
}

// Bytes from the current position of f to its end; -1 when the stream
// cannot seek (then the read itself reports a short file).
long long bytes_left(FILE *f) {
    const long pos = std::ftell(f);
    if (pos < 0 || std::fseek(f, 0, SEEK_END) != 0) return -1;
    const long end = std::ftell(f);
    if (std::fseek(f, pos, SEEK_SET) != 0 || end < pos) return -1;
    return (long long)(end - pos);
}

// The element count a header declares (the product of `factors`), checked
// for overflow and against what the file still holds before anything is
// allocated: a hostile header must not turn into a huge allocation.
int payload_count(long long left, const char *path, std::initializer_list<size_t> factors,
                  size_t elem, size_t *count) {
    size_t c = 1;
    for (size_t v : factors) {
        TDM_CHECK(v == 0 || c <= SIZE_MAX / elem / v, TDM_ERR_FORMAT,
                  "%s: header sizes overflow", path);
        c *= v;
    }
    TDM_CHECK(left < 0 || (unsigned long long)left >= c * elem, TDM_ERR_FORMAT,
              "%s is truncated: the header declares %zu bytes, %lld remain", path, c * elem, left);
    *count = c;
    return TDM_OK;
}

template <typename T>
int alloc_copy(size_t count, T **out) {
    *out = static_cast<T *>(std::malloc(count ? count * sizeof(T) : 1));
    TDM_CHECK(*out, TDM_ERR_NOMEM, "host allocation of %zu elements failed", count);
    return TDM_OK;
}

// ---- PNG (16-bit grayscale) --------------------------------------------------

uint32_t be32(const unsigned char *p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3];
}

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

int decode_png_u16(const char *path, std::vector<uint16_t> &pix, int &H, int &W) {
    std::ifstream in(path, std::ios::binary);
    TDM_CHECK(in, TDM_ERR_IO, "cannot open %s", path);
    std::vector<unsigned char> buf((std::istreambuf_iterator<char>(in)),
                                   std::istreambuf_iterator<char>());
    static const unsigned char sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    TDM_CHECK(buf.size() >= 8 && std::memcmp(buf.data(), sig, 8) == 0, TDM_ERR_FORMAT,
              "%s is not a PNG file", path);
    size_t pos = 8;
    int depth = 0, ctype = -1, interlace = 0;
    W = H = 0;
    std::vector<unsigned char> idat;
    bool seen_end = false;
    while (pos + 12 <= buf.size()) {
        uint32_t len = be32(&buf[pos]);
        TDM_CHECK(pos + 12 + (size_t)len <= buf.size(), TDM_ERR_FORMAT, "%s: truncated chunk", path);
        const unsigned char *type = &buf[pos + 4];
        const unsigned char *data = &buf[pos + 8];
        uint32_t crc = be32(&buf[pos + 8 + len]);
        TDM_CHECK(crc32(crc32(0L, Z_NULL, 0), type, len + 4) == crc, TDM_ERR_FORMAT,
                  "%s: chunk CRC mismatch", path);
        if (!std::memcmp(type, "IHDR", 4)) {
            TDM_CHECK(len == 13, TDM_ERR_FORMAT, "%s: bad IHDR", path);
            W = (int)be32(data);
            H = (int)be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            seen_end = true;
            break;
        }
        pos += 12 + (size_t)len;
    }
    TDM_CHECK(seen_end && W > 0 && H > 0, TDM_ERR_FORMAT, "%s: incomplete PNG", path);
    TDM_CHECK(depth == 16 && ctype == 0, TDM_ERR_UNSUPPORTED,
              "%s: only 16-bit grayscale depth PNGs are supported (bit depth %d, colour type %d)",
              path, depth, ctype);
    TDM_CHECK(interlace == 0, TDM_ERR_UNSUPPORTED, "%s: interlaced PNG", path);
    TDM_CHECK((long)W * H <= (1L << 28), TDM_ERR_FORMAT, "%s: image too large", path);
    const size_t stride = (size_t)W * 2;
    // deflate expands at most 1032:1, so a short IDAT cannot hold a large image
    TDM_CHECK((stride + 1) * H <= idat.size() * 1032 + 1024, TDM_ERR_FORMAT,
              "%s: image data too short for %d x %d", path, W, H);
    std::vector<unsigned char> raw((stride + 1) * H);
    z_stream zs;
    std::memset(&zs, 0, sizeof zs);
    TDM_CHECK(inflateInit(&zs) == Z_OK, TDM_ERR_FORMAT, "zlib init failed");
    zs.next_in = idat.data();
    zs.avail_in = (uInt)idat.size();
    zs.next_out = raw.data();
    zs.avail_out = (uInt)raw.size();
    int zr = inflate(&zs, Z_FINISH);
    size_t got = raw.size() - zs.avail_out;
    inflateEnd(&zs);
    TDM_CHECK((zr == Z_STREAM_END || zr == Z_BUF_ERROR) && got == raw.size(), TDM_ERR_FORMAT,
              "%s: bad image data", path);
    // Undo the per-row filters (bytes per pixel = 2).
    std::vector<unsigned char> prev(stride, 0), cur(stride);
    pix.resize((size_t)W * H);
    for (int r = 0; r < H; ++r) {
        const unsigned char *row = &raw[r * (stride + 1)];
        int ft = row[0];
        for (size_t i = 0; i < stride; ++i) {
            int a = i >= 2 ? cur[i - 2] : 0, b = prev[i], c = i >= 2 ? prev[i - 2] : 0;
            int x = row[1 + i], v;
            switch (ft) {
            case 0: v = x; break;
            case 1: v = x + a; break;
            case 2: v = x + b; break;
            case 3: v = x + ((a + b) >> 1); break;
            case 4: v = x + paeth(a, b, c); break;
            default: return tdm::fail(TDM_ERR_FORMAT, "%s: bad filter %d", path, ft);
            }
            cur[i] = (unsigned char)v;
        }
        for (int c = 0; c < W; ++c) pix[(size_t)r * W + c] = (uint16_t)(cur[2 * c] << 8 | cur[2 * c + 1]);
        prev.swap(cur);
    }
    return TDM_OK;
}

// Cofactor terms of the 4x4 inverse: entry e = s_e * (+t0 - t1 - t2 + t3 + t4 - t5)
// with t = m[a]*m[b]*m[c] ((a*b)*c), accumulated left to right, the sign
// pattern of depth-fusion/utils.hpp:112-238.
struct CofTerms {
    signed char first_sign;
    unsigned char idx[6][3];
};
const CofTerms kCof[16] = {
    {+1, {{5, 10, 15}, {5, 11, 14}, {9, 6, 15}, {9, 7, 14}, {13, 6, 11}, {13, 7, 10}}},   // 0
    {-1, {{1, 10, 15}, {1, 11, 14}, {9, 2, 15}, {9, 3, 14}, {13, 2, 11}, {13, 3, 10}}},   // 1
    {+1, {{1, 6, 15}, {1, 7, 14}, {5, 2, 15}, {5, 3, 14}, {13, 2, 7}, {13, 3, 6}}},       // 2
    {-1, {{1, 6, 11}, {1, 7, 10}, {5, 2, 11}, {5, 3, 10}, {9, 2, 7}, {9, 3, 6}}},         // 3
    {-1, {{4, 10, 15}, {4, 11, 14}, {8, 6, 15}, {8, 7, 14}, {12, 6, 11}, {12, 7, 10}}},   // 4
    {+1, {{0, 10, 15}, {0, 11, 14}, {8, 2, 15}, {8, 3, 14}, {12, 2, 11}, {12, 3, 10}}},   // 5
    {-1, {{0, 6, 15}, {0, 7, 14}, {4, 2, 15}, {4, 3, 14}, {12, 2, 7}, {12, 3, 6}}},       // 6
    {+1, {{0, 6, 11}, {0, 7, 10}, {4, 2, 11}, {4, 3, 10}, {8, 2, 7}, {8, 3, 6}}},         // 7
    {+1, {{4, 9, 15}, {4, 11, 13}, {8, 5, 15}, {8, 7, 13}, {12, 5, 11}, {12, 7, 9}}},     // 8
    {-1, {{0, 9, 15}, {0, 11, 13}, {8, 1, 15}, {8, 3, 13}, {12, 1, 11}, {12, 3, 9}}},     // 9
    {+1, {{0, 5, 15}, {0, 7, 13}, {4, 1, 15}, {4, 3, 13}, {12, 1, 7}, {12, 3, 5}}},       // 10
    {-1, {{0, 5, 11}, {0, 7, 9}, {4, 1, 11}, {4, 3, 9}, {8, 1, 7}, {8, 3, 5}}},           // 11
    {-1, {{4, 9, 14}, {4, 10, 13}, {8, 5, 14}, {8, 6, 13}, {12, 5, 10}, {12, 6, 9}}},     // 12
    {+1, {{0, 9, 14}, {0, 10, 13}, {8, 1, 14}, {8, 2, 13}, {12, 1, 10}, {12, 2, 9}}},     // 13
    {-1, {{0, 5, 14}, {0, 6, 13}, {4, 1, 14}, {4, 2, 13}, {12, 1, 6}, {12, 2, 5}}},       // 14
    {+1, {{0, 5, 10}, {0, 6, 9}, {4, 1, 10}, {4, 2, 9}, {8, 1, 6}, {8, 2, 5}}},           // 15
};
const signed char kCofPattern[6] = {+1, -1, -1, +1, +1, -1};

}  // namespace

namespace tdm {
int read_png_u16(const char *path, std::vector<uint16_t> &pix, int &H, int &W) {
    return decode_png_u16(path, pix, H, W);
}
}  // namespace tdm

extern "C" {

int tdm_read_ply_points(const char *path, float **pts, int *num) {
    tdm::clear_error();
    TDM_CHECK(path && pts && num, TDM_ERR_INVALID_ARG, "null pointer argument");
    std::ifstream in(path, std::ios::binary);
    TDM_CHECK(in, TDM_ERR_IO, "point cloud file %s not found", path);
    // core/demo.cu:77-88: 7 header lines; the third token of line 3 is N.
    long n = 0;
    for (int line = 0; line < 7; ++line) {
        std::string s;
        TDM_CHECK(std::getline(in, s), TDM_ERR_FORMAT, "%s: header shorter than 7 lines", path);
        if (line == 2) {
            std::istringstream t(s);
            std::string a, b;
            t >> a >> b >> n;
        }
    }
    TDM_CHECK(n > 0 && n < (1L << 30), TDM_ERR_FORMAT,
              "%s: third header line does not give a point count", path);
    long long left = -1;
    {
        const std::streamoff here = in.tellg();
        if (here >= 0 && in.seekg(0, std::ios::end)) {
            const std::streamoff end = in.tellg();
            if (end >= here) left = (long long)(end - here);
        }
        in.clear();
        in.seekg(here);
        TDM_CHECK(in, TDM_ERR_IO, "%s: cannot seek", path);
    }
    size_t count = 0;
    TDM_TRY(payload_count(left, path, {(size_t)n, 3}, sizeof(float), &count));
    float *p = nullptr;
    TDM_TRY(alloc_copy(count, &p));
    in.read(reinterpret_cast<char *>(p), (std::streamsize)(sizeof(float) * 3 * n));
    if (in.gcount() != (std::streamsize)(sizeof(float) * 3 * n)) {
        std::free(p);
        return tdm::fail(TDM_ERR_FORMAT, "%s: expected %ld points", path, n);
    }
    *pts = p;
    *num = (int)n;
    return TDM_OK;
}

int tdm_write_ply_points(const char *path, const float *pts, int num) {
    tdm::clear_error();
    TDM_CHECK(path && num >= 0 && (num == 0 || pts), TDM_ERR_INVALID_ARG, "bad arguments");
    File f(path, "wb");
    TDM_CHECK(f.f, TDM_ERR_IO, "cannot create %s", path);
    std::fprintf(f.f,
                 "ply\nformat binary_little_endian 1.0\nelement vertex %d\nproperty float x\n"
                 "property float y\nproperty float z\nend_header\n",
                 num);
    return write_all(f.f, pts, sizeof(float) * 3 * (size_t)num, path);
}

int tdm_write_keypts_bin(const char *path, const float *kp, int num) {
    tdm::clear_error();
    TDM_CHECK(path && num >= 0 && (num == 0 || kp), TDM_ERR_INVALID_ARG, "bad arguments");
    File f(path, "wb");
    TDM_CHECK(f.f, TDM_ERR_IO, "cannot create %s", path);
    float n = (float)num;
    TDM_TRY(write_all(f.f, &n, sizeof n, path));
    return write_all(f.f, kp, sizeof(float) * 3 * (size_t)num, path);
}

int tdm_read_keypts_bin(const char *path, float **kp, int *num) {
    tdm::clear_error();
    TDM_CHECK(path && kp && num, TDM_ERR_INVALID_ARG, "null pointer argument");
    File f(path, "rb");
    TDM_CHECK(f.f, TDM_ERR_IO, "cannot open %s", path);
    int n = 0;
    TDM_TRY(read_count(f.f, path, &n));
    size_t count = 0;
    TDM_TRY(payload_count(bytes_left(f.f), path, {(size_t)n, 3}, sizeof(float), &count));
    float *p = nullptr;
    TDM_TRY(alloc_copy(count, &p));
    int s = read_all(f.f, p, sizeof(float) * 3 * (size_t)n, path);
    if (s != TDM_OK) {
        std::free(p);
        return s;
    }
    *kp = p;
    *num = n;
    return TDM_OK;
}

int tdm_write_desc_bin(const char *path, const float *desc, int num, int dim) {
    tdm::clear_error();
    TDM_CHECK(path && num >= 0 && dim > 0 && (num == 0 || desc), TDM_ERR_INVALID_ARG,
              "bad arguments");
    File f(path, "wb");
    TDM_CHECK(f.f, TDM_ERR_IO, "cannot create %s", path);
    float h[2] = {(float)num, (float)dim};
    TDM_TRY(write_all(f.f, h, sizeof h, path));
    return write_all(f.f, desc, sizeof(float) * (size_t)num * dim, path);
}

int tdm_read_desc_bin(const char *path, float **desc, int *num, int *dim) {
    tdm::clear_error();
    TDM_CHECK(path && desc && num && dim, TDM_ERR_INVALID_ARG, "null pointer argument");
    File f(path, "rb");
    TDM_CHECK(f.f, TDM_ERR_IO, "cannot open %s", path);
    int n = 0, d = 0;
    TDM_TRY(read_count(f.f, path, &n));
    TDM_TRY(read_count(f.f, path, &d));
    size_t count = 0;
    TDM_TRY(payload_count(bytes_left(f.f), path, {(size_t)n, (size_t)d}, sizeof(float), &count));
    float *p = nullptr;
    TDM_TRY(alloc_copy(count, &p));
    int s = read_all(f.f, p, sizeof(float) * (size_t)n * d, path);
    if (s != TDM_OK) {
        std::free(p);
        return s;
    }
    *desc = p;
    *num = n;
    *dim = d;
    return TDM_OK;
}

int tdm_write_grid_bin(const char *path, const int dims[3], const float origin[3],
                       float voxel_size, float trunc_margin, const float *grid) {
    tdm::clear_error();
    TDM_CHECK(path && dims && origin && grid, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_CHECK(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, TDM_ERR_INVALID_ARG, "bad dims");
    File f(path, "wb");
    TDM_CHECK(f.f, TDM_ERR_IO, "cannot create %s", path);
    float h[8] = {(float)dims[0], (float)dims[1], (float)dims[2], origin[0], origin[1],
                  origin[2], voxel_size, trunc_margin};
    TDM_TRY(write_all(f.f, h, sizeof h, path));
    return write_all(f.f, grid, sizeof(float) * (size_t)dims[0] * dims[1] * dims[2], path);
}

int tdm_read_grid_bin(const char *path, int dims[3], float origin[3], float *voxel_size,
                      float *trunc_margin, float **grid) {
    tdm::clear_error();
    TDM_CHECK(path && dims && origin && voxel_size && trunc_margin && grid, TDM_ERR_INVALID_ARG,
              "null pointer argument");
    File f(path, "rb");
    TDM_CHECK(f.f, TDM_ERR_IO, "cannot open %s", path);
    float h[8];
    TDM_TRY(read_all(f.f, h, sizeof h, path));
    for (int a = 0; a < 3; ++a) {
        TDM_CHECK(h[a] >= 1 && h[a] <= 1.0e6f && h[a] == std::floor(h[a]), TDM_ERR_FORMAT,
                  "%s: bad grid dims", path);
        dims[a] = (int)h[a];
        origin[a] = h[3 + a];
    }
    *voxel_size = h[6];
    *trunc_margin = h[7];
    size_t n = 0;
    TDM_TRY(payload_count(bytes_left(f.f), path,
                          {(size_t)dims[0], (size_t)dims[1], (size_t)dims[2]}, sizeof(float), &n));
    float *p = nullptr;
    TDM_TRY(alloc_copy(n, &p));
    int s = read_all(f.f, p, sizeof(float) * n, path);
    if (s != TDM_OK) {
        std::free(p);
        return s;
    }
    *grid = p;
    return TDM_OK;
}

int tdm_read_patch_volumes(const char *path, float **data, int *num, int dims[3]) {
    tdm::clear_error();
    TDM_CHECK(path && data && num && dims, TDM_ERR_INVALID_ARG, "null pointer argument");
    File f(path, "rb");
    TDM_CHECK(f.f, TDM_ERR_IO, "cannot open %s", path);
    int n = 0;
    TDM_TRY(read_count(f.f, path, &n));
    for (int a = 0; a < 3; ++a) {
        TDM_TRY(read_count(f.f, path, &dims[a]));
        TDM_CHECK(dims[a] > 0, TDM_ERR_FORMAT, "%s: zero volume dimension", path);
    }
    size_t count = 0;
    TDM_TRY(payload_count(bytes_left(f.f), path,
                          {(size_t)n, (size_t)dims[0], (size_t)dims[1], (size_t)dims[2]},
                          sizeof(float), &count));
    float *p = nullptr;
    TDM_TRY(alloc_copy(count, &p));
    int s = read_all(f.f, p, sizeof(float) * count, path);
    if (s != TDM_OK) {
        std::free(p);
        return s;
    }
    *data = p;
    *num = n;
    return TDM_OK;
}

int tdm_write_feat_bin(const char *path, const float *feat, int num, int dim) {
    tdm::clear_error();
    TDM_CHECK(path && num >= 0 && dim > 0 && (num == 0 || feat), TDM_ERR_INVALID_ARG,
              "bad arguments");
    File f(path, "wb");
    TDM_CHECK(f.f, TDM_ERR_IO, "cannot create %s", path);
    // evaluation/keypoint-matching/get3DMatchDescriptors.cu:50-56: N then the
    // four trailing dims of the feat response [dim, 1, 1, 1].
    float h[5] = {(float)num, (float)dim, 1.0f, 1.0f, 1.0f};
    TDM_TRY(write_all(f.f, h, sizeof h, path));
    return write_all(f.f, feat, sizeof(float) * (size_t)num * dim, path);
}

int tdm_load_matrix(const char *path, int rows, int cols, float *out) {
    tdm::clear_error();
    TDM_CHECK(path && out && rows > 0 && cols > 0, TDM_ERR_INVALID_ARG, "bad arguments");
    File f(path, "r");
    TDM_CHECK(f.f, TDM_ERR_IO, "cannot open %s", path);
    for (int i = 0; i < rows * cols; ++i)
        TDM_CHECK(std::fscanf(f.f, "%f", &out[i]) == 1, TDM_ERR_FORMAT,
                  "%s: expected %d numbers, found %d", path, rows * cols, i);
    return TDM_OK;
}

int tdm_read_png_u16(const char *path, uint16_t **mm, int *height, int *width) {
    tdm::clear_error();
    TDM_CHECK(path && mm && height && width, TDM_ERR_INVALID_ARG, "null pointer argument");
    std::vector<uint16_t> pix;
    int H = 0, W = 0;
    TDM_TRY(decode_png_u16(path, pix, H, W));
    uint16_t *p = nullptr;
    TDM_TRY(alloc_copy(pix.size(), &p));
    std::memcpy(p, pix.data(), pix.size() * sizeof(uint16_t));
    *mm = p;
    *height = H;
    *width = W;
    return TDM_OK;
}

int tdm_read_depth_png(const char *path, int height, int width, float *depth) {
    tdm::clear_error();
    TDM_CHECK(path && depth && height > 0 && width > 0, TDM_ERR_INVALID_ARG, "bad arguments");
    std::vector<uint16_t> pix;
    int H = 0, W = 0;
    TDM_TRY(decode_png_u16(path, pix, H, W));
    TDM_CHECK(H == height && W == width, TDM_ERR_FORMAT, "%s is %dx%d, expected %dx%d", path, W,
              H, width, height);
    for (size_t i = 0; i < pix.size(); ++i) {
        float d = (float)pix[i] / 1000.0f;
        depth[i] = d > 6.0f ? 0.0f : d;
    }
    return TDM_OK;
}

int tdm_multiply_matrix(const float a[16], const float b[16], float out[16]) {
    TDM_CHECK(a && b && out, TDM_ERR_INVALID_ARG, "null pointer argument");
    float r[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            float s = a[i * 4] * b[j];
            for (int k = 1; k < 4; ++k) s = s + a[i * 4 + k] * b[k * 4 + j];
            r[i * 4 + j] = s;
        }
    std::memcpy(out, r, sizeof r);
    return TDM_OK;
}

int tdm_invert_matrix(const float m[16], float out[16]) {
    tdm::clear_error();
    TDM_CHECK(m && out, TDM_ERR_INVALID_ARG, "null pointer argument");
    float v[16];
    for (int e = 0; e < 16; ++e) {
        const CofTerms &c = kCof[e];
        float acc = 0.0f;
        for (int t = 0; t < 6; ++t) {
            float p = m[c.idx[t][0]] * m[c.idx[t][1]] * m[c.idx[t][2]];
            int sign = c.first_sign * kCofPattern[t];
            if (t == 0) acc = sign > 0 ? p : -p;
            else acc = sign > 0 ? acc + p : acc - p;
        }
        v[e] = acc;
    }
    float det = m[0] * v[0] + m[1] * v[4] + m[2] * v[8] + m[3] * v[12];
    TDM_CHECK(det != 0, TDM_ERR_INVALID_ARG, "singular matrix");
    float inv_det = (float)(1.0 / (double)det);
    for (int i = 0; i < 16; ++i) out[i] = v[i] * inv_det;
    return TDM_OK;
}

int tdm_save_surface_ply(const char *path, const int dims[3], float voxel_size,
                         const float origin[3], const float *tsdf, const float *weight,
                         float tsdf_thresh, float weight_thresh, int *num_written) {
    tdm::clear_error();
    TDM_CHECK(path && dims && origin && tsdf && weight, TDM_ERR_INVALID_ARG,
              "null pointer argument");
    const long dx = dims[0], dy = dims[1], dz = dims[2];
    TDM_CHECK(dx > 0 && dy > 0 && dz > 0, TDM_ERR_INVALID_ARG, "bad dims");
    const long n = dx * dy * dz;
    std::vector<float> pts;
    for (long i = 0; i < n; ++i) {
        if (std::fabs(tsdf[i]) < tsdf_thresh && weight[i] > weight_thresh) {
            long z = i / (dx * dy), y = (i - z * dx * dy) / dx, x = i - z * dx * dy - y * dx;
            pts.push_back(origin[0] + (float)x * voxel_size);
            pts.push_back(origin[1] + (float)y * voxel_size);
            pts.push_back(origin[2] + (float)z * voxel_size);
        }
    }
    int count = (int)(pts.size() / 3);
    if (num_written) *num_written = count;
    return tdm_write_ply_points(path, pts.data(), count);
}

int tdm_marvin_write(const char *path, int num_tensors, const char *const *names,
                     const int *ndims, const int *const *dims, const float *const *data) {
    tdm::clear_error();
    TDM_CHECK(path && num_tensors >= 0, TDM_ERR_INVALID_ARG, "bad arguments");
    std::vector<tdm::MarvinTensor> ts(num_tensors);
    for (int t = 0; t < num_tensors; ++t) {
        TDM_CHECK(names[t] && ndims[t] >= 0 && (ndims[t] == 0 || dims[t]), TDM_ERR_INVALID_ARG,
                  "bad tensor %d", t);
        ts[t].name = names[t];
        ts[t].dims.assign(dims[t], dims[t] + ndims[t]);
        size_t n = 1;
        for (int d : ts[t].dims) {
            TDM_CHECK(d >= 0, TDM_ERR_INVALID_ARG, "negative dim in tensor %d", t);
            n *= (size_t)d;
        }
        TDM_CHECK(n == 0 || data[t], TDM_ERR_INVALID_ARG, "tensor %d has no data", t);
        ts[t].data.assign(data[t], data[t] + n);
    }
    return tdm::marvin_write(path, ts);
}

int tdm_marvin_count(const char *path, int *num_tensors, size_t *num_values) {
    tdm::clear_error();
    TDM_CHECK(path && num_tensors && num_values, TDM_ERR_INVALID_ARG, "null pointer argument");
    std::vector<tdm::MarvinTensor> ts;
    TDM_TRY(tdm::marvin_read(path, ts));
    *num_tensors = (int)ts.size();
    size_t n = 0;
    for (auto &t : ts) n += t.data.size();
    *num_values = n;
    return TDM_OK;
}

}  // extern "C"
