// Point cloud -> TDF on gfx950, and the 30^3 patch gather.
//
// Reference: core/demo.cu:22-57 (ComputeTDF) searches, for every voxel, the
// whole 11^3 cube of a float occupancy grid (up to 1331 scattered loads per
// empty voxel).  Here:
//   1. points are scattered into a 1-bit occupancy grid (rows of 32-bit
//      words, x fastest; 1/32 of the float grid's bytes);
//   2. one workgroup per 64 x 16 x 16 output tile loads the bits of the tile
//      and its r-voxel halo once and computes the truncated squared distance
//      transform separably in LDS: nearest set bit along x with bit scans,
//      then min over dy of (f + dy^2), then over dz;
//   3. d^2 indexes a table of 1 - sqrtf(d^2)/r built on the host with the
//      reference's float operations, so values are bit-exact without any
//      device sqrt or division; tiles whose halo holds no point write zeros.
// The minimum of dx^2+dy^2+dz^2 over occupied voxels of the clipped cube is
// exactly what the reference's max over 1 - sqrt(d^2)/r selects, because the
// table is non-increasing in d^2.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "common.hpp"
#include "pipeline.hpp"

#pragma clang fp contract(off)

// Device memory of tdm_tdf_from_points_noalloc: the bit grid and its tile
// flags (bits_bytes) of one call.  Each GPU holds its own copy for the life of
// the process; host code reaches it only through the per-device address that
// static_scratch_ptr looks up, never through this name.
static __device__ __align__(256) unsigned char scratch[TDM_TDF_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxR = 8;          // largest supported search radius (voxels)
constexpr int kLutSize = 3 * kMaxR * kMaxR + 1;
constexpr int kTX = 64, kTY = 16, kTZ = 16;
constexpr int kTdfThreads = 256;       // a tile's workgroup in the batched form
constexpr int kTdfThreadsOne = 1024;   // and for one cloud: a tile whose halo holds points
                                       // is the critical path, so it gets 16 waves
constexpr unsigned char kNone = 255;  // no occupied voxel within reach

struct Lut {
    float v[kLutSize];
};

// Host-side table, core/demo.cu:51-54: value of an occupied voxel at squared
// distance d2, or 0 when 1 - dist is not positive (the grid starts at 0 and
// only grows).
Lut make_lut(int r) {
    Lut l;
    for (int d2 = 0; d2 < kLutSize; ++d2) {
        volatile float fd2 = (float)d2;  // keep host float semantics explicit
        float dist = sqrtf(fd2) / (float)r;
        float v = 1.0f - dist;
        l.v[d2] = v > 0.0f ? v : 0.0f;
    }
    return l;
}

__host__ __device__ inline long row_words(int dx) { return (dx + 31) / 32; }

// ---- bounding box -------------------------------------------------------------

__global__ void bbox_partial_kernel(const float *__restrict__ pts, int n, float *__restrict__ part) {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            float v = pts[(long)i * 3 + a];
            lo[a] = fminf(lo[a], v);
            hi[a] = fmaxf(hi[a], v);
        }
    }
    __shared__ float s[6][256];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        s[a][threadIdx.x] = lo[a];
        s[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                s[a][threadIdx.x] = fminf(s[a][threadIdx.x], s[a][threadIdx.x + off]);
                s[3 + a][threadIdx.x] = fmaxf(s[3 + a][threadIdx.x], s[3 + a][threadIdx.x + off]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// One block of 256 threads folds the per-block partials (min/max are exact,
// so the order does not matter).
__global__ void bbox_final_kernel(const float *__restrict__ part, int nblocks, float *__restrict__ out) {
    __shared__ float s[6][256];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
#pragma unroll
        for (int a = 0; a < 6; ++a)
            v[a] = a < 3 ? fminf(v[a], part[b * 6 + a]) : fmaxf(v[a], part[b * 6 + a]);
#pragma unroll
    for (int a = 0; a < 6; ++a) s[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
#pragma unroll
            for (int a = 0; a < 6; ++a)
                s[a][threadIdx.x] = a < 3 ? fminf(s[a][threadIdx.x], s[a][threadIdx.x + off])
                                          : fmaxf(s[a][threadIdx.x], s[a][threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

// ---- occupancy ---------------------------------------------------------------------

__global__ void fill_u32_kernel(unsigned *__restrict__ p, size_t n, unsigned v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

// Zeroing of n floats at p: 16-byte stores over the aligned body (n4 float4
// from p + head), single floats for the up to 3 before and after it.
struct ZeroJob {
    float *p;
    size_t n, head, n4;
};

ZeroJob zero_job(float *p, size_t n) {  // p may be null (nothing to zero)
    ZeroJob z{p, p ? n : 0, 0, 0};
    if (p) z.head = std::min<size_t>(((16 - (reinterpret_cast<uintptr_t>(p) & 15)) & 15) / 4, n);
    z.n4 = (z.n - z.head) / 4;
    return z;
}

// Thread t of `stride` threads.
__device__ inline void zero_floats_part(const ZeroJob &z, size_t t, size_t stride) {
    float4 *__restrict__ body = reinterpret_cast<float4 *>(z.p + z.head);
    for (size_t i = t; i < z.n4; i += stride) body[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (t < z.head) z.p[t] = 0.0f;
    const size_t tail0 = z.head + 4 * z.n4;
    if (t < z.n - tail0) z.p[tail0 + t] = 0.0f;
}

// The floats of z and nb bytes at the 16-byte aligned b, in one launch.
__global__ void clear_kernel(ZeroJob z, unsigned char *__restrict__ b, size_t nb) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    zero_floats_part(z, t, stride);
    const size_t nb16 = nb / 16;
    uint4 *__restrict__ bbody = reinterpret_cast<uint4 *>(b);
    for (size_t i = t; i < nb16; i += stride) bbody[i] = make_uint4(0u, 0u, 0u, 0u);
    if (t < nb - 16 * nb16) b[16 * nb16 + t] = 0u;
}

// p and b may be null (nothing is cleared there).
void clear_regions(float *p, size_t n, unsigned char *b, size_t nb, hipStream_t s) {
    const ZeroJob z = zero_job(p, n);
    if (!b) nb = 0;
    const unsigned blocks = (unsigned)std::min<size_t>((z.n4 + nb / 16) / 256 + 1, 16384);
    clear_kernel<<<blocks, 256, 0, s>>>(z, b, nb);
}

// core/demo.cu:129-134: voxel = round((p - origin)/vs) per axis.
__device__ inline bool point_voxel(const float *__restrict__ pts, long i, int dx, int dy, int dz,
                                   float ox, float oy, float oz, float vs, int &gx, int &gy,
                                   int &gz) {
    float fx = roundf((pts[i * 3 + 0] - ox) / vs);
    float fy = roundf((pts[i * 3 + 1] - oy) / vs);
    float fz = roundf((pts[i * 3 + 2] - oz) / vs);
    if (!(fx >= 0.0f && fx < (float)dx && fy >= 0.0f && fy < (float)dy && fz >= 0.0f &&
          fz < (float)dz))
        return false;
    gx = (int)fx;
    gy = (int)fy;
    gz = (int)fz;
    return true;
}

// Flag byte of the 64x16x16 tile holding voxel (gx, gy, gz): set to 1 by
// every point in it (the same value from every writer, so plain stores do).
__device__ inline long tile_of(int gx, int gy, int gz, int dx, int dy) {
    const int ntx = (dx + kTX - 1) / kTX, nty = (dy + kTY - 1) / kTY;
    return ((long)(gz / kTZ) * nty + gy / kTY) * ntx + gx / kTX;
}

__global__ void scatter_bits_kernel(const float *__restrict__ pts, int n, int dx, int dy, int dz,
                                    float ox, float oy, float oz, float vs,
                                    unsigned *__restrict__ bits, unsigned char *__restrict__ flags) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int gx, gy, gz;
    if (!point_voxel(pts, i, dx, dy, dz, ox, oy, oz, vs, gx, gy, gz)) return;
    long w = ((long)gz * dy + gy) * row_words(dx) + (gx >> 5);
    atomicOr(&bits[w], 1u << (gx & 31));
    flags[tile_of(gx, gy, gz, dx, dy)] = 1;
}

// The same scatter with a per-workgroup hash table in LDS: a workgroup ORs
// the bits of its kScatThreads x kScatPts points per word in LDS, then issues
// one device-scope atomicOr per distinct word.  Points cluster on surfaces
// (the 200k-point benchmark fragment puts up to 1286 points in one word), and
// device-scope atomics to one address serialise at the point all XCDs share;
// on that fragment this kernel takes 17 us against 28-34 us for one atomic
// per point.  Word indices must stay below 0xffffffff, the table's empty mark.
constexpr int kScatThreads = 1024, kScatPts = 4;
constexpr int kScatSlots = 8192;  // twice the keys of a workgroup: a probe always ends
constexpr unsigned kScatEmpty = 0xffffffffu;

//
// Workgroups from `nscat` on zero the output grid (z) instead: that store
// stream needs the memory bandwidth the scatter leaves idle, and sharing the
// launch runs the two side by side.
__global__ __launch_bounds__(kScatThreads) void scatter_bits_dedup_kernel(
    const float *__restrict__ pts, int n, int dx, int dy, int dz, float ox, float oy, float oz,
    float vs, unsigned *__restrict__ bits, unsigned char *__restrict__ flags, ZeroJob z,
    unsigned nscat) {
    __shared__ unsigned keys[kScatSlots];
    __shared__ unsigned vals[kScatSlots];
    if (blockIdx.x >= nscat) {
        zero_floats_part(z, (size_t)(blockIdx.x - nscat) * kScatThreads + threadIdx.x,
                         (size_t)(gridDim.x - nscat) * kScatThreads);
        return;
    }
    for (int s = threadIdx.x; s < kScatSlots; s += kScatThreads) {
        keys[s] = kScatEmpty;
        vals[s] = 0u;
    }
    const long rw = row_words(dx);
    const long base = (long)blockIdx.x * kScatThreads * kScatPts + threadIdx.x;
    // all loads first, then the decodes, then the inserts
    float px[kScatPts], py[kScatPts], pz[kScatPts];
#pragma unroll
    for (int k = 0; k < kScatPts; ++k) {
        const long i = min(base + (long)k * kScatThreads, (long)n - 1);
        px[k] = pts[i * 3 + 0];
        py[k] = pts[i * 3 + 1];
        pz[k] = pts[i * 3 + 2];
    }
    unsigned w[kScatPts], bit[kScatPts];
#pragma unroll
    for (int k = 0; k < kScatPts; ++k) {
        // point_voxel's arithmetic on the loaded coordinates
        const float fx = roundf((px[k] - ox) / vs);
        const float fy = roundf((py[k] - oy) / vs);
        const float fz = roundf((pz[k] - oz) / vs);
        const bool ok = base + (long)k * kScatThreads < n && fx >= 0.0f && fx < (float)dx &&
                        fy >= 0.0f && fy < (float)dy && fz >= 0.0f && fz < (float)dz;
        const int gx = ok ? (int)fx : 0, gy = ok ? (int)fy : 0, gz = ok ? (int)fz : 0;
        w[k] = ok ? (unsigned)(((long)gz * dy + gy) * rw + (gx >> 5)) : kScatEmpty;
        bit[k] = 1u << (gx & 31);
        if (ok) flags[tile_of(gx, gy, gz, dx, dy)] = 1;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScatPts; ++k) {
        if (w[k] == kScatEmpty) continue;
        unsigned h = (w[k] * 2654435761u) & (kScatSlots - 1);  // Knuth's multiplicative hash
        for (int probe = 0; probe < kScatSlots; ++probe) {
            const unsigned prev = atomicCAS(&keys[h], kScatEmpty, w[k]);
            if (prev == kScatEmpty || prev == w[k]) {
                atomicOr(&vals[h], bit[k]);
                break;
            }
            h = (h + 1) & (kScatSlots - 1);
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < kScatSlots; s += kScatThreads)
        if (keys[s] != kScatEmpty) atomicOr(&bits[keys[s]], vals[s]);
}

__global__ void scatter_occ_kernel(const float *__restrict__ pts, int n, int dx, int dy, int dz,
                                   float ox, float oy, float oz, float vs, float *__restrict__ occ) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int gx, gy, gz;
    if (!point_voxel(pts, i, dx, dy, dz, ox, oy, oz, vs, gx, gy, gz)) return;
    occ[((long)gz * dy + gy) * dx + gx] = 1.0f;
}

// Float occupancy -> bit rows.  One wave handles 64 consecutive x of a row
// (two words) with a ballot.
__global__ void pack_bits_kernel(const float *__restrict__ occ, int dx, long rows,
                                 unsigned *__restrict__ bits) {
    const long rw = row_words(dx);
    const long pairs_per_row = (rw + 1) / 2;
    const long wave = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= rows * pairs_per_row) return;
    const long row = wave / pairs_per_row;
    const int x = (int)(wave % pairs_per_row) * 64 + lane;
    bool on = x < dx && occ[row * dx + x] > 0.0f;
    unsigned long long m = __ballot(on);
    long w = row * rw + (x >> 5);
    if (lane == 0) bits[w] = (unsigned)m;
    if (lane == 32 && (x >> 5) < rw) bits[w] = (unsigned)(m >> 32);
}

// ---- TDF tile kernel --------------------------------------------------------------

// LDS of one tile: halo bit words, pass-y output, table.  Pass x lives in
// registers (a wave runs passes x and y over whole z-slabs of the halo), so
// a workgroup holds about 38 KB at R = 5 and four fit on a CU.
template <int R>
struct TileLds {
    static constexpr int HY = kTY + 2 * R, HZ = kTZ + 2 * R;
    unsigned sw[HZ * HY * 4];
    unsigned char s2[HZ * kTY * kTX];
    float lut[kLutSize];
};

// Squared distance (or kNone) along x from voxel `lane` of the tile to the
// nearest set bit of halo row words w0..w3 (bit 0 of w0 = tile x - 32).
template <int R>
__device__ inline unsigned char xpass(unsigned w0, unsigned w1, unsigned w2, unsigned w3, int lane) {
    if (!(w0 | w1 | w2 | w3)) return kNone;
    // The window of 2R+1 (<= 17) bits starting at bit s = lane + 32 - R of the
    // 128-bit row w3:w2:w1:w0 lies in words k = s / 32 and k + 1 (s % 32 + 2R
    // < 64), so one 32-bit funnel shift extracts it.
    const int s = lane + 32 - R;  // in [32 - R, 95 - R]: k is 0, 1 or 2
    const int k = s >> 5;
    const unsigned lo = k == 0 ? w0 : k == 1 ? w1 : w2;
    const unsigned hi = k == 0 ? w1 : k == 1 ? w2 : w3;
    unsigned win = __builtin_amdgcn_alignbit(hi, lo, (unsigned)(s & 31));
    win &= (1u << (2 * R + 1)) - 1u;
    if (!win) return kNone;
    const unsigned right = win >> R;               // bits at >= x
    const unsigned left = win & ((2u << R) - 1u);  // bits at <= x
    int d = 255;
    if (right) d = __builtin_ctz(right);
    if (left) d = min(d, R - (31 - __builtin_clz(left)));
    return (unsigned char)(d * d);
}

// Truncated squared-distance transform of the 64x16x16 tile at (X0, Y0, Z0).
// With `flags` (one byte per tile: any occupied voxel inside it) the grid has
// been zeroed beforehand with full-line stores: a tile whose 27 neighbouring
// tiles -- which contain its halo, R < 16 -- hold no occupied voxel, or whose
// halo turns out empty, writes nothing.  (A tile row is 256 bytes at a row
// pitch of 4*dx bytes, generally not line-aligned; storing a zero grid that
// way runs at well under half the rate of a contiguous fill.)  Without
// `flags` every tile writes all its voxels.
template <int R, int NT>
__device__ inline void tdf_tile(TileLds<R> &L, const unsigned *__restrict__ bits, int dx, int dy,
                                int dz, const Lut &lut, float *__restrict__ tdf, int X0, int Y0,
                                int Z0, const unsigned char *__restrict__ flags) {
    constexpr int HY = kTY + 2 * R, HZ = kTZ + 2 * R;
    unsigned *sw = L.sw;
    unsigned char *s2 = L.s2;
    float *slut = L.lut;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const long rw = row_words(dx);
    const long wbase = X0 >> 5;  // first word of the tile in a row

    if (flags) {
        bool near = false;
        if (threadIdx.x < 27) {
            const int ntx = (dx + kTX - 1) / kTX, nty = (dy + kTY - 1) / kTY, ntz = (dz + kTZ - 1) / kTZ;
            const int tx = X0 / kTX + (int)threadIdx.x % 3 - 1;
            const int ty = Y0 / kTY + ((int)threadIdx.x / 3) % 3 - 1;
            const int tz = Z0 / kTZ + (int)threadIdx.x / 9 - 1;
            if (tx >= 0 && tx < ntx && ty >= 0 && ty < nty && tz >= 0 && tz < ntz)
                near = flags[((long)tz * nty + ty) * ntx + tx] != 0;
        }
        if (__syncthreads_or(near) == 0) return;
    }

    for (int i = threadIdx.x; i < kLutSize; i += NT) slut[i] = lut.v[i];
    // All halo words at once (independent loads, so their latencies overlap).
    unsigned seen = 0;
    for (int i = threadIdx.x; i < HZ * HY * 4; i += NT) {
        const int row = i >> 2, wi = (int)(i & 3);
        const int y = Y0 + row % HY - R, z = Z0 + row / HY - R;
        const long w = wbase - 1 + wi;
        unsigned v = 0u;
        if (y >= 0 && y < dy && z >= 0 && z < dz && w >= 0 && w < rw)
            v = bits[((long)z * dy + y) * rw + w];
        sw[i] = v;
        seen |= v;
    }
    const bool empty = __syncthreads_or(seen != 0u) == 0;
    if (empty && flags) return;

    if (!empty) {
        // Passes x and y over the halo z-slab zz: pass x of its HY rows in
        // registers, then min over |d| <= R of f(y + d) + d^2 for the tile's y.
        for (int zz = wave; zz < HZ; zz += NT / 64) {
            unsigned char f[HY];
#pragma unroll
            for (int yy = 0; yy < HY; ++yy) {
                const unsigned *w = &sw[(zz * HY + yy) * 4];
                f[yy] = xpass<R>(w[0], w[1], w[2], w[3], lane);
            }
            // kNone (255) exceeds every reachable sum (2 R^2 <= 128), so it needs
            // no test: a result >= kNone means nothing within reach.
#pragma unroll
            for (int y = 0; y < kTY; ++y) {
                int best = kNone;
#pragma unroll
                for (int d = -R; d <= R; ++d) best = min(best, (int)f[y + R + d] + d * d);
                s2[(zz * kTY + y) * kTX + lane] = (unsigned char)min(best, (int)kNone);
            }
        }
        __syncthreads();
    }

    // Pass z and the table lookup; write the tile.
    const int x = X0 + lane;
    for (int row = wave; row < kTY * kTZ; row += NT / 64) {
        const int y = row % kTY, zl = row / kTY;
        const int gy = Y0 + y, gz = Z0 + zl;
        if (gy >= dy || gz >= dz || x >= dx) continue;
        float val = 0.0f;
        if (!empty) {
            int best = kNone;
#pragma unroll
            for (int d = -R; d <= R; ++d)
                best = min(best, (int)s2[((zl + R + d) * kTY + y) * kTX + lane] + d * d);
            if (best < kLutSize) val = slut[best];
        }
        tdf[((long)gz * dy + gy) * dx + x] = val;
    }
}

template <int R>
__global__ __launch_bounds__(kTdfThreadsOne) void tdf_tile_kernel(const unsigned *__restrict__ bits,
                                                               int dx, int dy, int dz, Lut lut,
                                                               float *__restrict__ tdf,
                                                               const unsigned char *__restrict__ flags) {
    __shared__ TileLds<R> L;
    tdf_tile<R, kTdfThreadsOne>(L, bits, dx, dy, dz, lut, tdf, blockIdx.x * kTX, blockIdx.y * kTY,
                blockIdx.z * kTZ, flags);
}

// One byte per 64x16x16 tile: whether any voxel of the tile is occupied
// (one wave per tile ORs the tile's 16 x 16 rows of two words).
__global__ void tile_flags_kernel(const unsigned *__restrict__ bits, int dx, int dy, int dz,
                                  unsigned char *__restrict__ flags) {
    const int ntx = (dx + kTX - 1) / kTX, nty = (dy + kTY - 1) / kTY, ntz = (dz + kTZ - 1) / kTZ;
    const long tile = ((long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (tile >= (long)ntx * nty * ntz) return;  // whole wave
    const int tx = (int)(tile % ntx), ty = (int)((tile / ntx) % nty), tz = (int)(tile / ((long)ntx * nty));
    const long rw = row_words(dx);
    const long w0 = (long)tx * (kTX / 32);
    unsigned v = 0u;
    for (int i = lane; i < kTY * kTZ * 2; i += 64) {
        const int row = i >> 1, wi = i & 1;
        const int y = ty * kTY + row % kTY, z = tz * kTZ + row / kTY;
        if (y < dy && z < dz && w0 + wi < rw) v |= bits[((long)z * dy + y) * rw + w0 + wi];
    }
    const bool any = __ballot(v != 0u) != 0ull;
    if (lane == 0) flags[tile] = any ? 1 : 0;
}

// ---- many clouds in one launch per step ---------------------------------------

// One cloud of a batch (device table, filled by the plan on the host).
struct CloudDev {
    const float *pts;
    float *tdf;
    long pt_off;    // first point of this cloud in the batch's point numbering
    long word_off;  // first bit word of this cloud's grid in the batch bit buffer
    long tile_off;  // first tile of this cloud in the batch's tile numbering
    long vox_off;   // first voxel of this cloud in the batch's voxel numbering
    int n;
    int dims[3];
    float origin[3];
    int tiles[3];
};

// Last cloud c with key(c) <= v (keys ascending, c in [0, num)).
template <typename Key>
__device__ inline int find_cloud(const CloudDev *__restrict__ clouds, int num, long v, Key key) {
    int lo = 0, hi = num - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (key(clouds[mid]) <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void scatter_bits_batch_kernel(const CloudDev *__restrict__ clouds, int num,
                                          long total, float vs, unsigned *__restrict__ bits,
                                          unsigned char *__restrict__ flags) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x) {
        const int c = find_cloud(clouds, num, i, [](const CloudDev &d) { return d.pt_off; });
        const CloudDev &cd = clouds[c];
        const long j = i - cd.pt_off;
        if (j >= cd.n) continue;  // (an empty cloud shares its offset with the next)
        int gx, gy, gz;
        if (!point_voxel(cd.pts, j, cd.dims[0], cd.dims[1], cd.dims[2], cd.origin[0],
                         cd.origin[1], cd.origin[2], vs, gx, gy, gz))
            continue;
        const long w =
            cd.word_off + ((long)gz * cd.dims[1] + gy) * row_words(cd.dims[0]) + (gx >> 5);
        atomicOr(&bits[w], 1u << (gx & 31));
        flags[cd.tile_off + tile_of(gx, gy, gz, cd.dims[0], cd.dims[1])] = 1;
    }
}

// The batch scatter with the words of a workgroup's 4096 points deduplicated
// in LDS before the device-scope atomics (as scatter_bits_dedup_kernel);
// keys are batch word indices, below 0xffffffff.
__global__ __launch_bounds__(kScatThreads) void scatter_bits_batch_dedup_kernel(
    const CloudDev *__restrict__ clouds, int num, long total, float vs,
    unsigned *__restrict__ bits, unsigned char *__restrict__ flags) {
    __shared__ unsigned keys[kScatSlots];
    __shared__ unsigned vals[kScatSlots];
    for (int s = threadIdx.x; s < kScatSlots; s += kScatThreads) {
        keys[s] = kScatEmpty;
        vals[s] = 0u;
    }
    const long base = (long)blockIdx.x * kScatThreads * kScatPts + threadIdx.x;
    // the workgroup's points are consecutive: one search for the first, then
    // each point steps forward over the (few) cloud boundaries before it
    const int c0 = find_cloud(clouds, num, (long)blockIdx.x * kScatThreads * kScatPts,
                              [](const CloudDev &d) { return d.pt_off; });
    unsigned w[kScatPts], bit[kScatPts];
#pragma unroll
    for (int k = 0; k < kScatPts; ++k) {
        w[k] = kScatEmpty;
        bit[k] = 0u;
        const long i = base + (long)k * kScatThreads;
        if (i >= total) continue;
        int c = c0;
        while (c + 1 < num && clouds[c + 1].pt_off <= i) ++c;
        const CloudDev &cd = clouds[c];
        const long j = i - cd.pt_off;
        if (j >= cd.n) continue;  // (an empty cloud shares its offset with the next)
        int gx, gy, gz;
        if (!point_voxel(cd.pts, j, cd.dims[0], cd.dims[1], cd.dims[2], cd.origin[0],
                         cd.origin[1], cd.origin[2], vs, gx, gy, gz))
            continue;
        w[k] = (unsigned)(cd.word_off + ((long)gz * cd.dims[1] + gy) * row_words(cd.dims[0]) +
                          (gx >> 5));
        bit[k] = 1u << (gx & 31);
        flags[cd.tile_off + tile_of(gx, gy, gz, cd.dims[0], cd.dims[1])] = 1;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScatPts; ++k) {
        if (w[k] == kScatEmpty) continue;
        unsigned h = (w[k] * 2654435761u) & (kScatSlots - 1);
        for (int probe = 0; probe < kScatSlots; ++probe) {
            const unsigned prev = atomicCAS(&keys[h], kScatEmpty, w[k]);
            if (prev == kScatEmpty || prev == w[k]) {
                atomicOr(&vals[h], bit[k]);
                break;
            }
            h = (h + 1) & (kScatSlots - 1);
        }
    }
    __syncthreads();
    for (int s = threadIdx.x; s < kScatSlots; s += kScatThreads)
        if (keys[s] != kScatEmpty) atomicOr(&bits[keys[s]], vals[s]);
}

// Zero every cloud's output grid: workgroup b takes voxels [b*chunk,
// (b+1)*chunk) of the batch's voxel numbering, cloud by cloud (a chunk may
// span clouds; the cloud of its start is found once).
__global__ void zero_clouds_kernel(const CloudDev *__restrict__ clouds, int num, long total_vox,
                                   long chunk) {
    long v = (long)blockIdx.x * chunk;
    const long end = min(v + chunk, total_vox);
    int c = find_cloud(clouds, num, v, [](const CloudDev &d) { return d.vox_off; });
    while (v < end) {
        const CloudDev &cd = clouds[c];
        const long cend = cd.vox_off + (long)cd.dims[0] * cd.dims[1] * cd.dims[2];
        const long stop = min(end, cend);
        for (long e = v + threadIdx.x; e < stop; e += blockDim.x) cd.tdf[e - cd.vox_off] = 0.0f;
        v = stop;
        ++c;
    }
}

template <int R>
__global__ __launch_bounds__(kTdfThreads) void tdf_tile_batch_kernel(
    const CloudDev *__restrict__ clouds, int num, const unsigned *__restrict__ bits, Lut lut,
    const unsigned char *__restrict__ flags) {
    __shared__ TileLds<R> L;
    const long t = blockIdx.x;
    const int c = find_cloud(clouds, num, t, [](const CloudDev &d) { return d.tile_off; });
    const CloudDev &cd = clouds[c];
    const long lt = t - cd.tile_off;
    const int tx = (int)(lt % cd.tiles[0]);
    const int ty = (int)((lt / cd.tiles[0]) % cd.tiles[1]);
    const int tz = (int)(lt / ((long)cd.tiles[0] * cd.tiles[1]));
    tdf_tile<R, kTdfThreads>(L, bits + cd.word_off, cd.dims[0], cd.dims[1], cd.dims[2], lut, cd.tdf, tx * kTX,
                ty * kTY, tz * kTZ, flags + cd.tile_off);
}

// float32 decode of core/demo.cu:32-34.
__device__ inline void float_decode(long v, int dx, int dy, int &x, int &y, int &z) {
    int vi = (int)v;
    z = (int)floorf((float)vi / ((float)dx * (float)dy));
    y = (int)floorf(((float)vi - ((float)z * (float)dx * (float)dy)) / (float)dx);
    x = (int)((float)vi - ((float)z * (float)dx * (float)dy) - ((float)y * (float)dx));
}

__device__ inline bool bit_at(const unsigned *__restrict__ bits, long rw, int dy, int x, int y,
                              int z) {
    return (bits[((long)z * dy + y) * rw + (x >> 5)] >> (x & 31)) & 1u;
}

// n floats from src to dst: 16-byte pieces over a grid-stride loop, then the
// tail (n need not be a multiple of 4; both buffers come from hipMalloc).
__global__ void copy_floats_kernel(const float *__restrict__ src, float *__restrict__ dst,
                                   long n) {
    const long n4 = n / 4;
    const long stride = (long)gridDim.x * blockDim.x;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride)
        reinterpret_cast<float4 *>(dst)[i] = reinterpret_cast<const float4 *>(src)[i];
    const long t = 4 * n4 + (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = src[t];
}

// Reference decode for one grid, from the exact TDF in a separate buffer.
// Voxel v searches around its float32 decode (x', y', z') but takes its
// occupancy at v (core/demo.cu:32-56).  Where the decode is v itself the value
// is exact[v]; where it is another voxel j inside the grid the search is the
// one that gives j its exact TDF (an occupied j is within reach at distance
// 0), so v gets 1 if occupied, else exact[j]; decodes outside the grid are
// searched here around the decoded voxel (core/demo.cu:44-56, bounds clipped).
// Rows of voxels per thread in tdf_refgather_kernel.
constexpr int kRefRows = 4;

__global__ void tdf_refgather_kernel(const unsigned *__restrict__ bits, int dx, int dy, int dz,
                                     int r, Lut lut, const float *__restrict__ exact,
                                     float *__restrict__ tdf) {
    // One workgroup per 256 x of kRefRows consecutive rows (z, y): a thread
    // takes the same x in each, the row coordinates need only (uniform)
    // divisions, and the loads of all kRefRows voxels go out before any is
    // used -- with one voxel per thread a wave made a single round trip to a
    // grid that does not fit in L2 and the kernel ran on latency.
    const long rw = row_words(dx);
    const unsigned nbx = (unsigned)(dx + 255) / 256;
    const unsigned rb = blockIdx.x / nbx;
    const int tx = (int)(blockIdx.x - rb * nbx) * 256 + (int)threadIdx.x;
    if (tx >= dx) return;
    const long nrows = (long)dy * dz;
    long v[kRefRows];
    int x[kRefRows], y[kRefRows], z[kRefRows], ty[kRefRows], tz[kRefRows];
    bool in[kRefRows];
    unsigned word[kRefRows];
    float ej[kRefRows];
#pragma unroll
    for (int k = 0; k < kRefRows; ++k) {
        const long row = min((long)rb * kRefRows + k, nrows - 1);
        ty[k] = (int)(row % dy);
        tz[k] = (int)(row / dy);
        v[k] = row * dx + tx;
        float_decode(v[k], dx, dy, x[k], y[k], z[k]);
        // the word holding v's occupancy bit, and the exact TDF at the
        // decoded voxel j (at v when the decode leaves the grid)
        in[k] = x[k] >= 0 && x[k] < dx && y[k] >= 0 && y[k] < dy && z[k] >= 0 && z[k] < dz;
        word[k] = bits[row * rw + (tx >> 5)];
        ej[k] = exact[in[k] ? ((long)z[k] * dy + y[k]) * dx + x[k] : v[k]];
    }
#pragma unroll
    for (int k = 0; k < kRefRows; ++k) {
        if ((long)rb * kRefRows + k >= nrows) break;
        float val;
        if (x[k] == tx && y[k] == ty[k] && z[k] == tz[k]) {
            val = ej[k];
        } else if ((word[k] >> (tx & 31)) & 1u) {
            val = 1.0f;
        } else if (in[k]) {
            val = ej[k];
        } else {
            int best = kLutSize;
            for (int ix = max(0, x[k] - r); ix < min(dx, x[k] + r + 1); ++ix)
                for (int iy = max(0, y[k] - r); iy < min(dy, y[k] + r + 1); ++iy)
                    for (int iz = max(0, z[k] - r); iz < min(dz, z[k] + r + 1); ++iz)
                        if (bit_at(bits, rw, dy, ix, iy, iz)) {
                            const int d2 = (x[k] - ix) * (x[k] - ix) + (y[k] - iy) * (y[k] - iy) +
                                           (z[k] - iz) * (z[k] - iz);
                            best = min(best, d2);
                        }
            val = best < kLutSize ? lut.v[best] : 0.0f;
        }
        tdf[v[k]] = val;
    }
}

// Workgroups of tdf_refgather_kernel for a dx x dy x dz grid.
inline unsigned refgather_blocks(int dx, int dy, int dz) {
    return (unsigned)((long)((dx + 255) / 256) * (((long)dy * dz + kRefRows - 1) / kRefRows));
}

// ---- patch gather --------------------------------------------------------------------

__global__ void extract_patches_kernel(const float *__restrict__ tdf, int dx, int dy, int dz,
                                       const int *__restrict__ kp, int num, int size,
                                       float *__restrict__ out) {
    const long vol = (long)size * size * size;
    const long total = vol * num;
    const int h = size / 2;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x) {
        const long p = i / vol;
        const int li = (int)(i % vol);
        const int lx = li % size, ly = (li / size) % size, lz = li / (size * size);
        const int x = kp[p * 3 + 0] - h + lx, y = kp[p * 3 + 1] - h + ly, z = kp[p * 3 + 2] - h + lz;
        float v = 0.0f;
        if (x >= 0 && y >= 0 && z >= 0 && x < dx && y < dy && z < dz)
            v = tdf[((long)z * dy + y) * dx + x];
        out[i] = v;
    }
}

// ---- drivers -----------------------------------------------------------------------

// ---- accurate TDF: distance from voxel centres to the nearest point ------------
//
// training/match.hpp:13-51 (Get1NNSearchDist) and
// deprecated/pointCloud2AccTDF.m: tdf = 1 - min(d, trunc) / trunc with d the
// Euclidean distance from the voxel centre to the nearest point.  Points are
// binned into 8x8x8-voxel tiles (count, scan, fill); one workgroup per tile
// stages the points of the tiles within reach through LDS and each thread
// keeps the minimum squared distance of its voxel.
constexpr int kAT = 8;  // tile edge (voxels)

__device__ inline long acc_tile_of(int gx, int gy, int gz, int tx, int ty) {
    return ((long)(gz / kAT) * ty + gy / kAT) * tx + gx / kAT;
}

// Binning of the points by 8x8x8 tile with the bins of a workgroup's 4096
// points aggregated in an LDS hash table (tile -> count), so the device-scope
// atomics on the clustered bins come once per distinct bin and workgroup.
constexpr int kBinThreads = 1024, kBinPts = 4, kBinSlots = 8192;  // slots: twice the keys

struct BinTable {
    unsigned key[kBinSlots];
    unsigned cnt[kBinSlots];  // count, then (fill) the slot's first index in binned
};

// Decodes this thread's kBinPts points (tile index or kScatEmpty) and counts
// them into the table; slot and rank per point.  Every thread calls it.
// A point whose voxel lies outside the grid but within `margin` voxels of it
// can still be the nearest point of a border voxel (the reference searches
// every point it is given, training/match.hpp:40-48): it is binned into the
// tile of its voxel clamped to the grid.  Clamping moves the voxel no
// farther, per axis, from any voxel of the grid, so every tile whose walk
// reaches the true voxel also reaches the clamped one.
__device__ inline void bin_points(BinTable &T, const float *__restrict__ pts, int n, int dx,
                                  int dy, int dz, float ox, float oy, float oz, float vs,
                                  int margin, int tx, int ty, float4 (&p)[kBinPts],
                                  int (&slot)[kBinPts], unsigned (&rank)[kBinPts]) {
    for (int i = threadIdx.x; i < kBinSlots; i += kBinThreads) {
        T.key[i] = kScatEmpty;
        T.cnt[i] = 0u;
    }
    const long base = (long)blockIdx.x * kBinThreads * kBinPts + threadIdx.x;
    unsigned key[kBinPts];
#pragma unroll
    for (int k = 0; k < kBinPts; ++k) {
        const long i = min(base + (long)k * kBinThreads, (long)n - 1);
        p[k] = make_float4(pts[i * 3 + 0], pts[i * 3 + 1], pts[i * 3 + 2], 0.0f);
    }
#pragma unroll
    for (int k = 0; k < kBinPts; ++k) {
        // point_voxel's arithmetic on the loaded coordinates
        const float fx = roundf((p[k].x - ox) / vs);
        const float fy = roundf((p[k].y - oy) / vs);
        const float fz = roundf((p[k].z - oz) / vs);
        const float m = (float)margin;
        const bool ok = base + (long)k * kBinThreads < n && fx >= -m && fx <= (float)(dx - 1) + m &&
                        fy >= -m && fy <= (float)(dy - 1) + m && fz >= -m &&
                        fz <= (float)(dz - 1) + m;
        const int gx = (int)fminf(fmaxf(fx, 0.0f), (float)(dx - 1));
        const int gy = (int)fminf(fmaxf(fy, 0.0f), (float)(dy - 1));
        const int gz = (int)fminf(fmaxf(fz, 0.0f), (float)(dz - 1));
        key[k] = ok ? (unsigned)acc_tile_of(gx, gy, gz, tx, ty) : kScatEmpty;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kBinPts; ++k) {
        slot[k] = -1;
        if (key[k] == kScatEmpty) continue;
        unsigned h = (key[k] * 2654435761u) & (kBinSlots - 1);
        for (int probe = 0; probe < kBinSlots; ++probe) {
            const unsigned prev = atomicCAS(&T.key[h], kScatEmpty, key[k]);
            if (prev == kScatEmpty || prev == key[k]) {
                rank[k] = atomicAdd(&T.cnt[h], 1u);
                slot[k] = (int)h;
                break;
            }
            h = (h + 1) & (kBinSlots - 1);
        }
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBinThreads) void acc_count_kernel(
    const float *__restrict__ pts, int n, int dx, int dy, int dz, float ox, float oy, float oz,
    float vs, int margin, int tx, int ty, int *__restrict__ counts) {
    __shared__ BinTable T;
    float4 p[kBinPts];
    int slot[kBinPts];
    unsigned rank[kBinPts];
    bin_points(T, pts, n, dx, dy, dz, ox, oy, oz, vs, margin, tx, ty, p, slot, rank);
    for (int i = threadIdx.x; i < kBinSlots; i += kBinThreads)
        if (T.key[i] != kScatEmpty) atomicAdd(&counts[T.key[i]], (int)T.cnt[i]);
}

// Exclusive scan of n counts in two launches: per-1024 block sums, then
// each block scans its own counts on top of the sum of the blocks before it.
// offsets[n] = total.
__device__ inline int block_sum_1024(int v, int *red) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    __syncthreads();
    if (lane == 0) red[wave] = v;
    __syncthreads();
    int tot = 0;
    for (int w = 0; w < 16; ++w) tot += red[w];
    return tot;
}

__global__ __launch_bounds__(1024) void acc_blocksum_kernel(const int *__restrict__ counts, long n,
                                                            int *__restrict__ bsum) {
    __shared__ int red[16];
    const long i = (long)blockIdx.x * 1024 + threadIdx.x;
    const int tot = block_sum_1024(i < n ? counts[i] : 0, red);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void acc_offsets_kernel(const int *__restrict__ counts, long n,
                                                           const int *__restrict__ bsum,
                                                           int *__restrict__ offsets) {
    __shared__ int red[16];
    __shared__ int wsum[16];
    int part = 0;
    for (unsigned b = threadIdx.x; b < blockIdx.x; b += 1024) part += bsum[b];
    const int base = block_sum_1024(part, red);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long i = (long)blockIdx.x * 1024 + threadIdx.x;
    const int v = i < n ? counts[i] : 0;
    int inc = v;  // inclusive scan within the wave
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(inc, off);
        if (lane >= off) inc += u;
    }
    if (lane == 63) wsum[wave] = inc;
    __syncthreads();
    int before = 0;
    for (int w = 0; w < wave; ++w) before += wsum[w];
    const int excl = base + before + inc - v;
    if (i < n) offsets[i] = excl;
    if (i == n - 1) offsets[n] = excl + v;
}

// Bin i's points go to binned[offsets[i] ..): a workgroup reserves the range
// of each of its bins with one atomic on cursor[] and places its points by
// their rank in the LDS count (order within a bin does not matter: the
// kernel below takes a minimum).
__global__ __launch_bounds__(kBinThreads) void acc_fill_kernel(
    const float *__restrict__ pts, int n, int dx, int dy, int dz, float ox, float oy, float oz,
    float vs, int margin, int tx, int ty, const int *__restrict__ offsets,
    int *__restrict__ cursor, float4 *__restrict__ binned) {
    __shared__ BinTable T;
    float4 p[kBinPts];
    int slot[kBinPts];
    unsigned rank[kBinPts];
    bin_points(T, pts, n, dx, dy, dz, ox, oy, oz, vs, margin, tx, ty, p, slot, rank);
    for (int i = threadIdx.x; i < kBinSlots; i += kBinThreads)
        if (T.key[i] != kScatEmpty)
            T.cnt[i] = (unsigned)offsets[T.key[i]] + (unsigned)atomicAdd(&cursor[T.key[i]], (int)T.cnt[i]);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kBinPts; ++k)
        if (slot[k] >= 0) binned[T.cnt[slot[k]] + rank[k]] = p[k];
}

// Candidates per work item of acc_tdf_kernel: a tile whose neighbour bins
// hold more points is split into segments of this many candidates, each a
// work item of its own (the points of a fragment crowd near its camera: the
// heaviest tile of a 200k-point view has ~50x the median's candidates).
constexpr int kAccSeg = 1024;

// The work items of acc_tdf_kernel: (tile, segment) for every segment of
// kAccSeg candidates of the tiles with a point within reach (candidates =
// the points of the bins within `reach` tiles, from the offsets of the bin
// scan).  A wave scan numbers the items of the wave's tiles, and one atomic
// per wave reserves their places in `list`; the order of the list does not
// matter, the items are independent.
__global__ __launch_bounds__(256) void acc_near_kernel(const int *__restrict__ offsets, int tx,
                                                       int ty, int tz, int reach,
                                                       int *__restrict__ count,
                                                       int2 *__restrict__ list) {
    const long t = (long)blockIdx.x * 256 + threadIdx.x;
    const long nt = (long)tx * ty * tz;
    int total = 0;
    if (t < nt) {
        const int bx = (int)(t % tx), by = (int)((t / tx) % ty), bz = (int)(t / ((long)tx * ty));
        const int x0 = max(0, bx - reach), x1 = min(tx - 1, bx + reach);
        for (int z = max(0, bz - reach); z <= min(tz - 1, bz + reach); ++z)
            for (int y = max(0, by - reach); y <= min(ty - 1, by + reach); ++y) {
                const long row = ((long)z * ty + y) * tx;
                total += offsets[row + x1 + 1] - offsets[row + x0];  // the points of the run
            }
    }
    const int n = (total + kAccSeg - 1) / kAccSeg;
    const int lane = threadIdx.x & 63;
    int v = n;  // inclusive scan over the wave
    for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    int base = 0;
    if (lane == 63 && v > 0) base = atomicAdd(count, v);
    base = __shfl(base, 63) + v - n;
    for (int k = 0; k < n; ++k) list[base + k] = make_int2((int)t, k);
}

// 8x8x8 tiles, one per workgroup of 256 threads, over the list of tiles with
// a point in reach.  The grid is zeroed first (a full-line fill), and a
// voxel's value is 0 whenever no point lies within the truncation distance,
// so only those tiles are visited and only voxels with a point in reach are
// stored.
// The candidates are the points of the tiles within `reach` tiles (reach*8
// voxels >= the truncation radius plus the half voxel a point may sit off
// its voxel): the workgroup reads the (2 reach + 1)^3 bin ranges at once,
// numbers their points consecutively (prefix sums in LDS), and streams them
// 256 at a time through LDS, keeping only those that can be within the
// truncation distance of one of the tile's voxel centres.  Each wave then
// keeps, in a list of its own, the ones within reach of its 4x4x8 block of
// the tile, and each lane takes two voxels of that block, (x, y, z) and
// (x, y, z + 4), which share the x and y terms of the distance.  The tests
// use the box of the centres, computed with the same float operations as
// the centres (so it contains them); float rounding is monotone, so every
// centre's d^2 is at least the box's, and the threshold trunc^2 (1 + 2^-20)
// clears the rounding of trunc^2 itself.  A tile with no candidate returns
// at once.
constexpr int kAccMaxBins = 9 * 9 * 9;  // (2 reach + 1)^3 with reach <= 4
constexpr int kAccThreads = 256;

struct AccBox {
    float lox, hix, loy, hiy, loz, hiz;
    // squared distance from q to the box, summed as the voxel distances are
    __device__ float d2(const float4 &q) const {
        const float ex = fmaxf(fmaxf(lox - q.x, q.x - hix), 0.0f);
        const float ey = fmaxf(fmaxf(loy - q.y, q.y - hiy), 0.0f);
        const float ez = fmaxf(fmaxf(loz - q.z, q.z - hiz), 0.0f);
        return ex * ex + ey * ey + ez * ez;
    }
};

// The box of the centres of voxels [x0, x0 + nx) x [y0, y0 + ny) x [z0, z0 + nz).
__device__ inline AccBox acc_box(int x0, int y0, int z0, int nx, int ny, int nz, float ox,
                                 float oy, float oz, float vs) {
    return {ox + (float)x0 * vs,        ox + (float)(x0 + nx - 1) * vs,
            oy + (float)y0 * vs,        oy + (float)(y0 + ny - 1) * vs,
            oz + (float)z0 * vs,        oz + (float)(z0 + nz - 1) * vs};
}

// One work item of acc_tdf_kernel: segment `seg` of tile t's candidates (all
// 256 threads of the workgroup call it).
__device__ inline void acc_tile(const float4 *__restrict__ binned, const int *__restrict__ offsets,
                                int t, int seg, int dx, int dy, int dz, int tx, int ty, int tz,
                                int reach,
                                float ox, float oy, float oz, float vs, float trunc, float4 *sp,
                                float4 *wsp, int *bstart, int *bpre, int *wave_n,
                                float *__restrict__ tdf) {
    const int bx = t % tx, by = (t / tx) % ty, bz = t / (tx * ty);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // wave w: x half (w & 1), y half (w >> 1), all of z; lane: 4x4x4
    const int sx = bx * kAT + (wave & 1) * 4, sy = by * kAT + (wave >> 1) * 4, sz = bz * kAT;
    const int x = sx + (lane & 3), y = sy + ((lane >> 2) & 3), z = sz + (lane >> 4);
    const int x0 = max(0, bx - reach), y0 = max(0, by - reach), z0 = max(0, bz - reach);
    const int nx = min(tx - 1, bx + reach) - x0 + 1, ny = min(ty - 1, by + reach) - y0 + 1,
              nz = min(tz - 1, bz + reach) - z0 + 1;
    const int nbins = nx * ny * nz;
    // bin ranges, then their exclusive prefix (wave 0, 64 bins per step)
    for (int k = threadIdx.x; k < nbins; k += kAccThreads) {
        const int kx = k % nx, ky = (k / nx) % ny, kz = k / (nx * ny);
        const long nt = ((long)(z0 + kz) * ty + (y0 + ky)) * tx + (x0 + kx);
        const int a = offsets[nt];
        bstart[k] = a;
        bpre[k + 1] = offsets[nt + 1] - a;  // count, scanned below
    }
    __syncthreads();
    if (wave == 0) {
        int run = 0;
        for (int k0 = 0; k0 < nbins; k0 += 64) {
            const int k = k0 + lane;
            int v = k < nbins ? bpre[k + 1] : 0;
            for (int off = 1; off < 64; off <<= 1) {
                const int u = __shfl_up(v, off);
                if (lane >= off) v += u;
            }
            if (k < nbins) bpre[k + 1] = run + v;
            run += __shfl(v, 63);
        }
        if (lane == 0) bpre[0] = 0;
    }
    __syncthreads();
    const int total = bpre[nbins];
    const int c_lo = seg * kAccSeg, c_hi = min(total, c_lo + kAccSeg);
    if (c_lo >= c_hi) return;  // uniform: nothing within reach, the grid holds 0 here
    // voxel centres (origin is the centre of voxel 0)
    const float cx = ox + (float)x * vs, cy = oy + (float)y * vs;
    const float cz0 = oz + (float)z * vs, cz1 = oz + (float)(z + 4) * vs;
    const AccBox tile = acc_box(bx * kAT, by * kAT, bz * kAT, kAT, kAT, kAT, ox, oy, oz, vs);
    const AccBox mine = acc_box(sx, sy, sz, 4, 4, kAT, ox, oy, oz, vs);
    const float lim = trunc * trunc * (1.0f + 0x1p-20f);
    const unsigned long long below = (1ull << lane) - 1ull;
    float4 *const wl = wsp + wave * kAccThreads;
    float best0 = __builtin_huge_valf(), best1 = __builtin_huge_valf();
    for (int c0 = c_lo; c0 < c_hi; c0 += kAccThreads) {
        const int c = c0 + (int)threadIdx.x;
        bool keep = false;
        float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (c < c_hi) {
            // the bin holding candidate c: last k with bpre[k] <= c
            int lo = 0, hi = nbins - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (bpre[mid] <= c) lo = mid;
                else hi = mid - 1;
            }
            q = binned[bstart[lo] + (c - bpre[lo])];
            keep = tile.d2(q) <= lim;
        }
        const unsigned long long m = __ballot(keep);
        __syncthreads();  // the previous chunk's reads of sp are done
        if (lane == 0) wave_n[wave] = __popcll(m);
        __syncthreads();
        int pos = __popcll(m & below), kept = 0;
        for (int w = 0; w < kAccThreads / 64; ++w) {
            const int cnt = wave_n[w];
            pos += w < wave ? cnt : 0;
            kept += cnt;
        }
        if (keep) sp[pos] = q;
        __syncthreads();
        // this wave's candidates: those within reach of its block
        int wn = 0;
        for (int i0 = 0; i0 < kept; i0 += 64) {
            const int i = i0 + lane;
            bool k2 = false;
            float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (i < kept) {
                p = sp[i];
                k2 = mine.d2(p) <= lim;
            }
            const unsigned long long m2 = __ballot(k2);
            if (k2) wl[wn + __popcll(m2 & below)] = p;
            wn += __popcll(m2);
        }
        __syncthreads();  // the lists are written
        for (int k = 0; k < wn; ++k) {
            const float4 p = wl[k];
            const float ex = cx - p.x, ey = cy - p.y;
            const float exy = ex * ex + ey * ey;
            const float ez0 = cz0 - p.z, ez1 = cz1 - p.z;
            best0 = fminf(best0, exy + ez0 * ez0);
            best1 = fminf(best1, exy + ez1 * ez1);
        }
    }
    if (x >= dx || y >= dy) return;
    // sqrt is monotone: the minimum of the distances
    const float v0 = 1.0f - fminf(sqrtf(best0), trunc) / trunc;
    const float v1 = 1.0f - fminf(sqrtf(best1), trunc) / trunc;
    float *const o0 = tdf + ((long)z * dy + y) * dx + x, *const o1 = o0 + 4L * dy * dx;
    if (total <= kAccSeg) {  // the tile's only segment: 0 is already there
        if (z < dz && v0 != 0.0f) *o0 = v0;
        if (z + 4 < dz && v1 != 0.0f) *o1 = v1;
    } else {
        // Segments meet in the grid by the largest value.  The value is a
        // non-increasing function of d^2 (each step is correctly rounded,
        // hence monotone), so the largest segment value is the value of the
        // smallest d^2; values lie in [0, 1], where the order of the float
        // bits as integers is the order of the floats.
        if (z < dz && v0 != 0.0f) atomicMax(reinterpret_cast<int *>(o0), __float_as_int(v0));
        if (z + 4 < dz && v1 != 0.0f) atomicMax(reinterpret_cast<int *>(o1), __float_as_int(v1));
    }
}

__global__ __launch_bounds__(kAccThreads) void acc_tdf_kernel(
    const float4 *__restrict__ binned, const int *__restrict__ offsets, int dx, int dy, int dz,
    int tx, int ty, int tz, int reach, float ox, float oy, float oz, float vs, float trunc,
    const int *__restrict__ item_count, int *__restrict__ next_item,
    const int2 *__restrict__ items, float *__restrict__ tdf) {
    __shared__ float4 sp[kAccThreads];
    __shared__ float4 wsp[kAccThreads / 64 * kAccThreads];
    __shared__ int bstart[kAccMaxBins];
    __shared__ int bpre[kAccMaxBins + 1];
    __shared__ int wave_n[kAccThreads / 64];
    __shared__ int item;
    // A resident grid takes the work items one at a time from a counter, so
    // heavy and light items balance as they come; every workgroup leaves
    // once the counter passes the (fixed) item count.
    const int n = *item_count;
    for (;;) {
        if (threadIdx.x == 0) item = atomicAdd(next_item, 1);
        __syncthreads();
        const int it = item;
        if (it >= n) break;  // uniform
        const int2 e = items[it];
        acc_tile(binned, offsets, e.x, e.y, dx, dy, dz, tx, ty, tz, reach, ox, oy, oz, vs, trunc,
                 sp, wsp, bstart, bpre, wave_n, tdf);
        __syncthreads();  // the LDS tables and `item` are reused
    }
}

int check_dims(int dx, int dy, int dz) {
    TDM_CHECK(dx > 0 && dy > 0 && dz > 0, TDM_ERR_INVALID_ARG, "bad grid %dx%dx%d", dx, dy, dz);
    TDM_CHECK((double)dx * dy * dz <= 2147483647.0, TDM_ERR_INVALID_ARG,
              "grid of %dx%dx%d exceeds 2^31 voxels", dx, dy, dz);
    return TDM_OK;
}

int search_radius(float vs, float trunc, int *r) {
    TDM_CHECK(vs > 0 && trunc > 0, TDM_ERR_INVALID_ARG, "voxel size and truncation must be > 0");
    volatile float q = trunc / vs;  // core/demo.cu:36, float division then round
    int rr = (int)std::round(q);
    TDM_CHECK(rr >= 1 && rr <= kMaxR, TDM_ERR_UNSUPPORTED,
              "search radius %d outside [1, %d] voxels", rr, kMaxR);
    *r = rr;
    return TDM_OK;
}

template <int R>
void launch_tile(const unsigned *bits, int dx, int dy, int dz, const Lut &lut, float *tdf,
                 const unsigned char *flags, hipStream_t s) {
    dim3 grid((dx + kTX - 1) / kTX, (dy + kTY - 1) / kTY, (dz + kTZ - 1) / kTZ);
    tdf_tile_kernel<R><<<grid, kTdfThreadsOne, 0, s>>>(bits, dx, dy, dz, lut, tdf, flags);
}

long num_tiles(int dx, int dy, int dz) {
    return (long)((dx + kTX - 1) / kTX) * ((dy + kTY - 1) / kTY) * ((dz + kTZ - 1) / kTZ);
}

// TDF from a bit grid in scratch into a grid that bits_scratch zeroed (the
// tile kernel writes only tiles whose halo holds a point).  `flags` (num_tiles
// bytes of scratch) holds the per-tile occupancy when flags_ready, else it is
// computed here from the bits.
// The exact TDF goes to `work` (already zeroed); in reference decode `work`
// is a separate buffer and tdf receives the gathered result, else work == tdf.
int tdf_from_bits(const unsigned *bits, int dx, int dy, int dz, int r, int decode, float *work,
                  float *tdf, unsigned char *flags, bool flags_ready, hipStream_t s) {
    Lut lut = make_lut(r);
    if (!flags_ready) {
        const long nt = num_tiles(dx, dy, dz);
        tile_flags_kernel<<<(unsigned)((nt + 3) / 4), 256, 0, s>>>(bits, dx, dy, dz, flags);
        TDM_LAUNCHED();
    }
    switch (r) {
    case 1: launch_tile<1>(bits, dx, dy, dz, lut, work, flags, s); break;
    case 2: launch_tile<2>(bits, dx, dy, dz, lut, work, flags, s); break;
    case 3: launch_tile<3>(bits, dx, dy, dz, lut, work, flags, s); break;
    case 4: launch_tile<4>(bits, dx, dy, dz, lut, work, flags, s); break;
    case 5: launch_tile<5>(bits, dx, dy, dz, lut, work, flags, s); break;
    case 6: launch_tile<6>(bits, dx, dy, dz, lut, work, flags, s); break;
    case 7: launch_tile<7>(bits, dx, dy, dz, lut, work, flags, s); break;
    case 8: launch_tile<8>(bits, dx, dy, dz, lut, work, flags, s); break;
    default: return tdm::fail(TDM_ERR_UNSUPPORTED, "radius %d", r);
    }
    TDM_LAUNCHED();
    if (decode == TDM_DECODE_REFERENCE) {
        // at most (dx/256 + 1) * dy * dz / kRefRows < 2^31 workgroups
        tdf_refgather_kernel<<<refgather_blocks(dx, dy, dz), 256, 0, s>>>(bits, dx, dy, dz, r,
                                                                          lut, work, tdf);
        TDM_LAUNCHED();
    }
    return TDM_OK;
}

// Where the exact TDF is computed: tdf itself, or for reference decode a
// dx*dy*dz-float buffer in the context's scratch slot 2 that the gather then
// reads.
int tdf_work(tdm_ctx *ctx, int decode, size_t nvox, float *tdf, hipStream_t s, float **work) {
    *work = tdf;
    if (decode != TDM_DECODE_REFERENCE) return TDM_OK;
    void *p = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 2, sizeof(float) * nvox, &p, s));
    *work = static_cast<float *>(p);
    return TDM_OK;
}

// Bytes of a bit grid plus its per-tile flags.
size_t bits_bytes(int dx, int dy, int dz) {
    return (size_t)row_words(dx) * dy * dz * sizeof(unsigned) + num_tiles(dx, dy, dz);
}

// Bit grid and, after it, the per-tile flags at p (bits_bytes, 16-byte
// aligned); one launch clears both and zeroes the output grid tdf (dx*dy*dz
// floats; none if null).
int clear_bits(void *p, int dx, int dy, int dz, float *tdf, unsigned **bits, size_t *words,
               unsigned char **flags, hipStream_t s) {
    *words = (size_t)row_words(dx) * dy * dz;
    *bits = static_cast<unsigned *>(p);
    *flags = reinterpret_cast<unsigned char *>(*bits + *words);
    clear_regions(tdf, (size_t)dx * dy * dz, static_cast<unsigned char *>(p), bits_bytes(dx, dy, dz),
                  s);
    TDM_LAUNCHED();
    return TDM_OK;
}

// The same with the bit grid in the context's scratch slot 0.
int bits_scratch(tdm_ctx *ctx, int dx, int dy, int dz, float *tdf, unsigned **bits,
                 size_t *words, unsigned char **flags, hipStream_t s) {
    void *p = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 0, bits_bytes(dx, dy, dz), &p, s));
    return clear_bits(p, dx, dy, dz, tdf, bits, words, flags, s);
}

// core/demo.cu:127-158 from points, given the bit grid region p (bits_bytes):
// clear it, scatter the points (the deduplicating scatter zeroes `work` in
// the same launch), then the tile pass; `work` and `tdf` as in tdf_from_bits.
int points_to_tdf(void *p, const float *pts, int num_pts, const int dims[3], const float origin[3],
                  float voxel_size, int r, int decode, float *work, float *tdf, hipStream_t s) {
    unsigned *bits = nullptr;
    size_t words = 0;
    unsigned char *flags = nullptr;
    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    const bool dedup = num_pts > 0 && (size_t)row_words(dims[0]) * dims[1] * dims[2] < kScatEmpty;
    TDM_TRY(clear_bits(p, dims[0], dims[1], dims[2], dedup ? nullptr : work, &bits, &words, &flags,
                       s));
    if (dedup) {
        const long per_block = (long)kScatThreads * kScatPts;
        const unsigned nscat = (unsigned)((num_pts + per_block - 1) / per_block);
        const ZeroJob z = zero_job(work, nvox);
        const unsigned nfill = (unsigned)std::min<size_t>(z.n4 / kScatThreads + 1, 512);
        scatter_bits_dedup_kernel<<<nscat + nfill, kScatThreads, 0, s>>>(
            pts, num_pts, dims[0], dims[1], dims[2], origin[0], origin[1], origin[2], voxel_size,
            bits, flags, z, nscat);
        TDM_LAUNCHED();
    } else if (num_pts > 0) {
        scatter_bits_kernel<<<tdm::blocks_for(num_pts, 256), 256, 0, s>>>(
            pts, num_pts, dims[0], dims[1], dims[2], origin[0], origin[1], origin[2], voxel_size,
            bits, flags);
        TDM_LAUNCHED();
    }
    return tdf_from_bits(bits, dims[0], dims[1], dims[2], r, decode, work, tdf, flags, true, s);
}

int check_points_args(const float *pts, int num_pts, const int dims[3], const float origin[3],
                      float *tdf) {
    TDM_CHECK((pts || num_pts == 0) && dims && origin && tdf, TDM_ERR_INVALID_ARG,
              "null pointer argument");
    TDM_CHECK(num_pts >= 0, TDM_ERR_INVALID_ARG, "negative point count");
    return check_dims(dims[0], dims[1], dims[2]);
}

constexpr int kMaxDevices = 64;
std::mutex g_static_mu;
void *g_static_ptr[kMaxDevices];

// The address of `scratch` on the current device, looked up once per device
// (with that device current) and kept.
int static_scratch_ptr(void **out) {
    int dev = 0;
    TDM_HIP(hipGetDevice(&dev));
    TDM_CHECK(dev >= 0 && dev < kMaxDevices, TDM_ERR_UNSUPPORTED, "device %d beyond %d", dev,
              kMaxDevices);
    std::lock_guard<std::mutex> lock(g_static_mu);
    if (!g_static_ptr[dev]) {
        void *p = nullptr;
        TDM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TDM_CHECK(p, TDM_ERR_HIP, "no device address for the static scratch");
        g_static_ptr[dev] = p;
    }
    *out = g_static_ptr[dev];
    return TDM_OK;
}

}  // namespace

namespace tdm {

int grid_from_box(const float box[6], float voxel_size, int padding, int dims[3],
                  float origin[3]) {
    // core/demo.cu:116-122, float arithmetic on the host.
    for (int a = 0; a < 3; ++a) {
        TDM_CHECK(std::isfinite(box[a]) && std::isfinite(box[3 + a]), TDM_ERR_INVALID_ARG,
                  "non-finite point coordinates");
        volatile float ext = (box[3 + a] - box[a]) / voxel_size;
        double dim = std::round((float)ext) + 1.0 + padding * 2.0;
        TDM_CHECK(dim < 2147483647.0, TDM_ERR_INVALID_ARG, "grid extent too large");
        dims[a] = (int)dim;
        volatile float padw = padding * voxel_size;
        volatile float o = box[a] - padw;
        origin[a] = o + voxel_size / 2;
    }
    return TDM_OK;
}

}  // namespace tdm

extern "C" {

int tdm_grid_params(tdm_ctx *ctx, const float *pts, int num_pts, float voxel_size, int padding,
                    int dims_out[3], float origin_out[3], void *stream) {
    tdm::clear_error();
    TDM_CHECK(pts && dims_out && origin_out, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_CHECK(num_pts > 0, TDM_ERR_INVALID_ARG, "empty point cloud");
    TDM_CHECK(voxel_size > 0 && padding >= 0, TDM_ERR_INVALID_ARG, "bad voxel size / padding");
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    hipStream_t s = tdm::as_stream(stream);
    const int nblocks = (int)std::min<long>(((long)num_pts + 255) / 256, 1024);
    void *area = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 1, sizeof(float) * 6 * (nblocks + 1), &area, s));
    float *part = static_cast<float *>(area);
    float *res = part + 6 * nblocks;
    bbox_partial_kernel<<<nblocks, 256, 0, s>>>(pts, num_pts, part);
    TDM_LAUNCHED();
    bbox_final_kernel<<<1, 256, 0, s>>>(part, nblocks, res);
    TDM_LAUNCHED();
    float box[6];
    TDM_HIP(hipMemcpyAsync(box, res, sizeof box, hipMemcpyDeviceToHost, s));
    TDM_HIP(hipStreamSynchronize(s));
    return tdm::grid_from_box(box, voxel_size, padding, dims_out, origin_out);
}

int tdm_occupancy(tdm_ctx *ctx, const float *pts, int num_pts, const int dims[3],
                  const float origin[3], float voxel_size, float *occ, void *stream) {
    tdm::clear_error();
    TDM_CHECK(pts && dims && origin && occ, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_CHECK(num_pts >= 0 && voxel_size > 0, TDM_ERR_INVALID_ARG, "bad arguments");
    TDM_TRY(check_dims(dims[0], dims[1], dims[2]));
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    hipStream_t s = tdm::as_stream(stream);
    size_t n = (size_t)dims[0] * dims[1] * dims[2];
    unsigned blocks = (unsigned)std::min<size_t>(tdm::blocks_for(n, 256), 16384);
    fill_u32_kernel<<<blocks, 256, 0, s>>>(reinterpret_cast<unsigned *>(occ), n, 0u);
    TDM_LAUNCHED();
    if (num_pts == 0) return TDM_OK;
    scatter_occ_kernel<<<tdm::blocks_for(num_pts, 256), 256, 0, s>>>(
        pts, num_pts, dims[0], dims[1], dims[2], origin[0], origin[1], origin[2], voxel_size, occ);
    TDM_LAUNCHED();
    return TDM_OK;
}

int tdm_compute_tdf(tdm_ctx *ctx, const float *occ, float *tdf, int dim_x, int dim_y, int dim_z,
                    float voxel_size, float trunc_margin, int decode, void *stream) {
    tdm::clear_error();
    TDM_CHECK(occ && tdf, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_CHECK(decode == TDM_DECODE_EXACT || decode == TDM_DECODE_REFERENCE, TDM_ERR_INVALID_ARG,
              "bad decode mode %d", decode);
    TDM_TRY(check_dims(dim_x, dim_y, dim_z));
    int r = 0;
    TDM_TRY(search_radius(voxel_size, trunc_margin, &r));
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    hipStream_t s = tdm::as_stream(stream);
    unsigned *bits = nullptr;
    size_t words = 0;
    unsigned char *flags = nullptr;
    TDM_TRY(bits_scratch(ctx, dim_x, dim_y, dim_z, nullptr, &bits, &words, &flags, s));
    const long rows = (long)dim_y * dim_z;
    const long pairs = (row_words(dim_x) + 1) / 2;
    const long waves = rows * pairs;
    pack_bits_kernel<<<tdm::blocks_for(waves * 64, 256), 256, 0, s>>>(occ, dim_x, rows, bits);
    TDM_LAUNCHED();
    // zeroed after the packing, so occ and tdf may be one buffer
    const size_t nvox = (size_t)dim_x * dim_y * dim_z;
    float *work = nullptr;
    TDM_TRY(tdf_work(ctx, decode, nvox, tdf, s, &work));
    clear_regions(work, nvox, nullptr, 0, s);
    TDM_LAUNCHED();
    return tdf_from_bits(bits, dim_x, dim_y, dim_z, r, decode, work, tdf, flags, false, s);
}

int tdm_tdf_from_points(tdm_ctx *ctx, const float *pts, int num_pts, const int dims[3],
                        const float origin[3], float voxel_size, float trunc_margin, int decode,
                        float *tdf, void *stream) {
    tdm::clear_error();
    TDM_TRY(check_points_args(pts, num_pts, dims, origin, tdf));
    TDM_CHECK(decode == TDM_DECODE_EXACT || decode == TDM_DECODE_REFERENCE, TDM_ERR_INVALID_ARG,
              "bad decode mode %d", decode);
    int r = 0;
    TDM_TRY(search_radius(voxel_size, trunc_margin, &r));
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    hipStream_t s = tdm::as_stream(stream);
    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    float *work = nullptr;
    TDM_TRY(tdf_work(ctx, decode, nvox, tdf, s, &work));
    void *p = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 0, bits_bytes(dims[0], dims[1], dims[2]), &p, s));
    return points_to_tdf(p, pts, num_pts, dims, origin, voxel_size, r, decode, work, tdf, s);
}

int tdm_tdf_from_points_noalloc(const float *pts, int num_pts, const int dims[3],
                                const float origin[3], float voxel_size, float trunc_margin,
                                int decode, float *tdf, void *stream) {
    tdm::clear_error();
    TDM_TRY(check_points_args(pts, num_pts, dims, origin, tdf));
    TDM_CHECK(decode == TDM_DECODE_EXACT || decode == TDM_DECODE_REFERENCE, TDM_ERR_INVALID_ARG,
              "bad decode mode %d", decode);
    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    // Below 2^24 voxels the float32 decode of core/demo.cu:32-34 is exact, so
    // both modes are the same computation; above it the reference decode
    // needs a whole-grid work buffer, which this entry point does not have.
    TDM_CHECK(decode == TDM_DECODE_EXACT || nvox <= ((size_t)1 << 24), TDM_ERR_UNSUPPORTED,
              "reference decode of a %zu-voxel grid needs tdm_tdf_from_points", nvox);
    const size_t need = bits_bytes(dims[0], dims[1], dims[2]);
    TDM_CHECK(need <= TDM_TDF_NOALLOC_SCRATCH_BYTES, TDM_ERR_UNSUPPORTED,
              "grid %d x %d x %d needs %zu bytes of bit grid, above the %d-byte static scratch",
              dims[0], dims[1], dims[2], need, TDM_TDF_NOALLOC_SCRATCH_BYTES);
    int r = 0;
    TDM_TRY(search_radius(voxel_size, trunc_margin, &r));
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return tdm::fail(TDM_ERR_NO_DEVICE, "no HIP device available");
    }
    (void)hipGetLastError();  // an earlier caller's error is not this call's
    void *p = nullptr;
    TDM_TRY(static_scratch_ptr(&p));
    return points_to_tdf(p, pts, num_pts, dims, origin, voxel_size, r, TDM_DECODE_EXACT, tdf, tdf,
                         tdm::as_stream(stream));
}

int tdm_keypoints_to_grid(const float *kp, int num, const float origin[3], float voxel_size,
                          int *grid_out) {
    tdm::clear_error();
    TDM_CHECK(num >= 0, TDM_ERR_INVALID_ARG, "negative count");
    TDM_CHECK(num == 0 || (kp && origin && grid_out), TDM_ERR_INVALID_ARG, "null pointer");
    TDM_CHECK(voxel_size > 0, TDM_ERR_INVALID_ARG, "bad voxel size");
    for (long i = 0; i < num; ++i)
        for (int a = 0; a < 3; ++a) {
            volatile float q = (kp[i * 3 + a] - origin[a]) / voxel_size;
            float r = std::round((float)q);
            TDM_CHECK(std::fabs(r) < 2.0e9f, TDM_ERR_INVALID_ARG, "keypoint %ld out of range", i);
            grid_out[i * 3 + a] = (int)r;
        }
    return TDM_OK;
}

int tdm_extract_patches(tdm_ctx *ctx, const float *tdf, int dim_x, int dim_y, int dim_z,
                        const int *kp_grid, int num, int patch_size, float *patches, void *stream) {
    tdm::clear_error();
    TDM_CHECK(num >= 0 && patch_size > 0 && patch_size <= 256, TDM_ERR_INVALID_ARG, "bad sizes");
    TDM_CHECK(num == 0 || (tdf && kp_grid && patches), TDM_ERR_INVALID_ARG, "null pointer");
    TDM_TRY(check_dims(dim_x, dim_y, dim_z));
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    if (num == 0) return TDM_OK;
    long total = (long)patch_size * patch_size * patch_size * num;
    unsigned blocks = (unsigned)std::min<long>((total + 255) / 256, 65536);
    extract_patches_kernel<<<blocks, 256, 0, tdm::as_stream(stream)>>>(
        tdf, dim_x, dim_y, dim_z, kp_grid, num, patch_size, patches);
    TDM_LAUNCHED();
    return TDM_OK;
}

int tdm_tdf_accurate(tdm_ctx *ctx, const float *pts, int num_pts, const int dims[3],
                     const float origin[3], float voxel_size, float trunc_margin, float *tdf,
                     void *stream) {
    tdm::clear_error();
    TDM_CHECK((pts || num_pts == 0) && dims && origin && tdf, TDM_ERR_INVALID_ARG,
              "null pointer argument");
    TDM_CHECK(num_pts >= 0, TDM_ERR_INVALID_ARG, "negative point count");
    TDM_CHECK(voxel_size > 0 && trunc_margin > 0, TDM_ERR_INVALID_ARG, "bad voxel size / margin");
    TDM_TRY(check_dims(dims[0], dims[1], dims[2]));
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    hipStream_t s = tdm::as_stream(stream);
    const int tx = (dims[0] + kAT - 1) / kAT, ty = (dims[1] + kAT - 1) / kAT,
              tz = (dims[2] + kAT - 1) / kAT;
    const long nt = (long)tx * ty * tz;
    // a point within trunc of a voxel centre is binned at most
    // trunc/vs + 1/2 voxels away from that voxel
    const double reach_vox = (double)trunc_margin / voxel_size + 0.5;
    const int reach = (int)std::ceil(reach_vox / kAT);
    TDM_CHECK(reach <= 4, TDM_ERR_UNSUPPORTED, "truncation of %g voxels is too wide",
              (double)trunc_margin / voxel_size);
    // points binned up to this many voxels outside the grid (>= reach_vox)
    const int margin = reach * kAT;
    const long nb = (nt + 1023) / 1024;  // scan blocks
    // work items: at most one per tile plus one per kAccSeg candidates, and
    // a point is a candidate of at most (2 reach + 1)^3 tiles
    const long side = 2 * reach + 1;
    const size_t max_items = (size_t)nt + (size_t)(side * side * side) * (size_t)num_pts / kAccSeg + 1;
    const size_t bytes = sizeof(int) * (2 * (size_t)nt + 3 + nb) + 32 +
                         sizeof(float4) * (size_t)num_pts + sizeof(int2) * max_items;
    void *area = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 0, bytes, &area, s));
    int *counts = static_cast<int *>(area);
    int *item_count = counts + nt;  // 1 entry, zeroed with the counts
    int *next_item = item_count + 1;  // 1 entry, zeroed with the counts
    int *offsets = next_item + 1;     // nt + 1 entries
    int *bsum = offsets + nt + 1;     // nb entries
    float4 *binned = reinterpret_cast<float4 *>(
        (reinterpret_cast<uintptr_t>(bsum + nb) + 15) & ~uintptr_t(15));
    int2 *items = reinterpret_cast<int2 *>(binned + num_pts);  // max_items entries
    const unsigned fb = (unsigned)std::min<long>((nt + 255) / 256, 16384);
    // one launch zeroes the output grid (full lines), the bin counts and the
    // two work-item counters
    clear_regions(tdf, (size_t)dims[0] * dims[1] * dims[2], reinterpret_cast<unsigned char *>(counts),
                  sizeof(int) * ((size_t)nt + 2), s);
    TDM_LAUNCHED();
    if (num_pts > 0) {
        acc_count_kernel<<<(unsigned)((num_pts + kBinThreads * kBinPts - 1) / (kBinThreads * kBinPts)),
                           kBinThreads, 0, s>>>(
            pts, num_pts, dims[0], dims[1], dims[2], origin[0], origin[1], origin[2], voxel_size,
            margin, tx, ty, counts);
        TDM_LAUNCHED();
    }
    acc_blocksum_kernel<<<(unsigned)nb, 1024, 0, s>>>(counts, nt, bsum);
    TDM_LAUNCHED();
    acc_offsets_kernel<<<(unsigned)nb, 1024, 0, s>>>(counts, nt, bsum, offsets);
    TDM_LAUNCHED();
    fill_u32_kernel<<<fb, 256, 0, s>>>(reinterpret_cast<unsigned *>(counts), (size_t)nt, 0u);
    TDM_LAUNCHED();
    if (num_pts > 0) {
        acc_fill_kernel<<<(unsigned)((num_pts + kBinThreads * kBinPts - 1) / (kBinThreads * kBinPts)),
                          kBinThreads, 0, s>>>(
            pts, num_pts, dims[0], dims[1], dims[2], origin[0], origin[1], origin[2], voxel_size,
            margin, tx, ty, offsets, counts, binned);
        TDM_LAUNCHED();
    }
    TDM_CHECK(nt < (1L << 31), TDM_ERR_INVALID_ARG, "grid too large");
    acc_near_kernel<<<(unsigned)((nt + 255) / 256), 256, 0, s>>>(offsets, tx, ty, tz, reach,
                                                                 item_count, items);
    TDM_LAUNCHED();
    // six workgroups per CU fit (LDS), on the device's CUs
    int cus = 0;
    TDM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const long grid = std::min<long>((long)std::max(cus, 1) * 6, (long)max_items);
    acc_tdf_kernel<<<(unsigned)grid, kAccThreads, 0, s>>>(
        binned, offsets, dims[0], dims[1], dims[2], tx, ty, tz, reach, origin[0], origin[1],
        origin[2], voxel_size, trunc_margin, item_count, next_item, items, tdf);
    TDM_LAUNCHED();
    return TDM_OK;
}

struct tdm_tdf_plan {
    int device = 0;
    int num = 0;
    int r = 0;
    int decode = TDM_DECODE_EXACT;
    float vs = 0.0f;
    long total_pts = 0, total_tiles = 0;
    size_t total_words = 0;
    CloudDev *table = nullptr;   // device
    unsigned *bits = nullptr;    // device, all clouds' bit grids, then the tile flags
    unsigned char *flags = nullptr;
    long total_vox = 0;
    std::vector<CloudDev> host;  // the same table on the host
    // reference decode: the exact TDF of one cloud above 2^24 voxels at a
    // time (the largest such cloud's size), gathered from by the decode
    float *ref_exact = nullptr;
};

int tdm_tdf_plan_create(tdm_ctx *ctx, const tdm_cloud_ref *clouds, int num, float voxel_size,
                        float trunc_margin, int decode, tdm_tdf_plan **out) {
    tdm::clear_error();
    TDM_CHECK(clouds && out && num >= 1, TDM_ERR_INVALID_ARG, "bad arguments");
    TDM_CHECK(decode == TDM_DECODE_EXACT || decode == TDM_DECODE_REFERENCE, TDM_ERR_INVALID_ARG,
              "bad decode mode %d", decode);
    int r = 0;
    TDM_TRY(search_radius(voxel_size, trunc_margin, &r));
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    auto plan = std::make_unique<tdm_tdf_plan>();
    plan->device = ctx->device;
    plan->num = num;
    plan->r = r;
    plan->decode = decode;
    plan->vs = voxel_size;
    plan->host.resize(num);
    long pts = 0, tiles = 0, vox = 0;
    size_t words = 0;
    for (int c = 0; c < num; ++c) {
        const tdm_cloud_ref &in = clouds[c];
        TDM_CHECK(in.tdf && in.num_pts >= 0 && (in.pts || in.num_pts == 0), TDM_ERR_INVALID_ARG,
                  "cloud %d: bad pointers", c);
        TDM_TRY(check_dims(in.dim_x, in.dim_y, in.dim_z));
        CloudDev &d = plan->host[c];
        d.pts = in.pts;
        d.tdf = in.tdf;
        d.n = in.num_pts;
        d.dims[0] = in.dim_x;
        d.dims[1] = in.dim_y;
        d.dims[2] = in.dim_z;
        d.origin[0] = in.origin_x;
        d.origin[1] = in.origin_y;
        d.origin[2] = in.origin_z;
        d.tiles[0] = (in.dim_x + kTX - 1) / kTX;
        d.tiles[1] = (in.dim_y + kTY - 1) / kTY;
        d.tiles[2] = (in.dim_z + kTZ - 1) / kTZ;
        d.pt_off = pts;
        d.word_off = (long)words;
        d.tile_off = tiles;
        d.vox_off = vox;
        vox += (long)in.dim_x * in.dim_y * in.dim_z;
        pts += in.num_pts;
        tiles += (long)d.tiles[0] * d.tiles[1] * d.tiles[2];
        words += (size_t)row_words(in.dim_x) * in.dim_y * in.dim_z;
    }
    TDM_CHECK(tiles < (1L << 31), TDM_ERR_INVALID_ARG, "batch too large");
    plan->total_pts = pts;
    plan->total_tiles = tiles;
    plan->total_words = words;
    plan->total_vox = vox;
    // the tile flags (one byte per tile) follow the bits, padded to a word
    const size_t bit_bytes = (words + ((size_t)tiles + 3) / 4) * sizeof(unsigned);
    hipError_t e = hipMalloc((void **)&plan->table, sizeof(CloudDev) * num);
    if (e == hipSuccess) e = hipMalloc((void **)&plan->bits, bit_bytes);
    if (e == hipSuccess)
        e = hipMemcpy(plan->table, plan->host.data(), sizeof(CloudDev) * num,
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        if (plan->table) (void)hipFree(plan->table);
        if (plan->bits) (void)hipFree(plan->bits);
        return tdm::fail(e == hipErrorOutOfMemory ? TDM_ERR_NOMEM : TDM_ERR_HIP,
                         "TDF plan of %zu bytes: %s", bit_bytes, hipGetErrorString(e));
    }
    plan->flags = reinterpret_cast<unsigned char *>(plan->bits + words);
    if (decode == TDM_DECODE_REFERENCE) {
        long big = 0;  // voxels of the largest cloud whose float32 decode can differ
        for (const CloudDev &d : plan->host) {
            const long n = (long)d.dims[0] * d.dims[1] * d.dims[2];
            if (n > (1L << 24)) big = std::max(big, n);
        }
        if (big > 0 && (e = hipMalloc((void **)&plan->ref_exact, sizeof(float) * big)) != hipSuccess) {
            (void)hipGetLastError();
            (void)hipFree(plan->table);
            (void)hipFree(plan->bits);
            return tdm::fail(e == hipErrorOutOfMemory ? TDM_ERR_NOMEM : TDM_ERR_HIP,
                             "TDF plan: %ld voxels of reference-decode scratch: %s", big,
                             hipGetErrorString(e));
        }
    }
    *out = plan.release();
    return TDM_OK;
}

int tdm_tdf_plan_run(tdm_tdf_plan *plan, void *stream) {
    tdm::clear_error();
    TDM_CHECK(plan, TDM_ERR_INVALID_ARG, "NULL plan");
    TDM_HIP(hipSetDevice(plan->device));
    hipStream_t s = tdm::as_stream(stream);
    // the grids are zeroed first; the tile kernel writes only tiles whose halo
    // holds a point
    const long chunk = 65536;
    zero_clouds_kernel<<<(unsigned)((plan->total_vox + chunk - 1) / chunk), 256, 0, s>>>(
        plan->table, plan->num, plan->total_vox, chunk);
    TDM_LAUNCHED();
    const size_t cw = plan->total_words + ((size_t)plan->total_tiles + 3) / 4;
    const unsigned fb = (unsigned)std::min<size_t>(tdm::blocks_for(cw, 256), 16384);
    fill_u32_kernel<<<fb, 256, 0, s>>>(plan->bits, cw, 0u);
    TDM_LAUNCHED();
    if (plan->total_pts > 0 && plan->total_words < (size_t)kScatEmpty) {
        const long per_block = (long)kScatThreads * kScatPts;
        scatter_bits_batch_dedup_kernel<<<(unsigned)((plan->total_pts + per_block - 1) / per_block),
                                          kScatThreads, 0, s>>>(plan->table, plan->num,
                                                                plan->total_pts, plan->vs,
                                                                plan->bits, plan->flags);
        TDM_LAUNCHED();
    } else if (plan->total_pts > 0) {
        const unsigned sb = (unsigned)std::min<long>((plan->total_pts + 255) / 256, 65536);
        scatter_bits_batch_kernel<<<sb, 256, 0, s>>>(plan->table, plan->num, plan->total_pts,
                                                      plan->vs, plan->bits, plan->flags);
        TDM_LAUNCHED();
    }
    const Lut lut = make_lut(plan->r);
    const unsigned tb = (unsigned)plan->total_tiles;
    switch (plan->r) {
#define TDM_TB(R) \
    case R: tdf_tile_batch_kernel<R><<<tb, kTdfThreads, 0, s>>>(plan->table, plan->num, plan->bits, lut, plan->flags); break;
        TDM_TB(1) TDM_TB(2) TDM_TB(3) TDM_TB(4) TDM_TB(5) TDM_TB(6) TDM_TB(7) TDM_TB(8)
#undef TDM_TB
    default: return tdm::fail(TDM_ERR_UNSUPPORTED, "radius %d", plan->r);
    }
    TDM_LAUNCHED();
    // Reference decode.  Up to 2^24 voxels the float32 decode of every index
    // is exact (all its operands are integers below 2^24, and the quotient
    // v / (dx*dy) = k - r/(dx*dy) lies more than half an ulp below the next
    // integer k), so those grids are done.  A larger grid's exact TDF is
    // copied to the plan's scratch and each voxel takes the value at its
    // decoded index from there, as the single-cloud path does.
    if (plan->decode == TDM_DECODE_REFERENCE)
        for (const CloudDev &d : plan->host) {
            const int dx = d.dims[0], dy = d.dims[1], dz = d.dims[2];
            const long n = (long)dx * dy * dz;
            if (n <= (1L << 24)) continue;
            const unsigned cb = (unsigned)std::min<long>((n / 4 + 255) / 256, 16384);
            copy_floats_kernel<<<cb, 256, 0, s>>>(d.tdf, plan->ref_exact, n);
            TDM_LAUNCHED();
            tdf_refgather_kernel<<<refgather_blocks(dx, dy, dz), 256, 0, s>>>(
                plan->bits + d.word_off, dx, dy, dz, plan->r, lut, plan->ref_exact, d.tdf);
            TDM_LAUNCHED();
        }
    return TDM_OK;
}

int tdm_tdf_plan_destroy(tdm_tdf_plan *plan) {
    tdm::clear_error();
    if (!plan) return TDM_OK;
    (void)hipSetDevice(plan->device);
    (void)hipDeviceSynchronize();
    if (plan->table) (void)hipFree(plan->table);
    if (plan->bits) (void)hipFree(plan->bits);
    if (plan->ref_exact) (void)hipFree(plan->ref_exact);
    delete plan;
    return TDM_OK;
}

}  // extern "C"
