// Alignment ratios of a registered fragment pair on gfx950
// (evaluation/geometric-registration/register2Fragments.m:68-69, 90-97):
// fragment 2 is moved by the estimated [R|t], then ratio 1 is the fraction
// of fragment-1 points whose nearest moved fragment-2 point lies closer than
// 0.05 m, and ratio 2 the same with the roles swapped.
//
// Only "is any target point within the radius" matters, so the 1-NN search
// of multiQueryKNNSearchImpl becomes a fixed-radius test: the target points
// are counting-sorted into a uniform grid of cells a little wider than the
// radius, and each query scans the 27 cells around its own, stopping at the
// first target within reach.  The transform runs in float with R and t
// rounded to float first (MATLAB evaluates the double matrix times the single
// point cloud in single), each coordinate summed left to right without
// contraction; a distance is sqrtf((dx*dx + dy*dy) + dz*dz) in float.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;
constexpr long kMaxCells = 1L << 22;  // wider cells beyond this

struct Box {
    float lo[3], hi[3];
};

// x1 = R x2 + t in float (rt: row-major 3x4 rounded to float).
__global__ __launch_bounds__(kThreads) void move_points_kernel(const float *__restrict__ p, int n,
                                                               float r00, float r01, float r02,
                                                               float r10, float r11, float r12,
                                                               float r20, float r21, float r22,
                                                               float t0, float t1, float t2,
                                                               float *__restrict__ out) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
    out[3 * i + 0] = ((r00 * x + r01 * y) + r02 * z) + t0;
    out[3 * i + 1] = ((r10 * x + r11 * y) + r12 * z) + t1;
    out[3 * i + 2] = ((r20 * x + r21 * y) + r22 * z) + t2;
}

// Per-workgroup min/max of each axis into part[blockIdx.x] (6 floats).
__global__ __launch_bounds__(kThreads) void box_partial_kernel(const float *__restrict__ p, int n,
                                                               float *__restrict__ part) {
    __shared__ float red[6][kThreads];
    float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * kThreads + threadIdx.x; i < n; i += gridDim.x * kThreads)
        for (int a = 0; a < 3; ++a) {
            v[a] = fminf(v[a], p[3 * i + a]);
            v[3 + a] = fmaxf(v[3 + a], p[3 * i + a]);
        }
    for (int a = 0; a < 6; ++a) red[a][threadIdx.x] = v[a];
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int a = 0; a < 6; ++a)
                red[a][threadIdx.x] = a < 3 ? fminf(red[a][threadIdx.x], red[a][threadIdx.x + s])
                                            : fmaxf(red[a][threadIdx.x], red[a][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

struct Grid {
    float lo[3];
    float inv;  // 1 / cell edge
    int dims[3];
};

__device__ inline int cell_axis(float v, float lo, float inv, int d) {
    const float c = floorf((v - lo) * inv);
    return (int)fminf(fmaxf(c, 0.0f), (float)(d - 1));
}

__device__ inline long cell_of(const Grid &g, float x, float y, float z) {
    const int cx = cell_axis(x, g.lo[0], g.inv, g.dims[0]);
    const int cy = cell_axis(y, g.lo[1], g.inv, g.dims[1]);
    const int cz = cell_axis(z, g.lo[2], g.inv, g.dims[2]);
    return ((long)cz * g.dims[1] + cy) * g.dims[0] + cx;
}

__global__ __launch_bounds__(kThreads) void fill_kernel(int *__restrict__ p, long n, int v) {
    for (long i = (long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long)gridDim.x * kThreads)
        p[i] = v;
}

__global__ __launch_bounds__(kThreads) void cell_count_kernel(const float *__restrict__ p, int n,
                                                              Grid g, int *__restrict__ count) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    atomicAdd(&count[cell_of(g, p[3 * i], p[3 * i + 1], p[3 * i + 2])], 1);
}

// Exclusive scan of count[0..n) into off[0..n] in three launches: sums of
// blocks of kScanPer cells, an exclusive scan of those sums in one workgroup
// (at most kMaxCells / kScanPer of them), then each block scans its cells
// from its prefix.
constexpr int kScanPer = 1024;  // cells per scan block
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kThreads) void scan_sums_kernel(const int *__restrict__ count, long n,
                                                             int *__restrict__ bsum) {
    __shared__ int red[kThreads];
    const long base = (long)blockIdx.x * kScanPer;
    int s = 0;
    for (int k = threadIdx.x; k < kScanPer; k += kThreads)
        if (base + k < n) s += count[base + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = red[0];
}

// In-place exclusive scan of bsum[0..nb), nb <= kScanThreads * 4.
__global__ __launch_bounds__(kScanThreads) void scan_block_sums_kernel(int *__restrict__ bsum,
                                                                       int nb) {
    __shared__ int part[kScanThreads];
    constexpr int PER = 4;
    const int base = threadIdx.x * PER;
    int local[PER];
    int s = 0;
    for (int k = 0; k < PER; ++k) {
        local[k] = base + k < nb ? bsum[base + k] : 0;
        s += local[k];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 1; w < kScanThreads; w <<= 1) {
        const int add = threadIdx.x >= w ? part[threadIdx.x - w] : 0;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    int run = part[threadIdx.x] - s;
    for (int k = 0; k < PER; ++k) {
        if (base + k < nb) bsum[base + k] = run;
        run += local[k];
    }
}

__global__ __launch_bounds__(kThreads) void scan_apply_kernel(const int *__restrict__ count, long n,
                                                              const int *__restrict__ bsum,
                                                              int *__restrict__ off,
                                                              int *__restrict__ cursor) {
    __shared__ int part[kThreads];
    // each thread owns kScanPer / kThreads consecutive cells
    constexpr int PER = kScanPer / kThreads;
    const long base = (long)blockIdx.x * kScanPer + (long)threadIdx.x * PER;
    int local[PER];
    int s = 0;
    for (int k = 0; k < PER; ++k) {
        local[k] = base + k < n ? count[base + k] : 0;
        s += local[k];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 1; w < kThreads; w <<= 1) {
        const int add = threadIdx.x >= w ? part[threadIdx.x - w] : 0;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    int run = bsum[blockIdx.x] + part[threadIdx.x] - s;
    for (int k = 0; k < PER; ++k) {
        if (base + k < n) {
            off[base + k] = run;
            cursor[base + k] = run;
        }
        run += local[k];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kThreads - 1) off[n] = run;
}

__global__ __launch_bounds__(kThreads) void cell_scatter_kernel(const float *__restrict__ p, int n,
                                                                Grid g, int *__restrict__ cursor,
                                                                float4 *__restrict__ sorted) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
    const int slot = atomicAdd(&cursor[cell_of(g, x, y, z)], 1);
    sorted[slot] = make_float4(x, y, z, 0.0f);
}

// hits += number of queries with a target closer than radius.
__global__ __launch_bounds__(kThreads) void within_radius_kernel(
    const float *__restrict__ q, int n, Grid g, const int *__restrict__ off,
    const float4 *__restrict__ sorted, float radius, unsigned long long *__restrict__ hits) {
    const int i = blockIdx.x * kThreads + threadIdx.x;
    bool hit = false;
    if (i < n) {
        const float x = q[3 * i], y = q[3 * i + 1], z = q[3 * i + 2];
        // the query's own cell, unclamped: a query more than one cell outside
        // the targets' box sees no cell
        int c[3];
        const float v[3] = {x, y, z};
        for (int a = 0; a < 3; ++a) {
            const float f = floorf((v[a] - g.lo[a]) * g.inv);
            c[a] = (int)fminf(fmaxf(f, -2.0f), (float)g.dims[a] + 1.0f);
        }
        for (int cz = max(c[2] - 1, 0); cz <= min(c[2] + 1, g.dims[2] - 1) && !hit; ++cz)
            for (int cy = max(c[1] - 1, 0); cy <= min(c[1] + 1, g.dims[1] - 1) && !hit; ++cy)
                for (int cx = max(c[0] - 1, 0); cx <= min(c[0] + 1, g.dims[0] - 1) && !hit; ++cx) {
                    const long cell = ((long)cz * g.dims[1] + cy) * g.dims[0] + cx;
                    for (int k = off[cell]; k < off[cell + 1]; ++k) {
                        const float4 t = sorted[k];
                        const float dx = x - t.x, dy = y - t.y, dz = z - t.z;
                        if (sqrtf((dx * dx + dy * dy) + dz * dz) < radius) {
                            hit = true;
                            break;
                        }
                    }
                }
    }
    const unsigned long long b = __ballot(hit);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(hits, (unsigned long long)__popcll(b));
}

size_t carve(size_t bytes) { return (bytes + 255) & ~size_t(255); }

// Cell edge and counts for a box of targets.  Cells are a sixteenth wider
// than the radius and at most 2^16 per axis, so the float cell coordinates
// (three roundings of values below 2^16: under 2^-6 of a cell) never put a
// pair within reach (less than 0.95 of a cell apart) two cells apart;
// beyond kMaxCells cells the edge doubles (a search over wider cells is still
// exact, only slower).
void plan_cells(const Box &box, float radius, double *edge_out, long dims[3]) {
    double edge = (double)radius * 1.0625;
    for (;;) {
        long total = 1;
        bool fits = true;
        for (int a = 0; a < 3; ++a) {
            dims[a] = (long)std::floor(((double)box.hi[a] - box.lo[a]) / edge) + 1;
            fits = fits && dims[a] <= 65536;
            total *= std::min(dims[a], 65537L);
        }
        if (fits && total <= kMaxCells) break;
        edge *= 2.0;
    }
    *edge_out = edge;
}

// Targets into cells (off: ncells + 1 ints, sorted: n float4), then the hit
// count of the queries into *hits (device).
int count_within(const float *targets, int nt, const float *queries, int nq, float radius,
                 unsigned char *scratch, size_t scratch_bytes, const Box &box,
                 unsigned long long *hits, hipStream_t s) {
    if (nq == 0 || nt == 0) return TDM_OK;
    double edge;
    long dims[3];
    plan_cells(box, radius, &edge, dims);
    Grid g;
    for (int a = 0; a < 3; ++a) {
        g.lo[a] = box.lo[a];
        g.dims[a] = (int)dims[a];
    }
    g.inv = (float)(1.0 / edge);
    const long ncells = dims[0] * dims[1] * dims[2];
    const long nblk = (ncells + kScanPer - 1) / kScanPer;
    unsigned char *p = scratch;
    int *count = reinterpret_cast<int *>(p);
    p += carve(sizeof(int) * ncells);
    int *off = reinterpret_cast<int *>(p);
    p += carve(sizeof(int) * (ncells + 1));
    int *cursor = reinterpret_cast<int *>(p);
    p += carve(sizeof(int) * ncells);
    int *bsum = reinterpret_cast<int *>(p);
    p += carve(sizeof(int) * nblk);
    float4 *sorted = reinterpret_cast<float4 *>(p);
    p += carve(sizeof(float4) * nt);
    TDM_CHECK((size_t)(p - scratch) <= scratch_bytes, TDM_ERR_INVALID_ARG, "scratch too small");
    const unsigned fb = (unsigned)std::min<long>((ncells + kThreads - 1) / kThreads, 16384);
    fill_kernel<<<fb, kThreads, 0, s>>>(count, ncells, 0);
    TDM_LAUNCHED();
    cell_count_kernel<<<tdm::blocks_for(nt, kThreads), kThreads, 0, s>>>(targets, nt, g, count);
    TDM_LAUNCHED();
    scan_sums_kernel<<<(unsigned)nblk, kThreads, 0, s>>>(count, ncells, bsum);
    TDM_LAUNCHED();
    scan_block_sums_kernel<<<1, kScanThreads, 0, s>>>(bsum, (int)nblk);
    TDM_LAUNCHED();
    scan_apply_kernel<<<(unsigned)nblk, kThreads, 0, s>>>(count, ncells, bsum, off, cursor);
    TDM_LAUNCHED();
    cell_scatter_kernel<<<tdm::blocks_for(nt, kThreads), kThreads, 0, s>>>(targets, nt, g, cursor,
                                                                          sorted);
    TDM_LAUNCHED();
    within_radius_kernel<<<tdm::blocks_for(nq, kThreads), kThreads, 0, s>>>(queries, nq, g, off,
                                                                            sorted, radius, hits);
    TDM_LAUNCHED();
    return TDM_OK;
}

size_t grid_bytes(const Box &box, float radius, int nt) {
    double edge;
    long dims[3];
    plan_cells(box, radius, &edge, dims);
    const long total = dims[0] * dims[1] * dims[2];
    const long nblk = (total + kScanPer - 1) / kScanPer;
    return carve(sizeof(int) * total) * 2 + carve(sizeof(int) * (total + 1)) +
           carve(sizeof(int) * nblk) + carve(sizeof(float4) * (size_t)nt);
}

}  // namespace

extern "C" {

int tdm_alignment_ratio(tdm_ctx *ctx, const float *pts1, int num1, const float *pts2, int num2,
                        const double rt[12], float radius, double ratio[2], void *stream) {
    tdm::clear_error();
    TDM_CHECK(pts1 && pts2 && rt && ratio, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_CHECK(num1 > 0 && num2 > 0, TDM_ERR_INVALID_ARG, "empty point cloud (%d, %d)", num1, num2);
    TDM_CHECK(num1 <= (1 << 29) && num2 <= (1 << 29), TDM_ERR_INVALID_ARG, "point clouds too large");
    TDM_CHECK(radius > 0 && std::isfinite(radius), TDM_ERR_INVALID_ARG, "bad radius");
    for (int k = 0; k < 12; ++k)
        TDM_CHECK(std::isfinite(rt[k]), TDM_ERR_INVALID_ARG, "non-finite transform");
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    hipStream_t s = tdm::as_stream(stream);

    // Scratch: moved fragment 2, box partials and results, then the cells of
    // the larger of the two searches.
    const int nbox = 512;
    const size_t head = carve(sizeof(float) * 3 * (size_t)num2) + carve(sizeof(float) * 6 * 2 * nbox) +
                        carve(sizeof(unsigned long long) * 2);
    void *base = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 3, head, &base, s));
    unsigned char *p = static_cast<unsigned char *>(base);
    float *moved = reinterpret_cast<float *>(p);
    float *part = reinterpret_cast<float *>(p + carve(sizeof(float) * 3 * (size_t)num2));

    float f[12];
    for (int k = 0; k < 12; ++k) f[k] = (float)rt[k];
    move_points_kernel<<<tdm::blocks_for(num2, kThreads), kThreads, 0, s>>>(
        pts2, num2, f[0], f[1], f[2], f[4], f[5], f[6], f[8], f[9], f[10], f[3], f[7], f[11], moved);
    TDM_LAUNCHED();
    const unsigned b1 = (unsigned)std::min(nbox, (num1 + kThreads - 1) / kThreads);
    const unsigned b2 = (unsigned)std::min(nbox, (num2 + kThreads - 1) / kThreads);
    box_partial_kernel<<<b1, kThreads, 0, s>>>(pts1, num1, part);
    TDM_LAUNCHED();
    box_partial_kernel<<<b2, kThreads, 0, s>>>(moved, num2, part + 6 * nbox);
    TDM_LAUNCHED();
    std::vector<float> hpart(6 * 2 * nbox);
    TDM_HIP(hipMemcpyAsync(hpart.data(), part, sizeof(float) * hpart.size(), hipMemcpyDeviceToHost, s));
    TDM_HIP(hipStreamSynchronize(s));
    Box box[2];
    for (int c = 0; c < 2; ++c) {
        const unsigned nb = c == 0 ? b1 : b2;
        for (int a = 0; a < 3; ++a) {
            box[c].lo[a] = INFINITY;
            box[c].hi[a] = -INFINITY;
        }
        for (unsigned b = 0; b < nb; ++b)
            for (int a = 0; a < 3; ++a) {
                box[c].lo[a] = std::min(box[c].lo[a], hpart[(c * nbox + b) * 6 + a]);
                box[c].hi[a] = std::max(box[c].hi[a], hpart[(c * nbox + b) * 6 + 3 + a]);
            }
        for (int a = 0; a < 3; ++a)
            TDM_CHECK(std::isfinite(box[c].lo[a]) && std::isfinite(box[c].hi[a]),
                      TDM_ERR_INVALID_ARG, "non-finite point coordinates in fragment %d", c + 1);
    }
    // ratio 1: fragment-1 queries against the moved fragment 2; ratio 2: the
    // moved fragment 2 against fragment 1.  Growing the scratch may replace
    // the buffer, so fragment 2 is moved again into the final one.
    const size_t g1 = grid_bytes(box[1], radius, num2), g2 = grid_bytes(box[0], radius, num1);
    const size_t need = head + std::max(g1, g2);
    TDM_TRY(tdm::ctx_scratch(ctx, 3, need, &base, s));
    p = static_cast<unsigned char *>(base);
    moved = reinterpret_cast<float *>(p);
    unsigned long long *hits = reinterpret_cast<unsigned long long *>(
        p + carve(sizeof(float) * 3 * (size_t)num2) + carve(sizeof(float) * 6 * 2 * nbox));
    unsigned char *cells = p + head;
    move_points_kernel<<<tdm::blocks_for(num2, kThreads), kThreads, 0, s>>>(
        pts2, num2, f[0], f[1], f[2], f[4], f[5], f[6], f[8], f[9], f[10], f[3], f[7], f[11], moved);
    TDM_LAUNCHED();
    fill_kernel<<<1, kThreads, 0, s>>>(reinterpret_cast<int *>(hits), 4, 0);
    TDM_LAUNCHED();
    TDM_TRY(count_within(moved, num2, pts1, num1, radius, cells, need - head, box[1], hits, s));
    TDM_TRY(count_within(pts1, num1, moved, num2, radius, cells, need - head, box[0], hits + 1, s));
    unsigned long long h[2];
    TDM_HIP(hipMemcpyAsync(h, hits, sizeof h, hipMemcpyDeviceToHost, s));
    TDM_HIP(hipStreamSynchronize(s));
    ratio[0] = (double)h[0] / num1;
    ratio[1] = (double)h[1] / num2;
    return TDM_OK;
}

}  // extern "C"
