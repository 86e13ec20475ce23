// Registration back end of the 3DMatch pipeline on gfx950:
//  - nearest neighbours in descriptor space both ways and the mutual filter
//    (knnsearch on KDTreeSearcher, core/demo.m:49-55,
//    evaluation/geometric-registration/register2Fragments.m:49-55);
//  - RANSAC over 3-point rigid fits (core/external/sfm/ransacfitRt.m:20-64,
//    core/external/peter/ransac.m:113-231) with the Horn-quaternion fit of
//    core/external/estimateRigidTransform/estimateRigidTransform.m:43-72
//    in fp64.
//
// Nearest neighbours.  A KD-tree search in double precision is exact: its
// answer is the argmin of the squared distance.  Here the distance matrix is
// one fp32 GEMM on MFMA (|a|^2 + |b|^2 - 2 a.b over 128 x 128 tiles) whose
// epilogue keeps, for every row and every column of the tile, the best and
// second-best distance and the best index.  A reduction over tiles then
// certifies each query: when its second-best exceeds its best by more than
// twice a bound on the fp32 error of one distance, the fp32 argmin is the
// exact argmin.  Queries that fail (near ties) are searched again with fp64
// sums of squared differences over all candidates, ties to the lower index.
//
// RANSAC.  Trials are independent given their samples, so all of them are
// fitted and scored at once (one thread per hypothesis; the points stream
// through LDS), and the sequential loop of ransac.m -- the `>=` update of the
// best model, the adaptive trial count N = log(1-p)/log(1-w^3) and the trial
// cap -- is replayed afterwards over the per-trial inlier counts to find
// where that loop stops and which trial it keeps.  The samples come from a
// counter-based generator (splitmix64 of seed and trial), not MATLAB's
// RandStream, which cannot be reproduced; everything downstream of the
// samples follows the reference.
#include <climits>
#include <cmath>

#include "common.hpp"
#include "pipeline.hpp"

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 128;     // descriptors of each set per workgroup tile
constexpr int kStage = 32;     // floats of k per LDS stage (128 bytes)
constexpr int kLdsRow = 144;   // LDS bytes per staged row (padded)
constexpr int kMaxDim = 2048;  // descriptor width limit of the fp64 search
constexpr float kInf = __builtin_huge_valf();

struct Top2 {
    float best;
    int idx;
    float second;
};

__device__ inline Top2 top2_empty() { return {kInf, INT_MAX, kInf}; }

// Adds candidate j at distance v.  A NaN distance never enters.
__device__ inline void top2_push(Top2 &t, float v, int j) {
    if (v < t.best || (v == t.best && j < t.idx)) {
        t.second = t.best;
        t.best = v;
        t.idx = j;
    } else if (v < t.second) {
        t.second = v;
    }
}

// Union of two disjoint candidate sets; symmetric in its arguments.
__device__ inline Top2 top2_merge(const Top2 &a, const Top2 &b) {
    if (b.best < a.best || (b.best == a.best && b.idx < a.idx))
        return {b.best, b.idx, fminf(a.best, b.second)};
    return {a.best, a.idx, fminf(a.second, b.best)};
}

__device__ inline Top2 top2_shfl_xor(const Top2 &t, int m) {
    return {__shfl_xor(t.best, m), __shfl_xor(t.idx, m), __shfl_xor(t.second, m)};
}

__device__ inline uint4 ld16(const float *p) { return *reinterpret_cast<const uint4 *>(p); }

__device__ inline float comp(const uint4 &v, int s) {
    const unsigned u = s == 0 ? v.x : s == 1 ? v.y : s == 2 ? v.z : v.w;
    return __uint_as_float(u);
}

__device__ inline double int_to_dbits(int i) {
    return __longlong_as_double((long long)(unsigned)i);
}
__device__ inline int dbits_to_int(double d) { return (int)__double_as_longlong(d); }

// Squared length of every descriptor, summed in fp64 (one wave per row):
// sq = float(|x|^2) feeds the fp32 distances, len = float(|x|) the error
// bound, and maxlen the largest len of the set (as float bits).
__global__ __launch_bounds__(256) void norms_kernel(const float *__restrict__ x, int n, int dim,
                                                    float *__restrict__ sq,
                                                    float *__restrict__ len,
                                                    unsigned *__restrict__ maxlen) {
    const int lane = threadIdx.x & 63;
    const long row = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const float *__restrict__ r = x + row * dim;
    double s = 0.0;
    for (int k = lane; k < dim; k += 64) {
        const double v = r[k];
        s = fma(v, v, s);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    if (lane == 0) {
        sq[row] = (float)s;
        const float l = (float)sqrt(s);
        len[row] = l;
        atomicMax(maxlen, __float_as_uint(l));
    }
}

__global__ void clear_ints_kernel(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// One 128 x 128 tile of fp32 squared distances between descriptors a0.. of A
// and b0.. of B, reduced to per-row and per-column top-2 partials:
//   part_a[tile_b][i] = top-2 of row i over the tile's columns,
//   part_b[tile_a][j] = top-2 of column j over the tile's rows,
// each as float4 {best, index bits, second, 0}.  Four waves in a 2 x 2 grid
// own 64 x 64 sub-tiles (4 x 4 blocks of v_mfma_f32_16x16x4_f32); k streams
// through double-buffered LDS 32 floats at a time.  Rows past the end of a
// set load the last row (in bounds) and are masked in the epilogue; k past
// dim loads zeros (dim is a multiple of 4, so a 16-byte piece is all in or
// all out).
__global__ __launch_bounds__(256, 2) void dist_top2_kernel(
    const float *__restrict__ A, int na, const float *__restrict__ B, int nb, int dim,
    const float *__restrict__ sqa, const float *__restrict__ sqb, float4 *__restrict__ part_a,
    float4 *__restrict__ part_b) {
    constexpr int IB = 4, JB = 4;
    __shared__ __attribute__((aligned(16))) unsigned char lds[2][2 * kTile * kLdsRow];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int ta = blockIdx.x, tb = blockIdx.y;
    const int a0 = ta * kTile, b0 = tb * kTile;

    const float *src[8];
    int dst[8], koff[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int idx = tid + 256 * j, row = idx >> 3, pc = idx & 7;
        if (row < kTile)
            src[j] = A + (long)min(a0 + row, na - 1) * dim;
        else
            src[j] = B + (long)min(b0 + row - kTile, nb - 1) * dim;
        koff[j] = pc * 4;
        dst[j] = row * kLdsRow + pc * 16;
    }
    const int ns = (dim + kStage - 1) / kStage;
    uint4 reg[8];
    auto fetch = [&](int s) {
        const int k0 = s * kStage;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + koff[j];
            reg[j] = k < dim ? ld16(src[j] + k) : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    auto stash = [&](int buf) {
#pragma unroll
        for (int j = 0; j < 8; ++j) *reinterpret_cast<uint4 *>(&lds[buf][dst[j]]) = reg[j];
    };

    floatx4 acc[IB][JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int r16 = lane & 15, g = lane >> 4;
    fetch(0);
    stash(0);
    __syncthreads();
    for (int s = 0; s < ns; ++s) {
        const int buf = s & 1;
        const bool more = s + 1 < ns;
        if (more) fetch(s + 1);
        const unsigned char *As = &lds[buf][(wm * 64 + r16) * kLdsRow + g * 16];
        const unsigned char *Bs = &lds[buf][(kTile + wn * 64 + r16) * kLdsRow + g * 16];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint4 a[IB], b[JB];
#pragma unroll
            for (int i = 0; i < IB; ++i)
                a[i] = *reinterpret_cast<const uint4 *>(As + 16 * i * kLdsRow + h * 64);
#pragma unroll
            for (int j = 0; j < JB; ++j)
                b[j] = *reinterpret_cast<const uint4 *>(Bs + 16 * j * kLdsRow + h * 64);
            // lane (r16, g) supplies k = h*16 + g*4 + step to both operands
#pragma unroll
            for (int st = 0; st < 4; ++st)
#pragma unroll
                for (int i = 0; i < IB; ++i)
#pragma unroll
                    for (int j = 0; j < JB; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                            comp(a[i], st), comp(b[j], st), acc[i][j], 0, 0, 0);
        }
        if (more) stash(buf ^ 1);
        __syncthreads();
    }

    // Epilogue.  Accumulator (i, j, r) of lane (r16, g) is row
    // wm*64 + 16i + 4g + r, column wn*64 + 16j + r16 of the tile.
    float sa[IB][4], sb[JB];
    bool va[IB][4], vb[JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int row = a0 + wm * 64 + 16 * i + 4 * g + r;
            va[i][r] = row < na;
            sa[i][r] = sqa[min(row, na - 1)];
        }
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        const int col = b0 + wn * 64 + 16 * j + r16;
        vb[j] = col < nb;
        sb[j] = sqb[min(col, nb - 1)];
    }
    auto dist = [&](int i, int j, int r) {
        return (va[i][r] && vb[j]) ? fmaf(-2.0f, acc[i][j][r], sa[i][r] + sb[j]) : kInf;
    };

    float4 *rowbuf = reinterpret_cast<float4 *>(&lds[0][0]);  // [2][kTile]
    float4 *colbuf = rowbuf + 2 * kTile;                        // [2][kTile]
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            Top2 t = top2_empty();
#pragma unroll
            for (int j = 0; j < JB; ++j) top2_push(t, dist(i, j, r), b0 + wn * 64 + 16 * j + r16);
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) t = top2_merge(t, top2_shfl_xor(t, m));
            if (r16 == 0)
                rowbuf[wn * kTile + wm * 64 + 16 * i + 4 * g + r] =
                    make_float4(t.best, __int_as_float(t.idx), t.second, 0.f);
        }
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        Top2 t = top2_empty();
#pragma unroll
        for (int i = 0; i < IB; ++i)
#pragma unroll
            for (int r = 0; r < 4; ++r) top2_push(t, dist(i, j, r), a0 + wm * 64 + 16 * i + 4 * g + r);
        t = top2_merge(t, top2_shfl_xor(t, 16));
        t = top2_merge(t, top2_shfl_xor(t, 32));
        if (g == 0)
            colbuf[wm * kTile + wn * 64 + 16 * j + r16] =
                make_float4(t.best, __int_as_float(t.idx), t.second, 0.f);
    }
    __syncthreads();
    const int e = tid & (kTile - 1);
    const float4 *buf = tid < kTile ? rowbuf : colbuf;
    const float4 u = buf[e], v = buf[kTile + e];
    const Top2 t = top2_merge({u.x, __float_as_int(u.y), u.z}, {v.x, __float_as_int(v.y), v.z});
    const float4 out = make_float4(t.best, __int_as_float(t.idx), t.second, 0.f);
    if (tid < kTile) {
        if (a0 + e < na) part_a[(long)tb * na + a0 + e] = out;
    } else {
        if (b0 + e < nb) part_b[(long)ta * nb + b0 + e] = out;
    }
}

// Merges a query's tile partials and certifies its fp32 argmin.
//
// The fp32 distances are taken between centred descriptors q' = fl(q - m),
// c' = fl(c - m) for one float vector m (the mean of both sets): distances
// do not change under a common shift, and the shift removes the common
// component that dominates the lengths, so the bounds below shrink with it.
// With u = 2^-24 and K = dim:
//  - centring: |q' - (q - m)| <= u(1+u)|q'| per vector, so
//    | |q'-c'| - |q-c| | <= eta = u(1+u)(|q'| + |c'|);
//  - the fp32 distance d~ = fl(fl(|q'|^2) + fl(|c'|^2) - 2 fl(q'.c')) is
//    within E = 8 K u |q'||c'| + 4 u (|q'| + |c'|)^2 of |q'-c'|^2 (eight
//    times the summation-order-free dot product bound K u |q'||c'|, plus the
//    rounded lengths and the final two operations).
// |c'| is replaced by the largest centred length of the candidate set.  Every
// rival has d~ >= second, so its exact distance is at least
// sqrt(max(second - E, 0)) - eta, while the best's is at most
// sqrt(best + E) + eta; when the first exceeds the second, the fp32 winner is
// the exact winner.  Other queries go on the refine list.
__global__ void nn_select_kernel(const float4 *__restrict__ part, int nparts, int n, int dim,
                                 const float *__restrict__ len,
                                 const unsigned *__restrict__ maxlen_other,
                                 int *__restrict__ nn, int *__restrict__ todo,
                                 int *__restrict__ todo_count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Top2 t = top2_empty();
    for (int p = 0; p < nparts; ++p) {
        const float4 v = part[(long)p * n + i];
        t = top2_merge(t, {v.x, __float_as_int(v.y), v.z});
    }
    const double u = 5.9604644775390625e-8;  // 2^-24
    // lengths are fp64 sums rounded to float: inflate by a few ulps
    const double q = len[i] * (1.0 + 4 * u), c = __uint_as_float(*maxlen_other) * (1.0 + 4 * u);
    const double E = 8.0 * dim * u * q * c + 4.0 * u * (q + c) * (q + c);
    const double eta = u * (1.0 + u) * (q + c);
    bool sure = false;
    if (t.idx != INT_MAX) {
        const double lo = sqrt(fmax((double)t.second - E, 0.0)) - eta;
        const double hi = sqrt((double)t.best + E) + eta;
        sure = lo > hi;
    }
    if (sure) {
        nn[i] = t.idx;
    } else {
        nn[i] = -1;
        todo[atomicAdd(todo_count, 1)] = i;
    }
}

// Column sums of rows [chunk*rows, (chunk+1)*rows) of X in fp64, for the
// centring mean; part[chunk][k].
__global__ __launch_bounds__(256) void colsum_kernel(const float *__restrict__ X, int n, int dim,
                                                     int rows, double *__restrict__ part) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= dim) return;
    const long r0 = (long)blockIdx.y * rows, r1 = min((long)n, r0 + rows);
    double s = 0.0;
    for (long r = r0; r < r1; ++r) s += X[r * dim + k];
    part[(long)blockIdx.y * dim + k] = s;
}

// mean[k] = float(sum of both sets' partials / (na + nb)), partials in order.
__global__ __launch_bounds__(256) void mean_kernel(const double *__restrict__ part, int nparts,
                                                   int dim, double count,
                                                   float *__restrict__ mean) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= dim) return;
    double s = 0.0;
    for (int p = 0; p < nparts; ++p) s += part[(long)p * dim + k];
    mean[k] = (float)(s / count);
}

// Y = fl(X - mean) row by row.
__global__ __launch_bounds__(256) void center_kernel(const float *__restrict__ X, long total,
                                                     int dim, const float *__restrict__ mean,
                                                     float *__restrict__ Y) {
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < total; i += (long)gridDim.x * 256)
        Y[i] = X[i] - mean[i % dim];
}

// fp64 squared distance between query q and descriptor r: lane-strided sums
// folded by a butterfly, so every lane holds the same value.
__device__ inline double sqdist64(const float *__restrict__ q, const float *__restrict__ r,
                                  int dim, int lane) {
    double s = 0.0;
    for (int k = lane; k < dim; k += 64) {
        const double d = (double)q[k] - (double)r[k];
        s = fma(d, d, s);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m);
    return s;
}

// Exact search for the queries on the refine list, as 64 x 64 tiles of
// (query, candidate) pairs: each pair's squared distance is the fp64 sum of
// (q_k - c_k)^2 in k order, so its value does not depend on the tiling.  The
// tiles stage 32 k of both sides through LDS (k-major, so a thread reads its
// 4 queries and 4 candidates as one float4 each).  Each tile leaves, per
// query, the lowest (distance, index) of its candidates in
// part[tile_c][slot]; tiles past the list's length (read on the device)
// return at once.
constexpr int kRT = 64;
constexpr int kRK = 32;

__global__ __launch_bounds__(256) void nn_refine_tile_kernel(
    const float *__restrict__ Q, const float *__restrict__ R, int nr, int dim,
    const int *__restrict__ todo, const int *__restrict__ todo_count, int slots,
    double2 *__restrict__ part) {
    __shared__ __attribute__((aligned(16))) float sq[kRK][kRT + 4];
    __shared__ __attribute__((aligned(16))) float sc[kRK][kRT + 4];
    const int count = *todo_count;
    const int e0 = blockIdx.x * kRT;
    if (e0 >= count) return;  // uniform over the block
    const int c0 = blockIdx.y * kRT;
    const int tid = threadIdx.x, tq = tid >> 4, tc = tid & 15;
    // loaders: row tid/4 of each tile, 8 consecutive k at (tid%4)*8
    const int lr = tid >> 2, lk = (tid & 3) * 8;
    const float *qrow = Q + (long)todo[min(e0 + lr, count - 1)] * dim;
    const float *crow = R + (long)min(c0 + lr, nr - 1) * dim;
    double acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
    for (int k0 = 0; k0 < dim; k0 += kRK) {
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int k = k0 + lk + j;
            sq[lk + j][lr] = k < dim ? qrow[k] : 0.0f;
            sc[lk + j][lr] = k < dim ? crow[k] : 0.0f;
        }
        __syncthreads();
        const int kn = min(kRK, dim - k0);
        for (int kk = 0; kk < kn; ++kk) {
            const float4 qv = *reinterpret_cast<const float4 *>(&sq[kk][tq * 4]);
            const float4 cv = *reinterpret_cast<const float4 *>(&sc[kk][tc * 4]);
            const double qd[4] = {qv.x, qv.y, qv.z, qv.w};
            const double cd[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const double d = qd[a] - cd[b];
                    acc[a][b] = fma(d, d, acc[a][b]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        double best = __builtin_huge_val();
        int bi = INT_MAX;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int c = c0 + tc * 4 + b;
            if (c < nr && acc[a][b] < best) {
                best = acc[a][b];
                bi = c;
            }
        }
        // lowest (distance, index) over the 16 threads of this query group
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) {
            const double ob = __shfl_xor(best, m);
            const int oi = __shfl_xor(bi, m);
            if (ob < best || (ob == best && oi < bi)) {
                best = ob;
                bi = oi;
            }
        }
        const int e = e0 + tq * 4 + a;
        if (tc == 0 && e < count)
            part[(long)blockIdx.y * slots + e] = make_double2(best, int_to_dbits(bi));
    }
}

// Lowest (distance, index) over the candidate tiles for each listed query.
__global__ void nn_refine_merge_kernel(const double2 *__restrict__ part, int ctiles, int slots,
                                       const int *__restrict__ todo,
                                       const int *__restrict__ todo_count, int *__restrict__ nn) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= *todo_count) return;
    double best = __builtin_huge_val();
    int bi = INT_MAX;
    for (int t = 0; t < ctiles; ++t) {
        const double2 v = part[(long)t * slots + e];
        const int vi = dbits_to_int(v.y);
        if (v.x < best || (v.x == best && vi < bi)) {
            best = v.x;
            bi = vi;
        }
    }
    nn[todo[e]] = bi == INT_MAX ? -1 : bi;
}

// Euclidean distance of each query to its neighbour (knnsearch's second
// output), from the same fp64 sums; NaN where there is no neighbour.
__global__ __launch_bounds__(256) void nn_dist_kernel(const float *__restrict__ Q,
                                                      const float *__restrict__ R, int n,
                                                      int dim, const int *__restrict__ nn,
                                                      float *__restrict__ dist) {
    const int lane = threadIdx.x & 63;
    const long i = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (i >= n) return;
    const int c = nn[i];
    const double s = c >= 0 ? sqdist64(Q + i * dim, R + (long)c * dim, dim, lane)
                            : __builtin_nan("");
    if (lane == 0) dist[i] = (float)sqrt(s);
}

// Mutual pairs in ascending order of the B index: j is kept when
// nn_ab[nn_ba[j]] == j (core/demo.m:53-55).  One workgroup; ballots give each
// kept pair its slot.
__global__ __launch_bounds__(1024) void mutual_kernel(const int *__restrict__ nn_ab, int na,
                                                      const int *__restrict__ nn_ba, int nb,
                                                      int *__restrict__ pairs,
                                                      int *__restrict__ n_pairs) {
    __shared__ int wsum[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int base = 0;
    for (int j0 = 0; j0 < nb; j0 += 1024) {
        const int j = j0 + tid;
        int i = -1;
        bool keep = false;
        if (j < nb) {
            i = nn_ba[j];
            keep = i >= 0 && i < na && nn_ab[i] == j;
        }
        const unsigned long long m = __ballot(keep);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int off = base, total = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wave) off += wsum[w];
            total += wsum[w];
        }
        if (keep) {
            pairs[2 * (off + before)] = i;
            pairs[2 * (off + before) + 1] = j;
        }
        base += total;
        __syncthreads();
    }
    if (tid == 0) *n_pairs = base;
}

// ---------------------------------------------------------------------------
// Rigid fit (estimateRigidTransform.m:43-72) in fp64.

// Adds A_i^T A_i of one centred pair to the 10 upper-triangle entries of B,
// with d = y - x and s = y + x (the R12/R21 and crossTimesMatrix(R22_1)
// blocks of A_i = [0, d'; -d, [s]x]).
__device__ inline void add_pair_B(double Bm[10], const double d[3], const double s[3]) {
    const double A[4][4] = {{0.0, d[0], d[1], d[2]},
                            {-d[0], 0.0, -s[2], s[1]},
                            {-d[1], s[2], 0.0, -s[0]},
                            {-d[2], -s[1], s[0], 0.0}};
    int e = 0;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = p; q < 4; ++q) {
            double v = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) v += A[r][p] * A[r][q];
            Bm[e++] += v;
        }
}

// Unit eigenvector of the smallest eigenvalue of symmetric B (packed upper
// triangle) by cyclic Jacobi rotations; equal to V(:,4) of svd(B) for the
// positive semi-definite B, up to sign.
__device__ inline void smallest_eigvec(const double Bm[10], double v[4]) {
    double a[4][4], V[4][4];
    int e = 0;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = p; q < 4; ++q) {
            a[p][q] = a[q][p] = Bm[e++];
        }
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) V[p][q] = p == q ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 32; ++sweep) {
        bool rotated = false;
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = p + 1; q < 4; ++q) {
                const double apq = a[p][q];
                if (apq == 0.0) continue;
                if (fabs(apq) <= 1e-18 * (fabs(a[p][p]) + fabs(a[q][q]))) {
                    a[p][q] = a[q][p] = 0.0;
                    continue;
                }
                rotated = true;
                const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
                const double t =
                    (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
                for (int r = 0; r < 4; ++r) {  // columns p, q
                    const double arp = a[r][p], arq = a[r][q];
                    a[r][p] = c * arp - s * arq;
                    a[r][q] = s * arp + c * arq;
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {  // rows p, q
                    const double apr = a[p][r], aqr = a[q][r];
                    a[p][r] = c * apr - s * aqr;
                    a[q][r] = s * apr + c * aqr;
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const double vrp = V[r][p], vrq = V[r][q];
                    V[r][p] = c * vrp - s * vrq;
                    V[r][q] = s * vrp + c * vrq;
                }
            }
        if (!rotated) break;
    }
    double lo = a[0][0];
#pragma unroll
    for (int r = 0; r < 4; ++r) v[r] = V[r][0];
#pragma unroll
    for (int k = 1; k < 4; ++k)
        if (a[k][k] < lo) {
            lo = a[k][k];
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = V[r][k];
        }
}

// [R | t] (row-major 3x4) mapping y onto x: R = quat2rot(q), t = xc - R yc
// (the T3 * T2 * T1 product of estimateRigidTransform.m:64-68).
__device__ inline void rt_from_B(const double Bm[10], const double xc[3], const double yc[3],
                                 double *rt) {
    double q[4];
    smallest_eigvec(Bm, q);
    const double q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const double R[3][3] = {
        {q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3, 2 * (q1 * q2 - q0 * q3), 2 * (q1 * q3 + q0 * q2)},
        {2 * (q1 * q2 + q0 * q3), q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3, 2 * (q2 * q3 - q0 * q1)},
        {2 * (q1 * q3 - q0 * q2), 2 * (q2 * q3 + q0 * q1), q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3}};
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) rt[4 * r + c] = R[r][c];
        rt[4 * r + 3] = xc[r] - (R[r][0] * yc[0] + R[r][1] * yc[1] + R[r][2] * yc[2]);
    }
}

struct PairSrc {
    const float *x1;    // keypoints of fragment 1 (the fit's x), 3 floats each
    const float *x2;    // keypoints of fragment 2 (the fit's y)
    const int *pairs;   // [k][2] = (index into x1, index into x2), or NULL: k, k
    const int *n_dev;   // pair count on the device, or NULL: n_host
    int n_host;
};

__device__ inline int pair_count(const PairSrc &p) {
    return p.n_dev ? min(*p.n_dev, p.n_host) : p.n_host;
}

__device__ inline void load_pair(const PairSrc &p, int k, double x[3], double y[3]) {
    const int i = p.pairs ? p.pairs[2 * k] : k;
    const int j = p.pairs ? p.pairs[2 * k + 1] : k;
    for (int c = 0; c < 3; ++c) {
        x[c] = p.x1[3 * (long)i + c];
        y[c] = p.x2[3 * (long)j + c];
    }
}

// d = |x - (R y + t)|, inlier when d < threshold (euc3Ddist,
// core/external/sfm/ransacfitRt.m:40-62).
__device__ inline bool is_inlier(const double *rt, const double x[3], const double y[3],
                                 double thr) {
    double s = 0.0;
    for (int r = 0; r < 3; ++r) {
        const double p = rt[4 * r] * y[0] + rt[4 * r + 1] * y[1] + rt[4 * r + 2] * y[2] + rt[4 * r + 3];
        const double e = x[r] - p;
        s += e * e;
    }
    return sqrt(s) < thr;
}

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Three distinct indices of [0, n) for one trial (what randsample(npts, 3)
// provides, ransac.m:136-141): r_k = mix64(mix64(seed) + 3*trial + k).
__device__ inline void sample3(uint64_t seed, long trial, int n, int idx[3]) {
    const uint64_t base = mix64(seed) + 3ull * (uint64_t)trial;
    const int i0 = (int)(mix64(base) % (uint64_t)n);
    int i1 = (int)(mix64(base + 1) % (uint64_t)(n - 1));
    if (i1 >= i0) ++i1;
    const int lo = min(i0, i1), hi = max(i0, i1);
    int i2 = (int)(mix64(base + 2) % (uint64_t)(n - 2));
    if (i2 >= lo) ++i2;
    if (i2 >= hi) ++i2;
    idx[0] = i0;
    idx[1] = i1;
    idx[2] = i2;
}

// Fit of a small set (estimateRt on the three sampled pairs): centroids as
// sum / count, then B over the centred pairs.
__device__ inline void fit_points(const double (*x)[3], const double (*y)[3], int n, double *rt) {
    double xc[3] = {0, 0, 0}, yc[3] = {0, 0, 0};
    for (int k = 0; k < n; ++k)
        for (int c = 0; c < 3; ++c) {
            xc[c] += x[k][c];
            yc[c] += y[k][c];
        }
    for (int c = 0; c < 3; ++c) {
        xc[c] /= n;
        yc[c] /= n;
    }
    double Bm[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < n; ++k) {
        double d[3], s[3];
        for (int c = 0; c < 3; ++c) {
            const double xs = x[k][c] - xc[c], ys = y[k][c] - yc[c];
            d[c] = ys - xs;
            s[c] = ys + xs;
        }
        add_pair_B(Bm, d, s);
    }
    rt_from_B(Bm, xc, yc, rt);
}

// One hypothesis per thread: sample, fit, and clear its inlier count.
__global__ __launch_bounds__(256) void ransac_hyp_kernel(PairSrc ps, long trials, uint64_t seed,
                                                         double *__restrict__ hyp,
                                                         int *__restrict__ counts) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= trials) return;
    counts[t] = 0;
    const int n = pair_count(ps);
    if (n < 4) return;
    int idx[3];
    sample3(seed, t, n, idx);
    double x[3][3], y[3][3];
    for (int k = 0; k < 3; ++k) load_pair(ps, idx[k], x[k], y[k]);
    fit_points(x, y, 3, hyp + 12 * t);
}

// Inlier counts: blockIdx.y picks a chunk of pairs staged in LDS, each thread
// scores its hypothesis on it and adds to the trial's count.
constexpr int kChunk = 512;

__global__ __launch_bounds__(256) void ransac_count_kernel(PairSrc ps, long trials, double thr,
                                                           const double *__restrict__ hyp,
                                                           int *__restrict__ counts) {
    __shared__ double px[kChunk][3], py[kChunk][3];
    const int n = pair_count(ps);
    if (n < 4) return;
    const int c0 = blockIdx.y * kChunk;
    if (c0 >= n) return;  // uniform over the block
    const int cn = min(kChunk, n - c0);
    for (int k = threadIdx.x; k < cn; k += blockDim.x) load_pair(ps, c0 + k, px[k], py[k]);
    __syncthreads();
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= trials) return;
    double rt[12];
    for (int e = 0; e < 12; ++e) rt[e] = hyp[12 * t + e];
    int cnt = 0;
    for (int k = 0; k < cn; ++k) cnt += is_inlier(rt, px[k], py[k], thr) ? 1 : 0;
    if (cnt) atomicAdd(&counts[t], cnt);
}

// Replays the trial loop of ransac.m:125-221 over the counts.  After trial t
// (trialcount = t + 1) the loop stops when trialcount > max_trials or when
// N <= trialcount, N being set from the best count so far:
//   N = max(10, log(1 - 0.99) / log(clamp(1 - (best/n)^3, eps, 1 - eps))).
// best only changes on `count >= best`, so the kept trial is the last one up
// to the stop whose count equals the maximum count up to the stop.
// stats: [0] inliers of the kept trial, [1] kept trial, [2] trials run.
__device__ inline double trials_needed(int best, int n) {
    const double eps = 2.220446049250313e-16;  // 2^-52
    const double frac = (double)best / (double)n;
    double pno = 1.0 - pow(frac, 3.0);
    pno = fmax(eps, pno);
    pno = fmin(1.0 - eps, pno);
    const double N = log(1.0 - 0.99) / log(pno);
    return fmax(N, 10.0);
}

__global__ __launch_bounds__(1024) void ransac_select_kernel(const PairSrc ps, long max_trials,
                                                             const int *__restrict__ counts,
                                                             int *__restrict__ stats) {
    __shared__ int scan[1024];
    __shared__ long stop_s;
    __shared__ int wmax[16], wlast[16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = pair_count(ps);
    if (n < 4) {
        if (tid == 0) {
            stats[0] = n < 3 ? 0 : 3;
            stats[1] = -1;
            stats[2] = 0;
        }
        return;
    }
    const long T = max_trials + 1;
    const long per = (T + 1023) / 1024;
    const long t0 = min(T, tid * per), t1 = min(T, t0 + per);
    int m = 0;
    for (long t = t0; t < t1; ++t) m = max(m, counts[t]);
    scan[tid] = m;
    if (tid == 0) stop_s = T - 1;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive prefix max
        const int v = tid >= off ? scan[tid - off] : 0;
        __syncthreads();
        scan[tid] = max(scan[tid], v);
        __syncthreads();
    }
    int pm = tid > 0 ? scan[tid - 1] : 0;
    int last_pm = -1;
    double N = 0.0;
    for (long t = t0; t < t1; ++t) {
        pm = max(pm, counts[t]);
        if (pm != last_pm) {
            N = trials_needed(pm, n);
            last_pm = pm;
        }
        const double tc = (double)(t + 1);
        if (t + 1 > max_trials || !(N > tc)) {
            atomicMin((unsigned long long *)&stop_s, (unsigned long long)t);
            break;
        }
    }
    __syncthreads();
    const long stop = stop_s;
    // maximum over [0, stop], then the last trial reaching it
    int mv = 0;
    for (long t = tid; t <= stop; t += 1024) mv = max(mv, counts[t]);
    for (int o = 32; o >= 1; o >>= 1) mv = max(mv, __shfl_xor(mv, o));
    if (lane == 0) wmax[wave] = mv;
    __syncthreads();
    int best = 0;
    for (int w = 0; w < 16; ++w) best = max(best, wmax[w]);
    int last = -1;
    for (long t = tid; t <= stop; t += 1024)
        if (counts[t] == best) last = (int)t;
    for (int o = 32; o >= 1; o >>= 1) last = max(last, __shfl_xor(last, o));
    if (lane == 0) wlast[wave] = last;
    __syncthreads();
    if (tid == 0) {
        int l = -1;
        for (int w = 0; w < 16; ++w) l = max(l, wlast[w]);
        stats[0] = best;
        stats[1] = l;
        stats[2] = (int)(stop + 1);
    }
}

// Final model: the kept hypothesis' inliers (ascending, as find() returns
// them) and estimateRt over them (ransacfitRt.m:30).  With exactly three
// pairs, or with fit_all, the fit is over every pair (ransacfitRt.m:22-26);
// with fewer than three the reference's caller falls back to the identity
// (register2Fragments.m:61-64) and stats[3] = 1 records it.
__global__ __launch_bounds__(1024) void ransac_refit_kernel(const PairSrc ps, double thr,
                                                            bool fit_all,
                                                            const double *__restrict__ hyp,
                                                            int *__restrict__ stats,
                                                            double *__restrict__ rt_out,
                                                            int *__restrict__ inliers) {
    __shared__ int wsum[16];
    __shared__ double red[16][10];
    __shared__ double tot[10];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = pair_count(ps);
    if (n < 3) {
        if (tid < 12) rt_out[tid] = (tid % 5 == 0) ? 1.0 : 0.0;
        if (tid == 0) {
            stats[0] = 0;
            stats[1] = -1;
            stats[2] = 0;
            stats[3] = 1;
        }
        return;
    }
    const bool all = fit_all || n == 3;
    double rt_h[12];
    for (int e = 0; e < 12; ++e) rt_h[e] = all ? 0.0 : hyp[12 * (long)stats[1] + e];
    // Fixed-order sum of v[0..nv) over the block; every thread gets the sums.
    auto block_sum = [&](double *v, int nv) {
        for (int e = 0; e < nv; ++e) {
            double s = v[e];
            for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
            if (lane == 0) red[wave][e] = s;
        }
        __syncthreads();
        if (tid < nv) {
            double s = 0.0;
            for (int w = 0; w < 16; ++w) s += red[w][tid];
            tot[tid] = s;
        }
        __syncthreads();
        for (int e = 0; e < nv; ++e) v[e] = tot[e];
        __syncthreads();
    };
    // pass 1: inlier list and centroid sums
    double c6[6] = {0, 0, 0, 0, 0, 0};
    int base = 0;
    for (int k0 = 0; k0 < n; k0 += 1024) {
        const int k = k0 + tid;
        bool in = false;
        double x[3], y[3];
        if (k < n) {
            load_pair(ps, k, x, y);
            in = all || is_inlier(rt_h, x, y, thr);
        }
        if (in)
            for (int c = 0; c < 3; ++c) {
                c6[c] += x[c];
                c6[3 + c] += y[c];
            }
        const unsigned long long m = __ballot(in);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int off = base, total = 0;
        for (int w = 0; w < 16; ++w) {
            if (w < wave) off += wsum[w];
            total += wsum[w];
        }
        if (in) inliers[off + before] = k;
        base += total;
        __syncthreads();
    }
    const int cnt = base;
    block_sum(c6, 6);
    double xc[3], yc[3];
    for (int c = 0; c < 3; ++c) {
        xc[c] = c6[c] / cnt;
        yc[c] = c6[3 + c] / cnt;
    }
    // pass 2: B over the inliers (the list written above by this block)
    double Bm[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int e = tid; e < cnt; e += 1024) {
        double x[3], y[3], d[3], s[3];
        load_pair(ps, inliers[e], x, y);
        for (int c = 0; c < 3; ++c) {
            const double xs = x[c] - xc[c], ys = y[c] - yc[c];
            d[c] = ys - xs;
            s[c] = ys + xs;
        }
        add_pair_B(Bm, d, s);
    }
    block_sum(Bm, 10);
    if (tid == 0) {
        rt_from_B(Bm, xc, yc, rt_out);
        stats[0] = cnt;
        stats[3] = 0;
        if (all) {
            stats[1] = -1;
            stats[2] = 0;
        }
    }
}

int check_aligned(const void *p, const char *what) {
    TDM_CHECK(p, TDM_ERR_INVALID_ARG, "%s is NULL", what);
    TDM_CHECK(((uintptr_t)p & 15) == 0, TDM_ERR_INVALID_ARG, "%s must be 16-byte aligned", what);
    return TDM_OK;
}

// Carves consecutive 256-byte aligned arrays out of a scratch buffer.
template <typename T>
T *carve(unsigned char *&cur, size_t count) {
    T *p = reinterpret_cast<T *>(cur);
    cur += (count * sizeof(T) + 255) & ~size_t(255);
    return p;
}

size_t carved(size_t count, size_t elem) { return (count * elem + 255) & ~size_t(255); }

}  // namespace

namespace tdm {

size_t match_scratch_bytes(int num_a, int num_b, int dim) {
    const size_t na = num_a, nb = num_b, d = dim;
    const size_t tiles_a = (na + kTile - 1) / kTile, tiles_b = (nb + kTile - 1) / kTile;
    const size_t rt_a = (na + kRT - 1) / kRT, rt_b = (nb + kRT - 1) / kRT;
    const size_t chunks = (na + kRT - 1) / kRT + (nb + kRT - 1) / kRT;
    return carved(tiles_b * na, 16) + carved(tiles_a * nb, 16) + carved(4, 4) +
           2 * carved(na, 4) + 2 * carved(nb, 4) + 2 * carved(na, 4) + 2 * carved(nb, 4) +
           carved(na * d, 4) + carved(nb * d, 4) + carved(chunks * d, 8) + carved(d, 4) +
           carved(rt_b * na, 16) + carved(rt_a * nb, 16);
}

size_t ransac_scratch_bytes(long max_trials) {
    const size_t T = (size_t)max_trials + 1;
    return carved(12 * T, 8) + carved(T, 4);
}

}  // namespace tdm

extern "C" {

int tdm_match_nn(tdm_ctx *ctx, const float *desc_a, int num_a, const float *desc_b, int num_b,
                 int dim, int *nn_ab, float *dist_ab, int *nn_ba, float *dist_ba, int *refined,
                 void *stream) {
    tdm::clear_error();
    TDM_CHECK(num_a >= 1 && num_b >= 1, TDM_ERR_INVALID_ARG, "empty descriptor set (%d, %d)",
              num_a, num_b);
    TDM_CHECK(dim >= 4 && dim % 4 == 0 && dim <= kMaxDim, TDM_ERR_INVALID_ARG,
              "descriptor width %d must be a multiple of 4 in [4, %d]", dim, kMaxDim);
    TDM_TRY(check_aligned(desc_a, "desc_a"));
    TDM_TRY(check_aligned(desc_b, "desc_b"));
    TDM_CHECK(nn_ab || nn_ba, TDM_ERR_INVALID_ARG, "no output requested");
    TDM_CHECK(!dist_ab || nn_ab, TDM_ERR_INVALID_ARG, "dist_ab needs nn_ab");
    TDM_CHECK(!dist_ba || nn_ba, TDM_ERR_INVALID_ARG, "dist_ba needs nn_ba");
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    hipStream_t s = tdm::as_stream(stream);
    const int tiles_a = (num_a + kTile - 1) / kTile, tiles_b = (num_b + kTile - 1) / kTile;
    TDM_CHECK(tiles_b <= 65535 && (num_a + kRT - 1) / kRT <= 65535 &&
                  (num_b + kRT - 1) / kRT <= 65535,
              TDM_ERR_INVALID_ARG, "descriptor sets too large (%d, %d)", num_a, num_b);

    const size_t na = num_a, nb = num_b;
    const size_t bytes = tdm::match_scratch_bytes(num_a, num_b, dim);
    void *base = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 2, bytes, &base, s));
    unsigned char *cur = static_cast<unsigned char *>(base);
    float4 *part_a = carve<float4>(cur, (size_t)tiles_b * na);
    float4 *part_b = carve<float4>(cur, (size_t)tiles_a * nb);
    int *small = carve<int>(cur, 4);  // max length of A, of B; refine counts of A, B
    float *sq_a = carve<float>(cur, na), *len_a = carve<float>(cur, na);
    float *sq_b = carve<float>(cur, nb), *len_b = carve<float>(cur, nb);
    int *todo_a = carve<int>(cur, na), *todo_b = carve<int>(cur, nb);
    int *own_ab = carve<int>(cur, na), *own_ba = carve<int>(cur, nb);
    float *cen_a = carve<float>(cur, na * dim), *cen_b = carve<float>(cur, nb * dim);
    const int chunks_a = (num_a + kRT - 1) / kRT, chunks_b = (num_b + kRT - 1) / kRT;
    double *colsum = carve<double>(cur, (size_t)(chunks_a + chunks_b) * dim);
    float *mean = carve<float>(cur, dim);
    double2 *rpart_a = carve<double2>(cur, (size_t)chunks_b * na);
    double2 *rpart_b = carve<double2>(cur, (size_t)chunks_a * nb);
    int *out_ab = nn_ab ? nn_ab : own_ab, *out_ba = nn_ba ? nn_ba : own_ba;
    unsigned *maxlen = reinterpret_cast<unsigned *>(small);

    clear_ints_kernel<<<1, 64, 0, s>>>(small, 4);
    TDM_LAUNCHED();
    // centring shift: the mean of both sets
    const unsigned cb = tdm::blocks_for(dim, 256);
    colsum_kernel<<<dim3(cb, chunks_a), 256, 0, s>>>(desc_a, num_a, dim, kRT, colsum);
    TDM_LAUNCHED();
    colsum_kernel<<<dim3(cb, chunks_b), 256, 0, s>>>(desc_b, num_b, dim, kRT,
                                                     colsum + (size_t)chunks_a * dim);
    TDM_LAUNCHED();
    mean_kernel<<<cb, 256, 0, s>>>(colsum, chunks_a + chunks_b, dim, (double)na + (double)nb,
                                   mean);
    TDM_LAUNCHED();
    const unsigned ctb = (unsigned)std::min<size_t>(tdm::blocks_for(na * dim, 256), 4096);
    center_kernel<<<ctb, 256, 0, s>>>(desc_a, (long)(na * dim), dim, mean, cen_a);
    TDM_LAUNCHED();
    const unsigned ctb2 = (unsigned)std::min<size_t>(tdm::blocks_for(nb * dim, 256), 4096);
    center_kernel<<<ctb2, 256, 0, s>>>(desc_b, (long)(nb * dim), dim, mean, cen_b);
    TDM_LAUNCHED();
    norms_kernel<<<tdm::blocks_for(na, 4), 256, 0, s>>>(cen_a, num_a, dim, sq_a, len_a, maxlen);
    TDM_LAUNCHED();
    norms_kernel<<<tdm::blocks_for(nb, 4), 256, 0, s>>>(cen_b, num_b, dim, sq_b, len_b,
                                                         maxlen + 1);
    TDM_LAUNCHED();
    dist_top2_kernel<<<dim3(tiles_a, tiles_b), 256, 0, s>>>(cen_a, num_a, cen_b, num_b, dim,
                                                            sq_a, sq_b, part_a, part_b);
    TDM_LAUNCHED();
    nn_select_kernel<<<tdm::blocks_for(na, 256), 256, 0, s>>>(part_a, tiles_b, num_a, dim, len_a,
                                                              maxlen + 1, out_ab, todo_a,
                                                              small + 2);
    TDM_LAUNCHED();
    nn_select_kernel<<<tdm::blocks_for(nb, 256), 256, 0, s>>>(part_b, tiles_a, num_b, dim, len_b,
                                                              maxlen, out_ba, todo_b, small + 3);
    TDM_LAUNCHED();
    // exact fp64 search of the uncertified queries, on the original vectors
    nn_refine_tile_kernel<<<dim3(chunks_a, chunks_b), 256, 0, s>>>(
        desc_a, desc_b, num_b, dim, todo_a, small + 2, num_a, rpart_a);
    TDM_LAUNCHED();
    nn_refine_merge_kernel<<<tdm::blocks_for(na, 256), 256, 0, s>>>(rpart_a, chunks_b, num_a,
                                                                    todo_a, small + 2, out_ab);
    TDM_LAUNCHED();
    nn_refine_tile_kernel<<<dim3(chunks_b, chunks_a), 256, 0, s>>>(
        desc_b, desc_a, num_a, dim, todo_b, small + 3, num_b, rpart_b);
    TDM_LAUNCHED();
    nn_refine_merge_kernel<<<tdm::blocks_for(nb, 256), 256, 0, s>>>(rpart_b, chunks_a, num_b,
                                                                    todo_b, small + 3, out_ba);
    TDM_LAUNCHED();
    if (dist_ab) {
        nn_dist_kernel<<<tdm::blocks_for(na, 4), 256, 0, s>>>(desc_a, desc_b, num_a, dim, out_ab,
                                                              dist_ab);
        TDM_LAUNCHED();
    }
    if (dist_ba) {
        nn_dist_kernel<<<tdm::blocks_for(nb, 4), 256, 0, s>>>(desc_b, desc_a, num_b, dim, out_ba,
                                                              dist_ba);
        TDM_LAUNCHED();
    }
    if (refined) TDM_HIP(hipMemcpyAsync(refined, small + 2, 2 * sizeof(int),
                                        hipMemcpyDeviceToDevice, s));
    return TDM_OK;
}

int tdm_mutual_pairs(tdm_ctx *ctx, const int *nn_ab, int num_a, const int *nn_ba, int num_b,
                     int *pairs, int *num_pairs, void *stream) {
    tdm::clear_error();
    TDM_CHECK(nn_ab && nn_ba && pairs && num_pairs, TDM_ERR_INVALID_ARG, "NULL argument");
    TDM_CHECK(num_a >= 1 && num_b >= 1, TDM_ERR_INVALID_ARG, "empty set (%d, %d)", num_a,
              num_b);
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    mutual_kernel<<<1, 1024, 0, tdm::as_stream(stream)>>>(nn_ab, num_a, nn_ba, num_b, pairs,
                                                          num_pairs);
    TDM_LAUNCHED();
    return TDM_OK;
}

int tdm_ransac_rt(tdm_ctx *ctx, const float *kp1, const float *kp2, const int *pairs,
                  const int *num_pairs, int max_pairs, double threshold, long max_trials,
                  uint64_t seed, double *rt, int *inliers, int *stats, int *trial_counts,
                  void *stream) {
    tdm::clear_error();
    TDM_CHECK(kp1 && kp2 && rt && inliers && stats, TDM_ERR_INVALID_ARG, "NULL argument");
    TDM_CHECK(max_pairs >= 0, TDM_ERR_INVALID_ARG, "max_pairs %d < 0", max_pairs);
    TDM_CHECK(threshold > 0.0, TDM_ERR_INVALID_ARG, "threshold must be > 0");
    TDM_CHECK(max_trials >= 1 && max_trials < (1L << 30), TDM_ERR_INVALID_ARG,
              "max_trials %ld out of range", max_trials);
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    hipStream_t s = tdm::as_stream(stream);
    const long T = max_trials + 1;  // ransac.m runs trials until trialcount > maxTrials
    void *base = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 3, tdm::ransac_scratch_bytes(max_trials), &base, s));
    unsigned char *cur = static_cast<unsigned char *>(base);
    double *hyp = carve<double>(cur, 12 * (size_t)T);
    int *counts = trial_counts ? trial_counts : carve<int>(cur, T);
    PairSrc ps{kp1, kp2, pairs, num_pairs, max_pairs};
    const unsigned tb = tdm::blocks_for(T, 256);
    ransac_hyp_kernel<<<tb, 256, 0, s>>>(ps, T, seed, hyp, counts);
    TDM_LAUNCHED();
    const unsigned chunks = std::max(1u, tdm::blocks_for(max_pairs, kChunk));
    TDM_CHECK(chunks <= 65535, TDM_ERR_INVALID_ARG, "max_pairs %d too large", max_pairs);
    ransac_count_kernel<<<dim3(tb, chunks), 256, 0, s>>>(ps, T, threshold, hyp, counts);
    TDM_LAUNCHED();
    ransac_select_kernel<<<1, 1024, 0, s>>>(ps, max_trials, counts, stats);
    TDM_LAUNCHED();
    ransac_refit_kernel<<<1, 1024, 0, s>>>(ps, threshold, false, hyp, stats, rt, inliers);
    TDM_LAUNCHED();
    return TDM_OK;
}

int tdm_estimate_rt(tdm_ctx *ctx, const float *x1, const float *x2, int num, double *rt,
                    void *stream) {
    tdm::clear_error();
    TDM_CHECK(x1 && x2 && rt, TDM_ERR_INVALID_ARG, "NULL argument");
    TDM_CHECK(num >= 3, TDM_ERR_INVALID_ARG, "at least 3 point pairs are needed (%d)", num);
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    hipStream_t s = tdm::as_stream(stream);
    void *base = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 3, carved(4, 4) + carved(num, 4), &base, s));
    unsigned char *cur = static_cast<unsigned char *>(base);
    int *stats = carve<int>(cur, 4);
    int *inl = carve<int>(cur, num);
    PairSrc ps{x1, x2, nullptr, nullptr, num};
    ransac_refit_kernel<<<1, 1024, 0, s>>>(ps, 1.0, true, nullptr, stats, rt, inl);
    TDM_LAUNCHED();
    return TDM_OK;
}

}  // extern "C"
