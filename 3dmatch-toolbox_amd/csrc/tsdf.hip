// TSDF depth fusion on gfx950.
//
// The reference kernel (depth-fusion/demo.cu:18-68) runs one thread per
// (z, y) column with a serial loop over x, so the 64 lanes of a wave touch
// addresses dim_x*4 bytes apart.  Here a lane owns one (x, y) and sweeps a
// run of z: the lanes of a wave cover consecutive x, every grid access is a
// coalesced 256-byte row segment, and the per-column part of the camera
// transform is computed once per lane instead of once per voxel.
//
// Bit-exactness: this file is compiled with -ffp-contract=off and IEEE
// division/sqrt, and every per-voxel expression keeps the reference's
// operation order, so pixel indices, the skip tests and the weights are
// bit-identical and the TSDF values equal the reference arithmetic.
// powf(v, 2) of the reference is evaluated as v*v (the single correctly
// rounded product that gcc and clang both fold it to).
#include "common.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kThreads = 256;
constexpr int kZRun = 16;  // z slices swept by one lane (one frame)
constexpr int kZRunF = 8;  // z slices kept in registers across frames

// Per-frame camera constants as the kernel uses them.
struct Cam {
    float r00, r10, r20, tx;  // cam2base[0], [4], [8], [3]
    float r01, r11, r21, ty;  // cam2base[1], [5], [9], [7]
    float r02, r12, r22, tz;  // cam2base[2], [6], [10], [11]
    float fx, cx, fy, cy;     // K[0], K[2], K[4], K[5]
};

__device__ inline Cam load_cam(const float *__restrict__ K, const float *__restrict__ m) {
    Cam c;
    c.r00 = m[0]; c.r10 = m[4]; c.r20 = m[8];  c.tx = m[3];
    c.r01 = m[1]; c.r11 = m[5]; c.r21 = m[9];  c.ty = m[7];
    c.r02 = m[2]; c.r12 = m[6]; c.r22 = m[10]; c.tz = m[11];
    c.fx = K[0]; c.cx = K[2]; c.fy = K[4]; c.cy = K[5];
    return c;
}

// Column constants of one (x, y) for one frame: the first two terms of each
// row of R^T (p - t), summed left to right as the reference does.
struct ColPart {
    float sx, sy, sz;
};

__device__ inline ColPart col_part(const Cam &c, float bx, float by) {
    float t0 = bx - c.tx;
    float t1 = by - c.ty;
    ColPart p;
    p.sx = c.r00 * t0 + c.r10 * t1;
    p.sy = c.r01 * t0 + c.r11 * t1;
    p.sz = c.r02 * t0 + c.r12 * t1;
    return p;
}

// The per-voxel body of depth-fusion/demo.cu:34-58 for one frame, split into
// project() (camera coordinates, pixel) and finish() (depth tests, distance).
//
// Shortcuts skip work whose outcome is already decided, without changing any
// result:
//  - a voxel more than one pixel outside the image is rejected with products
//    instead of the two divisions (the margin dwarfs the rounding error of
//    the products; voxels near the border take the exact path);
//  - with e = d - pcz, the reference's diff = e*s has s = sqrtf(1+qx^2+qy^2)
//    >= 1, so e <= -trunc implies diff <= -trunc (skip), and e >= trunc
//    implies diff/trunc >= 1, i.e. dist = 1; the square root and the
//    division are only evaluated between the two.
struct Proj {
    int pix = -1;  // depth-image index, or -1 when the voxel is not seen
    float pcz = 0.0f, qx = 0.0f, qy = 0.0f;
};

// depth-fusion/demo.cu:34-49: camera coordinates and pixel of a voxel.
__device__ inline Proj project(const Cam &c, const ColPart &p, float bz, int H, int W) {
    Proj r;
    float t2 = bz - c.tz;
    float pcx = p.sx + c.r20 * t2;
    float pcy = p.sy + c.r21 * t2;
    float pcz = p.sz + c.r22 * t2;
    if (!(pcz > 0.0f)) return r;
    // fx*pcx/pcz + cx < -1.5, > W + 0.5 (and the same for y): certainly outside.
    const float lx = c.fx * pcx, ly = c.fy * pcy;
    if (lx + (c.cx + 1.5f) * pcz < 0.0f || lx + (c.cx - ((float)W + 0.5f)) * pcz > 0.0f ||
        ly + (c.cy + 1.5f) * pcz < 0.0f || ly + (c.cy - ((float)H + 0.5f)) * pcz > 0.0f)
        return r;
    const float qx = pcx / pcz;
    const float qy = pcy / pcz;
    // roundf = half away from zero; the range test runs on the rounded float,
    // which rejects exactly what the reference's int conversion + test rejects.
    const float rx = roundf(c.fx * qx + c.cx);
    const float ry = roundf(c.fy * qy + c.cy);
    if (!(rx >= 0.0f && rx < (float)W && ry >= 0.0f && ry < (float)H)) return r;
    r.pix = (int)ry * W + (int)rx;
    r.qx = qx;
    r.qy = qy;
    r.pcz = pcz;
    return r;
}

// depth-fusion/demo.cu:51-62 given the depth sample d of a seen voxel.
__device__ inline bool finish(const Proj &r, float d, float trunc, float &dist) {
    if (d <= 0.0f || d > 6.0f) return false;
    const float e = d - r.pcz;
    if (e <= -trunc) return false;
    if (e >= trunc) {
        dist = 1.0f;
        return true;
    }
    float diff = e * sqrtf(1.0f + r.qx * r.qx + r.qy * r.qy);
    if (diff <= -trunc) return false;
    dist = fminf(1.0f, diff / trunc);
    return true;
}

// (t*w + dist)/(w + 1) of depth-fusion/demo.cu:63-66.  When the voxel has
// only ever been in front of the surface (t == 1) and is again (dist == 1),
// the quotient is (w + 1)/(w + 1) == 1 exactly, so the division is skipped.
__device__ inline float running_mean(float t, float w, float w_new, float dist) {
    if (t == 1.0f && dist == 1.0f) return 1.0f;
    return (t * w + dist) / w_new;
}

// One frame; lane = (x, y), blockIdx.y = run of kZRun z slices.
__global__ __launch_bounds__(kThreads) void integrate_frame_kernel(
    const float *__restrict__ K, const float *__restrict__ cam2base,
    const float *__restrict__ depth, int H, int W, int dx, int dy, int dz, float ox, float oy,
    float oz, float vs, float trunc, float *__restrict__ tsdf, float *__restrict__ weight) {
    const long plane = (long)dx * dy;
    const long col = (long)blockIdx.x * kThreads + threadIdx.x;
    if (col >= plane) return;
    const int x = (int)(col % dx);
    const int y = (int)(col / dx);
    const Cam c = load_cam(K, cam2base);
    const ColPart p = col_part(c, ox + x * vs, oy + y * vs);
    const int z0 = blockIdx.y * kZRun;
    const int z1 = min(dz, z0 + kZRun);
    for (int z = z0; z < z1; ++z) {
        const Proj r = project(c, p, oz + z * vs, H, W);
        if (r.pix < 0) continue;
        float dist;
        if (!finish(r, depth[r.pix], trunc, dist)) continue;
        const long i = z * plane + col;
        const float w_old = weight[i];
        const float w_new = w_old + 1.0f;
        const float t_old = tsdf[i];
        const float t_new = running_mean(t_old, w_old, w_new, dist);
        weight[i] = w_new;
        // an unchanged value (same bits) needs no store
        if (__float_as_uint(t_new) != __float_as_uint(t_old)) tsdf[i] = t_new;
    }
}

// All frames in one pass: a lane keeps the kZRunF voxels of its run in
// registers, applies the frames in order, reads a voxel the first time it is
// updated and writes back only updated voxels.  Per voxel the sequence of
// updates is the reference's, so the result is bit-identical.
__global__ __launch_bounds__(kThreads) void integrate_frames_kernel(
    const float *__restrict__ K, const float *__restrict__ cam2base,
    const float *__restrict__ depths, int F, int H, int W, int dx, int dy, int dz, float ox,
    float oy, float oz, float vs, float trunc, float *__restrict__ tsdf,
    float *__restrict__ weight, const float *__restrict__ origins) {
    const long plane = (long)dx * dy;
    const long col = (long)blockIdx.x * kThreads + threadIdx.x;
    if (col >= plane) return;
    // scene blockIdx.z of a batch: its own poses, frames, grid and origin
    if (gridDim.z > 1 || origins) {
        const long sc = blockIdx.z;
        cam2base += sc * F * 16;
        depths += sc * F * (long)H * W;
        tsdf += sc * plane * dz;
        weight += sc * plane * dz;
        if (origins) {
            ox = origins[3 * sc + 0];
            oy = origins[3 * sc + 1];
            oz = origins[3 * sc + 2];
        }
    }
    const int x = (int)(col % dx);
    const int y = (int)(col / dx);
    const float bx = ox + x * vs;
    const float by = oy + y * vs;
    const int z0 = blockIdx.y * kZRunF;
    const int nz = min(dz - z0, kZRunF);
    const long base = z0 * plane + col;
    float t[kZRunF], w[kZRunF];
    unsigned loaded = 0;
    const long frame_px = (long)H * W;
    for (int f = 0; f < F; ++f) {
        const Cam c = load_cam(K, cam2base + f * 16);
        const ColPart p = col_part(c, bx, by);
        const float *__restrict__ depth = depths + f * frame_px;
        Proj pr[kZRunF];
#pragma unroll
        for (int j = 0; j < kZRunF; ++j)
            pr[j] = j < nz ? project(c, p, oz + (z0 + j) * vs, H, W) : Proj{};
        float d[kZRunF];
#pragma unroll
        for (int j = 0; j < kZRunF; ++j) d[j] = depth[max(pr[j].pix, 0)];
        unsigned upd = 0;
        float dist[kZRunF];
#pragma unroll
        for (int j = 0; j < kZRunF; ++j)
            if (pr[j].pix >= 0 && finish(pr[j], d[j], trunc, dist[j])) upd |= 1u << j;
        const unsigned fresh = upd & ~loaded;
#pragma unroll
        for (int j = 0; j < kZRunF; ++j)
            if (fresh & (1u << j)) {
                t[j] = tsdf[base + j * plane];
                w[j] = weight[base + j * plane];
            }
        loaded |= upd;
#pragma unroll
        for (int j = 0; j < kZRunF; ++j)
            if (upd & (1u << j)) {
                const float w_new = w[j] + 1.0f;
                t[j] = running_mean(t[j], w[j], w_new, dist[j]);
                w[j] = w_new;
            }
    }
#pragma unroll
    for (int j = 0; j < kZRunF; ++j) {
        if (loaded & (1u << j)) {
            tsdf[base + j * plane] = t[j];
            weight[base + j * plane] = w[j];
        }
    }
}

__global__ void tsdf_init_kernel(float *__restrict__ tsdf, float *__restrict__ weight, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        tsdf[i] = 1.0f;
        weight[i] = 0.0f;
    }
}

// depth-fusion/utils.hpp:80-84: (float)mm / 1000.0f, above 6 m -> 0.
__global__ void depth_from_mm_kernel(const uint16_t *__restrict__ mm, float *__restrict__ out,
                                     size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float d = (float)mm[i] / 1000.0f;
    out[i] = d > 6.0f ? 0.0f : d;
}

int check_grid(int H, int W, int dx, int dy, int dz, float vs, float trunc) {
    TDM_CHECK(H > 0 && W > 0, TDM_ERR_INVALID_ARG, "bad image size %dx%d", H, W);
    TDM_CHECK(dx > 0 && dy > 0 && dz > 0, TDM_ERR_INVALID_ARG, "bad grid %dx%dx%d", dx, dy, dz);
    TDM_CHECK((double)dx * dy * dz < 9.0e18, TDM_ERR_INVALID_ARG, "grid too large");
    TDM_CHECK((long)H * W < (1L << 31), TDM_ERR_INVALID_ARG, "image too large");
    TDM_CHECK(vs > 0 && trunc > 0, TDM_ERR_INVALID_ARG, "voxel size and truncation must be > 0");
    return TDM_OK;
}

// ---- surface points (SaveVoxelGrid2SurfacePointCloud, depth-fusion/utils.hpp:14-56) ----
//
// Stream compaction in linear-index order: a workgroup owns a chunk of
// kSurfChunk voxels read as 16 coalesced sweeps of 256; pass 1 counts the
// chunk's surface voxels, one workgroup scans the counts, pass 2 recounts
// sweep by sweep (ballots give each voxel its rank) and writes the points.
constexpr int kSurfSweeps = 16;
constexpr long kSurfChunk = 256L * kSurfSweeps;

__device__ inline bool surface_voxel(const float *__restrict__ tsdf,
                                     const float *__restrict__ weight, long i, float tt,
                                     float wt) {
    return fabsf(tsdf[i]) < tt && weight[i] > wt;
}

__global__ __launch_bounds__(256) void surface_count_kernel(const float *__restrict__ tsdf,
                                                            const float *__restrict__ weight,
                                                            long n, float tt, float wt,
                                                            long long *__restrict__ counts) {
    __shared__ int wsum[4];
    const long base = (long)blockIdx.x * kSurfChunk;
    int c = 0;
    for (int k = 0; k < kSurfSweeps; ++k) {
        const long i = base + 256L * k + threadIdx.x;
        c += (i < n && surface_voxel(tsdf, weight, i, tt, wt)) ? 1 : 0;
    }
    for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

// Exclusive scan of the chunk counts in place; total into counts[nb].
__global__ __launch_bounds__(1024) void surface_scan_kernel(long long *__restrict__ counts,
                                                            long nb) {
    __shared__ long long part[1024];
    const long per = (nb + 1023) / 1024;
    const long b0 = min(nb, (long)threadIdx.x * per), b1 = min(nb, b0 + per);
    long long s = 0;
    for (long b = b0; b < b1; ++b) s += counts[b];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const long long v = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += v;
        __syncthreads();
    }
    long long run = threadIdx.x > 0 ? part[threadIdx.x - 1] : 0;
    for (long b = b0; b < b1; ++b) {
        const long long c = counts[b];
        counts[b] = run;
        run += c;
    }
    if (threadIdx.x == 1023) counts[nb] = part[1023];
}

__global__ __launch_bounds__(256) void surface_write_kernel(
    const float *__restrict__ tsdf, const float *__restrict__ weight, long n, float tt, float wt,
    int dx, int dy, float ox, float oy, float oz, float vs, const long long *__restrict__ offsets,
    float *__restrict__ pts, long long capacity) {
    __shared__ int wsum[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long base = (long)blockIdx.x * kSurfChunk;
    const long plane = (long)dx * dy;
    long long at = offsets[blockIdx.x];
    for (int k = 0; k < kSurfSweeps; ++k) {
        const long i = base + 256L * k + threadIdx.x;
        const bool keep = i < n && surface_voxel(tsdf, weight, i, tt, wt);
        const unsigned long long m = __ballot(keep);
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        long long off = at;
        int total = 0;
        for (int w = 0; w < 4; ++w) {
            if (w < wave) off += wsum[w];
            total += wsum[w];
        }
        off += __popcll(m & ((1ull << lane) - 1ull));
        if (keep && off < capacity) {
            // integer decode (the reference's floor of an int quotient)
            const long z = i / plane, r = i - z * plane;
            const long y = r / dx, x = r - y * dx;
            pts[3 * off + 0] = ox + (float)x * vs;
            pts[3 * off + 1] = oy + (float)y * vs;
            pts[3 * off + 2] = oz + (float)z * vs;
        }
        at += total;
        __syncthreads();
    }
}

}  // namespace

extern "C" {

int tdm_tsdf_init(tdm_ctx *ctx, float *tsdf, float *weight, size_t n, void *stream) {
    tdm::clear_error();
    TDM_CHECK(tsdf && weight, TDM_ERR_INVALID_ARG, "null grid");
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    if (n == 0) return TDM_OK;
    unsigned blocks = (unsigned)std::min<size_t>(tdm::blocks_for(n, 256), 65536);
    tsdf_init_kernel<<<blocks, 256, 0, tdm::as_stream(stream)>>>(tsdf, weight, n);
    TDM_LAUNCHED();
    return TDM_OK;
}

int tdm_integrate(tdm_ctx *ctx, const float *cam_K, const float *cam2base, const float *depth_im,
                  int im_height, int im_width, int dim_x, int dim_y, int dim_z, float origin_x,
                  float origin_y, float origin_z, float voxel_size, float trunc_margin,
                  float *tsdf, float *weight, void *stream) {
    tdm::clear_error();
    TDM_CHECK(cam_K && cam2base && depth_im && tsdf && weight, TDM_ERR_INVALID_ARG,
              "null pointer argument");
    TDM_TRY(check_grid(im_height, im_width, dim_x, dim_y, dim_z, voxel_size, trunc_margin));
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    long plane = (long)dim_x * dim_y;
    dim3 grid(tdm::blocks_for(plane, kThreads), (dim_z + kZRun - 1) / kZRun);
    TDM_CHECK(grid.x < (1u << 31) && grid.y < 65536, TDM_ERR_INVALID_ARG, "grid too large");
    integrate_frame_kernel<<<grid, kThreads, 0, tdm::as_stream(stream)>>>(
        cam_K, cam2base, depth_im, im_height, im_width, dim_x, dim_y, dim_z, origin_x, origin_y,
        origin_z, voxel_size, trunc_margin, tsdf, weight);
    TDM_LAUNCHED();
    return TDM_OK;
}

int tdm_integrate_frames(tdm_ctx *ctx, const float *cam_K, const float *cam2base,
                         const float *depth_ims, int num_frames, int im_height, int im_width,
                         int dim_x, int dim_y, int dim_z, float origin_x, float origin_y,
                         float origin_z, float voxel_size, float trunc_margin, float *tsdf,
                         float *weight, void *stream) {
    tdm::clear_error();
    TDM_CHECK(cam_K && cam2base && depth_ims && tsdf && weight, TDM_ERR_INVALID_ARG,
              "null pointer argument");
    TDM_CHECK(num_frames >= 0, TDM_ERR_INVALID_ARG, "negative frame count");
    TDM_TRY(check_grid(im_height, im_width, dim_x, dim_y, dim_z, voxel_size, trunc_margin));
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    if (num_frames == 0) return TDM_OK;
    long plane = (long)dim_x * dim_y;
    dim3 grid(tdm::blocks_for(plane, kThreads), (dim_z + kZRunF - 1) / kZRunF);
    TDM_CHECK(grid.x < (1u << 31) && grid.y < 65536, TDM_ERR_INVALID_ARG, "grid too large");
    integrate_frames_kernel<<<grid, kThreads, 0, tdm::as_stream(stream)>>>(
        cam_K, cam2base, depth_ims, num_frames, im_height, im_width, dim_x, dim_y, dim_z,
        origin_x, origin_y, origin_z, voxel_size, trunc_margin, tsdf, weight, nullptr);
    TDM_LAUNCHED();
    return TDM_OK;
}

int tdm_integrate_scenes(tdm_ctx *ctx, const float *cam_K, const float *cam2base,
                         const float *depth_ims, int num_scenes, int num_frames, int im_height,
                         int im_width, int dim_x, int dim_y, int dim_z, const float *origins,
                         float voxel_size, float trunc_margin, float *tsdf, float *weight,
                         void *stream) {
    tdm::clear_error();
    TDM_CHECK(cam_K && cam2base && depth_ims && origins && tsdf && weight, TDM_ERR_INVALID_ARG,
              "null pointer argument");
    TDM_CHECK(num_frames >= 0 && num_scenes >= 0 && num_scenes < 65536, TDM_ERR_INVALID_ARG,
              "bad scene/frame count");
    TDM_TRY(check_grid(im_height, im_width, dim_x, dim_y, dim_z, voxel_size, trunc_margin));
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    if (num_frames == 0 || num_scenes == 0) return TDM_OK;
    long plane = (long)dim_x * dim_y;
    dim3 grid(tdm::blocks_for(plane, kThreads), (dim_z + kZRunF - 1) / kZRunF, num_scenes);
    TDM_CHECK(grid.x < (1u << 31) && grid.y < 65536, TDM_ERR_INVALID_ARG, "grid too large");
    integrate_frames_kernel<<<grid, kThreads, 0, tdm::as_stream(stream)>>>(
        cam_K, cam2base, depth_ims, num_frames, im_height, im_width, dim_x, dim_y, dim_z, 0.0f,
        0.0f, 0.0f, voxel_size, trunc_margin, tsdf, weight, origins);
    TDM_LAUNCHED();
    return TDM_OK;
}

int tdm_depth_from_mm(tdm_ctx *ctx, const uint16_t *depth_mm, size_t n, float *depth_m,
                      void *stream) {
    tdm::clear_error();
    TDM_CHECK(depth_mm && depth_m, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    if (n == 0) return TDM_OK;
    depth_from_mm_kernel<<<tdm::blocks_for(n, 256), 256, 0, tdm::as_stream(stream)>>>(
        depth_mm, depth_m, n);
    TDM_LAUNCHED();
    return TDM_OK;
}

int tdm_surface_points(tdm_ctx *ctx, const float *tsdf, const float *weight, const int dims[3],
                       const float origin[3], float voxel_size, float tsdf_thresh,
                       float weight_thresh, float *pts, size_t capacity, size_t *num_points,
                       void *stream) {
    tdm::clear_error();
    TDM_CHECK(tsdf && weight && dims && origin && num_points, TDM_ERR_INVALID_ARG,
              "null pointer argument");
    TDM_CHECK(pts || capacity == 0, TDM_ERR_INVALID_ARG, "no output buffer");
    TDM_CHECK(dims[0] > 0 && dims[1] > 0 && dims[2] > 0, TDM_ERR_INVALID_ARG, "bad dims");
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    hipStream_t s = tdm::as_stream(stream);
    const long n = (long)dims[0] * dims[1] * dims[2];
    const long nb = (n + kSurfChunk - 1) / kSurfChunk;
    TDM_CHECK(nb < (1L << 31), TDM_ERR_INVALID_ARG, "grid too large");
    void *scratch = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 1, sizeof(long long) * (nb + 1), &scratch, s));
    long long *counts = static_cast<long long *>(scratch);
    surface_count_kernel<<<(unsigned)nb, 256, 0, s>>>(tsdf, weight, n, tsdf_thresh,
                                                       weight_thresh, counts);
    TDM_LAUNCHED();
    surface_scan_kernel<<<1, 1024, 0, s>>>(counts, nb);
    TDM_LAUNCHED();
    if (capacity > 0) {
        surface_write_kernel<<<(unsigned)nb, 256, 0, s>>>(
            tsdf, weight, n, tsdf_thresh, weight_thresh, dims[0], dims[1], origin[0], origin[1],
            origin[2], voxel_size, counts, pts, (long long)capacity);
        TDM_LAUNCHED();
    }
    long long total = 0;
    TDM_HIP(hipMemcpyAsync(&total, counts + nb, sizeof total, hipMemcpyDeviceToHost, s));
    TDM_HIP(hipStreamSynchronize(s));
    *num_points = (size_t)total;
    return TDM_OK;
}

}  // extern "C"
