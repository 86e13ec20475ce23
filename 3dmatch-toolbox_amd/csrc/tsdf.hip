// TSDF depth fusion on gfx950.
//
// The reference kernel (depth-fusion/demo.cu:18-68) runs one thread per
// (z, y) column with a serial loop over x, so the 64 lanes of a wave touch
// addresses dim_x*4 bytes apart.  Here a wave owns a tile of 64 consecutive x
// of one row y and a run of kZRun slices z: every grid access is a coalesced
// 256-byte row segment, the per-column part of the camera transform is
// computed once per lane, and the per-slice part once per tile.
//
// Work per voxel is cut three ways, none of which changes a result:
//  - a tile whose voxels all project outside the image (or behind the
//    camera) is rejected from its four corners before any voxel is touched;
//  - the pixel index comes from a reciprocal instead of two divisions when
//    the pixel coordinate is far enough from a rounding boundary for the
//    reference's rounding to agree (certified below); the rest, and every
//    voxel whose distance needs them, use the reference's own divisions;
//  - voxels in front of the truncation band skip the square root and the
//    divisions whose outcome is already fixed.
//
// Bit-exactness: this file is compiled with -ffp-contract=off and IEEE
// division/sqrt, and every value that reaches the grid (camera coordinates,
// pixel index, distance, running mean) is computed with the reference's
// operations in the reference's order, so pixel indices, the skip tests and
// the weights are bit-identical and the TSDF values equal the reference
// arithmetic.  powf(v, 2) of the reference is evaluated as v*v (the single
// correctly rounded product that gcc and clang both fold it to).
#include "common.hpp"

#pragma clang fp contract(off)

namespace {

constexpr int kRows = 4;   // rows (one per wave) of a workgroup
constexpr int kZRun = 8;   // z slices of a tile
constexpr float kU = 0x1p-24f;  // unit roundoff of float

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
// row of R^T (p - t), summed left to right as the reference does
// (depth-fusion/demo.cu:35-41).
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

// Third term of each row for slice z (uniform over a tile's lanes):
// pt_cam = ColPart + SliceTerm, as (a + b) + c of the reference.
__device__ inline float4 slice_term(const Cam &c, float oz, float vs, int z) {
    const float t2 = (oz + z * vs) - c.tz;
    return make_float4(c.r20 * t2, c.r21 * t2, c.r22 * t2, 0.0f);
}

// ---- Bounds that make the shortcuts exact ----
//
// Pixel coordinate.  The reference computes q = fl(pc_x / pc_z),
// v = fl(fl(fx * q) + cx) and the pixel roundf(v).  Here
// va = fma(fx, fl(pc_x * rcp(pc_z)), cx), where v_rcp_f32 errs by at most
// one ulp (2u relative; u = 2^-24).  Then |fl(pc_x * rcp) - q| <= 4u|Q|
// (Q = pc_x / pc_z exactly) and |va - v| <= 5u|fx Q| + 2u|v| to first order.
// Only va in [-1, W] matters (elsewhere both sides reject the voxel), where
// |v| <= W + 1 and |fx Q| <= W + 1 + |cx|, so
//   B = u (7 (W + 1 + |cx|) + 2 (W + 1)) + 2^-20
// bounds |va - v| with room for a 2-ulp reciprocal.  If |va - rint(va)| <
// 0.5 - B, v lies in the same open rounding interval as va, so
// roundf(v) = rint(va).  Otherwise (about 1 voxel in 2000) the reference's
// own division decides.
//
// Tile rejection.  With S_a = |o_a| + dim_a vs + |t_a|, the float camera
// coordinates of any voxel differ from the exact affine function of its
// index, R^T(o + i vs - t), by at most eta = 8u max_r (sum_a |R_ar| S_a)
// (three roundings in the differences, two in the products and sums).  A
// linear form of the exact coordinates that is negative at a tile's four
// corners is negative on the whole tile.  The image-edge forms
// f pc + (c + 1.5) pc_z (and the three others) put the edge one pixel
// outside the reference's rejection boundary -0.5 (resp. W - 0.5); where
// pc_z >= zeta the float pixel coordinate is within half a pixel of the
// exact one, so a tile with all corners beyond a margin that covers the
// float evaluation of the form is rejected by the reference voxel by voxel.
struct FrameBounds {
    float h;               // certified when |va - rint(va)| < h on both axes
    float eta, zeta;
    float cL, cR, cT, cB;  // constant terms of the four edge forms
    float mL, mR, mT, mB;  // their margins
    float Wf, Hf;
    int edges;             // fx > 0 and fy > 0 (the edge forms assume it)
};

__device__ inline FrameBounds frame_bounds(const Cam &c, int H, int W, int dx, int dy, int dz,
                                           float ox, float oy, float oz, float vs) {
    FrameBounds f;
    const float Wf = (float)W, Hf = (float)H;
    const float acx = fabsf(c.cx), acy = fabsf(c.cy);
    // each bound is rounded up by a factor covering its own float evaluation
    const float hx = 0.5f - (kU * (7.0f * (Wf + 1.0f + acx) + 2.0f * (Wf + 1.0f)) + 0x1p-20f) * 1.01f;
    const float hy = 0.5f - (kU * (7.0f * (Hf + 1.0f + acy) + 2.0f * (Hf + 1.0f)) + 0x1p-20f) * 1.01f;
    f.h = fminf(hx, hy);
    const float Sx = fabsf(ox) + (float)dx * vs + fabsf(c.tx);
    const float Sy = fabsf(oy) + (float)dy * vs + fabsf(c.ty);
    const float Sz = fabsf(oz) + (float)dz * vs + fabsf(c.tz);
    const float ex = fabsf(c.r00) * Sx + fabsf(c.r10) * Sy + fabsf(c.r20) * Sz;
    const float ey = fabsf(c.r01) * Sx + fabsf(c.r11) * Sy + fabsf(c.r21) * Sz;
    const float ez = fabsf(c.r02) * Sx + fabsf(c.r12) * Sy + fabsf(c.r22) * Sz;
    const float pmax = fmaxf(ex, fmaxf(ey, ez)) * 1.01f;  // bounds |camera coordinate|
    const float eta = 8.0f * kU * pmax + 1e-30f;
    f.eta = eta * 1.01f;
    f.zeta = 1.02f * fmaxf(4.0f * eta, 8.0f * eta * (fmaxf(c.fx, c.fy) + fmaxf(Wf, Hf) +
                                                    fmaxf(acx, acy) + 1.0f));
    f.cL = c.cx + 1.5f;
    f.cR = Wf + 0.5f - c.cx;
    f.cT = c.cy + 1.5f;
    f.cB = Hf + 0.5f - c.cy;
    const float g = 1.03f * (eta + 4.0f * kU * pmax);
    f.mL = (c.fx + fabsf(f.cL)) * g + 1e-6f;
    f.mR = (c.fx + fabsf(f.cR)) * g + 1e-6f;
    f.mT = (c.fy + fabsf(f.cT)) * g + 1e-6f;
    f.mB = (c.fy + fabsf(f.cB)) * g + 1e-6f;
    f.Wf = Wf;
    f.Hf = Hf;
    f.edges = c.fx > 0.0f && c.fy > 0.0f;
    return f;
}

// Everything the tile kernels need of one frame.  The multi-frame kernel gets
// them from frame_consts_kernel, once per frame: this GPU has no scalar float
// unit, so uniform float arithmetic costs a full vector instruction in every
// wave that does it, and the kernel reads the block with scalar loads instead.
// The one-frame kernel computes them itself (a separate launch per frame
// costs more than the instructions it saves).
struct alignas(16) FrameConst {
    Cam c;
    FrameBounds f;
    float pad[2];
};
static_assert(sizeof(FrameConst) == 128, "FrameConst is two 64-byte scalar loads");

// Entry i = scene * F + frame; origins (3 per scene) or the single origin.
__global__ void frame_consts_kernel(const float *__restrict__ K, const float *__restrict__ cam2base,
                                    int n, int F, int H, int W, int dx, int dy, int dz,
                                    const float *__restrict__ origins, float ox, float oy,
                                    float oz, float vs, FrameConst *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (origins) {
        const int sc = i / F;
        ox = origins[3 * sc + 0];
        oy = origins[3 * sc + 1];
        oz = origins[3 * sc + 2];
    }
    FrameConst r;
    r.c = load_cam(K, cam2base + (size_t)i * 16);
    r.f = frame_bounds(r.c, H, W, dx, dy, dz, ox, oy, oz, vs);
    r.pad[0] = r.pad[1] = 0.0f;
    out[i] = r;
}

// True when no voxel of the tile (x in [bx0, bx1] as world coordinates, row
// by, slice terms z0 / z1 of its first and last slice) can be seen.  Lanes
// take the corners (lane & 1: x end, lane & 2: z end); every lane of the wave
// must call it.
__device__ inline bool tile_hidden(const Cam &c, const FrameBounds &f, float bx0, float bx1,
                                   float by, float4 z0, float4 z1) {
    const int lane = threadIdx.x & 63;
    const ColPart q = col_part(c, (lane & 1) ? bx1 : bx0, by);
    const float4 t = (lane & 2) ? z1 : z0;
    const float px = q.sx + t.x, py = q.sy + t.y, pz = q.sz + t.z;
    if (__all(pz <= -2.5f * f.eta)) return true;  // behind the camera
    if (!f.edges || !__all(pz >= f.zeta + 1.5f * f.eta)) return false;
    const float hL = c.fx * px + f.cL * pz, hR = f.cR * pz - c.fx * px;
    const float hT = c.fy * py + f.cT * pz, hB = f.cB * pz - c.fy * py;
    return __all(hL < -f.mL) || __all(hR < -f.mR) || __all(hT < -f.mT) || __all(hB < -f.mB);
}

// roundf(f * (pc / pz) + c) of depth-fusion/demo.cu:46-47 with the range test
// of :48-49; -1 when outside [0, n).
__device__ inline int exact_index(float pc, float pz, float f, float c, int n) {
    const float q = pc / pz;
    const float r = roundf(f * q + c);
    return (r >= 0.0f && r < (float)n) ? (int)r : -1;
}

// Depth-image index of a voxel with camera coordinates (pcx, pcy, pcz), or
// -1 when the reference skips it at depth-fusion/demo.cu:42-49.
// cxv, cyv: c.cx, c.cy held in vector registers (an FMA reads one scalar).
__device__ inline int pixel_of(const Cam &c, const FrameBounds &f, float cxv, float cyv,
                               float pcx, float pcy, float pcz, int H, int W) {
    const float r = __builtin_amdgcn_rcpf(pcz);
    const float vx = fmaf(c.fx, pcx * r, cxv);
    const float vy = fmaf(c.fy, pcy * r, cyv);
    const float rx = rintf(vx), ry = rintf(vy);
    // one test for both axes against the smaller of the two bounds; bitwise
    // ands keep the common path free of branches
    const bool cert = (pcz >= 0x1p-64f) & (fmaxf(fabsf(vx - rx), fabsf(vy - ry)) < f.h);
    const bool in = (rx >= 0.0f) & (rx < f.Wf) & (ry >= 0.0f) & (ry < f.Hf);
    // in range: 0 <= ry < H < 2^24 and 0 <= rx < W < 2^24, ry*W + rx < 2^31
    int pix = (cert & in) ? (int)__umul24((unsigned)ry, (unsigned)W) + (int)rx : -1;
    if (!cert && pcz > 0.0f) {
        const int ix = exact_index(pcx, pcz, c.fx, c.cx, W);
        const int iy = exact_index(pcy, pcz, c.fy, c.cy, H);
        pix = (ix >= 0 && iy >= 0) ? iy * W + ix : -1;
    }
    return pix;
}

// depth-fusion/demo.cu:51-62 for a seen voxel with depth sample d.  With
// e = d - pcz, the reference's diff = e*s has s = sqrtf(1+qx^2+qy^2) >= 1, so
// e <= -trunc implies diff <= -trunc (skip) and e >= trunc implies
// diff/trunc >= 1, i.e. dist = 1; the divisions and the square root are only
// evaluated between the two.
__device__ inline bool tsdf_dist(float d, float pcx, float pcy, float pcz, float trunc,
                                 float &dist) {
    if (d <= 0.0f || d > 6.0f) return false;
    const float e = d - pcz;
    if (e <= -trunc) return false;
    dist = 1.0f;
    if (e < trunc) {
        const float qx = pcx / pcz, qy = pcy / pcz;
        const float diff = e * sqrtf(1.0f + qx * qx + qy * qy);
        if (diff <= -trunc) return false;
        dist = fminf(1.0f, diff / trunc);
    }
    return true;
}

// (t*w + dist)/(w + 1) of depth-fusion/demo.cu:63-66.  When the voxel has
// only ever been in front of the surface (t == 1) and is again (dist == 1),
// the quotient is (w + 1)/(w + 1) == 1 exactly, so the division is skipped.
__device__ inline float running_mean(float t, float w, float w_new, float dist) {
    if (t == 1.0f && dist == 1.0f) return 1.0f;
    return (t * w + dist) / w_new;
}

// Grid element at a byte offset from a (uniform) slice base: the address is a
// scalar base plus a 32-bit lane offset.
__device__ inline float &at_byte(float *base, unsigned byte_off) {
    return *reinterpret_cast<float *>(reinterpret_cast<char *>(base) + byte_off);
}
__device__ inline float at_byte(const float *base, unsigned byte_off) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + byte_off);
}

// One frame.  Workgroup = kRows waves; wave = row y = blockIdx.y*kRows + wave,
// lanes = x in [blockIdx.x*64, +64), slices z in [blockIdx.z*kZRun, +kZRun).
__global__ __launch_bounds__(64 * kRows) void integrate_frame_kernel(
    const float *__restrict__ K, const float *__restrict__ cam2base,
    const float *__restrict__ depth, int H, int W, int dx, int dy, int dz, float ox, float oy,
    float oz, float vs, float trunc, float *__restrict__ tsdf, float *__restrict__ weight) {
    __shared__ float4 slices[kRows][kZRun];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int y = blockIdx.y * kRows + wave;
    if (y >= dy) return;  // whole wave
    const Cam c = load_cam(K, cam2base);
    const int z0 = blockIdx.z * kZRun;
    const int nz = min(kZRun, dz - z0);
    const int x0 = blockIdx.x * 64;
    const int x = x0 + lane;
    float4 *zt = slices[wave];
    const float4 mine = slice_term(c, oz, vs, z0 + min(lane, nz - 1));
    if (lane < kZRun) zt[lane] = mine;
    const FrameBounds f = frame_bounds(c, H, W, dx, dy, dz, ox, oy, oz, vs);
    const float by = oy + y * vs;
    float4 first, last;
    first.x = __shfl(mine.x, 0); first.y = __shfl(mine.y, 0); first.z = __shfl(mine.z, 0);
    last.x = __shfl(mine.x, nz - 1); last.y = __shfl(mine.y, nz - 1); last.z = __shfl(mine.z, nz - 1);
    if (tile_hidden(c, f, ox + x0 * vs, ox + min(x0 + 63, dx - 1) * vs, by, first, last)) return;
    if (x >= dx) return;
    float cxv = c.cx, cyv = c.cy;
    asm volatile("" : "+v"(cxv), "+v"(cyv));
    const ColPart p = col_part(c, ox + x * vs, by);
    const unsigned plane = (unsigned)dx * (unsigned)dy;
    const unsigned boff = ((unsigned)y * (unsigned)dx + (unsigned)x) * 4u;
    int pix[kZRun];
    float pz[kZRun];
#pragma unroll
    for (int j = 0; j < kZRun; ++j) {
        const float4 t = zt[j];
        pz[j] = p.sz + t.z;
        const int v = pixel_of(c, f, cxv, cyv, p.sx + t.x, p.sy + t.y, pz[j], H, W);
        pix[j] = j < nz ? v : -1;
    }
    float d[kZRun];
#pragma unroll
    for (int j = 0; j < kZRun; ++j) d[j] = at_byte(depth, (unsigned)max(pix[j], 0) * 4u);
    unsigned upd = 0;
    float dist[kZRun];
#pragma unroll
    for (int j = 0; j < kZRun; ++j) {
        const float4 t = zt[j];
        if (pix[j] >= 0 && tsdf_dist(d[j], p.sx + t.x, p.sy + t.y, pz[j], trunc, dist[j]))
            upd |= 1u << j;
    }
    float w[kZRun], tt[kZRun];
#pragma unroll
    for (int j = 0; j < kZRun; ++j)
        if (upd & (1u << j)) {
            const size_t zoff = (size_t)(z0 + j) * plane;
            w[j] = at_byte(weight + zoff, boff);
            tt[j] = at_byte(tsdf + zoff, boff);
        }
#pragma unroll
    for (int j = 0; j < kZRun; ++j)
        if (upd & (1u << j)) {
            const size_t zoff = (size_t)(z0 + j) * plane;
            const float w_new = w[j] + 1.0f;
            const float t_new = running_mean(tt[j], w[j], w_new, dist[j]);
            at_byte(weight + zoff, boff) = w_new;
            // an unchanged value (same bits) needs no store
            if (__float_as_uint(t_new) != __float_as_uint(tt[j])) at_byte(tsdf + zoff, boff) = t_new;
        }
}

// All frames in one pass: a wave keeps its tile's voxels in registers,
// applies the frames in order, reads a voxel the first time it is updated and
// writes back only updated voxels.  Per voxel the sequence of updates is the
// reference's, so the result is bit-identical.  Scene blockIdx.x / ntx of a
// batch (ntx = x tiles per row) uses its own poses, frames, grid and origin.
__global__ __launch_bounds__(64 * kRows) void integrate_frames_kernel(
    const FrameConst *__restrict__ fc, const float *__restrict__ depths, int F, int H, int W,
    int dx, int dy, int dz, float ox, float oy, float oz, float vs, float trunc,
    float *__restrict__ tsdf, float *__restrict__ weight, const float *__restrict__ origins,
    int ntx) {
    __shared__ float4 slices[kRows][kZRun];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sc = blockIdx.x / ntx;
    const int y = blockIdx.y * kRows + wave;
    if (y >= dy) return;  // whole wave
    const unsigned plane = (unsigned)dx * (unsigned)dy;
    fc += (size_t)sc * F;
    if (origins) {
        depths += (size_t)sc * F * H * W;
        tsdf += (size_t)sc * plane * dz;
        weight += (size_t)sc * plane * dz;
        ox = origins[3 * sc + 0];
        oy = origins[3 * sc + 1];
        oz = origins[3 * sc + 2];
    }
    const int z0 = blockIdx.z * kZRun;
    const int nz = min(kZRun, dz - z0);
    const int x0 = (blockIdx.x - sc * ntx) * 64;
    const int x = x0 + lane;
    const bool live = x < dx;
    const float bx = ox + x * vs, by = oy + y * vs;
    const float bx0 = ox + x0 * vs, bx1 = ox + min(x0 + 63, dx - 1) * vs;
    const unsigned boff = ((unsigned)y * (unsigned)dx + (unsigned)min(x, dx - 1)) * 4u;
    const size_t frame_px = (size_t)H * W;
    float4 *zt = slices[wave];
    float t[kZRun], w[kZRun];
    unsigned loaded = 0;
    for (int fr = 0; fr < F; ++fr) {
        const Cam c = fc[fr].c;
        const float4 mine = slice_term(c, oz, vs, z0 + min(lane, nz - 1));
        // the previous frame's reads of the table are done (same wave, in order)
        if (lane < kZRun) zt[lane] = mine;
        const FrameBounds f = fc[fr].f;
        float4 first, last;
        first.x = __shfl(mine.x, 0); first.y = __shfl(mine.y, 0); first.z = __shfl(mine.z, 0);
        last.x = __shfl(mine.x, nz - 1); last.y = __shfl(mine.y, nz - 1); last.z = __shfl(mine.z, nz - 1);
        if (tile_hidden(c, f, bx0, bx1, by, first, last) || !live) continue;
        float cxv = c.cx, cyv = c.cy;
        asm volatile("" : "+v"(cxv), "+v"(cyv));
        const ColPart p = col_part(c, bx, by);
        const float *__restrict__ depth = depths + fr * frame_px;
        int pix[kZRun];
        float pz[kZRun];
#pragma unroll
        for (int j = 0; j < kZRun; ++j) {
            const float4 q = zt[j];
            pz[j] = p.sz + q.z;
            const int v = pixel_of(c, f, cxv, cyv, p.sx + q.x, p.sy + q.y, pz[j], H, W);
            pix[j] = j < nz ? v : -1;
        }
        float d[kZRun];
#pragma unroll
        for (int j = 0; j < kZRun; ++j) d[j] = at_byte(depth, (unsigned)max(pix[j], 0) * 4u);
        unsigned upd = 0;
        float dist[kZRun];
#pragma unroll
        for (int j = 0; j < kZRun; ++j) {
            const float4 q = zt[j];
            if (pix[j] >= 0 && tsdf_dist(d[j], p.sx + q.x, p.sy + q.y, pz[j], trunc, dist[j]))
                upd |= 1u << j;
        }
        const unsigned fresh = upd & ~loaded;
#pragma unroll
        for (int j = 0; j < kZRun; ++j)
            if (fresh & (1u << j)) {
                const size_t zoff = (size_t)(z0 + j) * plane;
                t[j] = at_byte(tsdf + zoff, boff);
                w[j] = at_byte(weight + zoff, boff);
            }
        loaded |= upd;
#pragma unroll
        for (int j = 0; j < kZRun; ++j)
            if (upd & (1u << j)) {
                const float w_new = w[j] + 1.0f;
                t[j] = running_mean(t[j], w[j], w_new, dist[j]);
                w[j] = w_new;
            }
    }
#pragma unroll
    for (int j = 0; j < kZRun; ++j)
        if (loaded & (1u << j)) {
            const size_t zoff = (size_t)(z0 + j) * plane;
            at_byte(tsdf + zoff, boff) = t[j];
            at_byte(weight + zoff, boff) = w[j];
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
    // the kernels address a slice and the depth image with 32-bit byte offsets
    TDM_CHECK((double)dx * dy * 4.0 < 4294967296.0, TDM_ERR_INVALID_ARG,
              "grid slice too large (%d x %d)", dx, dy);
    TDM_CHECK((double)H * W * 4.0 < 4294967296.0 && W < (1 << 24) && H < (1 << 24),
              TDM_ERR_INVALID_ARG, "image too large");
    return TDM_OK;
}

// Launch grid of the tile kernels: x = 64-voxel tiles of a row (times the
// scene count), y = groups of kRows rows, z = runs of kZRun slices.
int tile_grid(int dx, int dy, int dz, int scenes, dim3 *grid) {
    const long gx = (long)((dx + 63) / 64) * scenes;
    const long gy = (dy + kRows - 1) / kRows, gz = (dz + kZRun - 1) / kZRun;
    TDM_CHECK(gx < (1L << 31) && gy < 65536 && gz < 65536, TDM_ERR_INVALID_ARG, "grid too large");
    *grid = dim3((unsigned)gx, (unsigned)gy, (unsigned)gz);
    return TDM_OK;
}

// Per-frame constants of n = scenes * F frames into the context's scratch
// slot 1, on stream s ahead of the tile kernel that reads them.
int frame_consts(tdm_ctx *ctx, const float *K, const float *cam2base, int n, int F, int H, int W,
                 int dx, int dy, int dz, const float *origins, float ox, float oy, float oz,
                 float vs, hipStream_t s, const FrameConst **out) {
    TDM_CHECK(n > 0 && n < (1 << 24), TDM_ERR_INVALID_ARG, "bad frame count %d", n);
    void *p = nullptr;
    TDM_TRY(tdm::ctx_scratch(ctx, 1, sizeof(FrameConst) * (size_t)n, &p, s));
    frame_consts_kernel<<<(n + 63) / 64, 64, 0, s>>>(K, cam2base, n, F, H, W, dx, dy, dz, origins,
                                                     ox, oy, oz, vs, static_cast<FrameConst *>(p));
    TDM_LAUNCHED();
    *out = static_cast<const FrameConst *>(p);
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
    dim3 grid;
    TDM_TRY(tile_grid(dim_x, dim_y, dim_z, 1, &grid));
    integrate_frame_kernel<<<grid, 64 * kRows, 0, tdm::as_stream(stream)>>>(
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
    dim3 grid;
    TDM_TRY(tile_grid(dim_x, dim_y, dim_z, 1, &grid));
    const hipStream_t s = tdm::as_stream(stream);
    const FrameConst *fc = nullptr;
    TDM_TRY(frame_consts(ctx, cam_K, cam2base, num_frames, num_frames, im_height, im_width, dim_x,
                         dim_y, dim_z, nullptr, origin_x, origin_y, origin_z, voxel_size, s, &fc));
    integrate_frames_kernel<<<grid, 64 * kRows, 0, s>>>(
        fc, depth_ims, num_frames, im_height, im_width, dim_x, dim_y, dim_z, origin_x, origin_y,
        origin_z, voxel_size, trunc_margin, tsdf, weight, nullptr, (int)grid.x);
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
    dim3 grid;
    TDM_TRY(tile_grid(dim_x, dim_y, dim_z, num_scenes, &grid));
    TDM_CHECK((long)num_scenes * num_frames < (1L << 24), TDM_ERR_INVALID_ARG,
              "too many scene frames");
    const hipStream_t s = tdm::as_stream(stream);
    const FrameConst *fc = nullptr;
    TDM_TRY(frame_consts(ctx, cam_K, cam2base, num_scenes * num_frames, num_frames, im_height,
                         im_width, dim_x, dim_y, dim_z, origins, 0.0f, 0.0f, 0.0f, voxel_size, s,
                         &fc));
    integrate_frames_kernel<<<grid, 64 * kRows, 0, s>>>(
        fc, depth_ims, num_frames, im_height, im_width, dim_x, dim_y, dim_z, 0.0f, 0.0f, 0.0f,
        voxel_size, trunc_margin, tsdf, weight, origins, (int)(grid.x / num_scenes));
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
