/*
 * oracle.c -- CPU restatement of the 3DMatch-toolbox hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only as the checker or
 * as the timed CPU baseline.  The product (lib3dmatch_amd.so, the HIP path)
 * never links, loads or calls anything here.
 *
 * Build: oracle/Makefile  (gcc -O3 -fopenmp -ffp-contract=off).  Contraction
 * is off so that every float expression rounds once per operation, in the
 * source order of the reference kernels; that is the defined reference
 * arithmetic for parity (the CUDA build's FMA contraction cannot be
 * reproduced without nvcc).  gcc folds powf(v, 2) to v*v (exact, one
 * rounding); the restatement keeps the reference's powf spelling.
 *
 * Each function names the reference lines it restates.  Pinning of each
 * function (golden vectors, closed forms, library routines) lives in
 * tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* TSDF fusion                                                              */
/* ------------------------------------------------------------------------ */

/* depth-fusion/demo.cu:18-68 (Integrate).  The CUDA grid is <<<dim_z, dim_y>>>
 * with a serial loop over x; here the (z, y) columns are an OpenMP loop.
 * Each voxel is touched by exactly one column, so the result does not depend
 * on the thread schedule. */
ORC_API void orc_integrate(const float *cam_K, const float *cam2base, const float *depth_im,
                           int im_height, int im_width, int dim_x, int dim_y, int dim_z,
                           float origin_x, float origin_y, float origin_z, float voxel_size,
                           float trunc_margin, float *tsdf, float *weight)
{
    long ncol = (long)dim_z * dim_y;
#pragma omp parallel for schedule(static)
    for (long col = 0; col < ncol; ++col) {
        int gz = (int)(col / dim_y);
        int gy = (int)(col % dim_y);
        for (int gx = 0; gx < dim_x; ++gx) {
            /* demo.cu:30-32 voxel corner in base-camera coordinates */
            float bx = origin_x + gx * voxel_size;
            float by = origin_y + gy * voxel_size;
            float bz = origin_z + gz * voxel_size;
            /* demo.cu:35-41 base -> current camera: R^T (p - t) */
            float t0 = bx - cam2base[0 * 4 + 3];
            float t1 = by - cam2base[1 * 4 + 3];
            float t2 = bz - cam2base[2 * 4 + 3];
            float cx = cam2base[0 * 4 + 0] * t0 + cam2base[1 * 4 + 0] * t1 + cam2base[2 * 4 + 0] * t2;
            float cy = cam2base[0 * 4 + 1] * t0 + cam2base[1 * 4 + 1] * t1 + cam2base[2 * 4 + 1] * t2;
            float cz = cam2base[0 * 4 + 2] * t0 + cam2base[1 * 4 + 2] * t1 + cam2base[2 * 4 + 2] * t2;
            if (cz <= 0) /* demo.cu:43-44 */
                continue;
            /* demo.cu:46-49 projection; roundf = half away from zero */
            int px = (int)roundf(cam_K[0 * 3 + 0] * (cx / cz) + cam_K[0 * 3 + 2]);
            int py = (int)roundf(cam_K[1 * 3 + 1] * (cy / cz) + cam_K[1 * 3 + 2]);
            if (px < 0 || px >= im_width || py < 0 || py >= im_height)
                continue;
            float d = depth_im[py * im_width + px]; /* demo.cu:51-54 */
            if (d <= 0 || d > 6)
                continue;
            /* demo.cu:56-58 */
            float diff = (d - cz) * sqrtf(1 + powf((cx / cz), 2) + powf((cy / cz), 2));
            if (diff <= -trunc_margin)
                continue;
            /* demo.cu:61-66 running average; weight is an integer-valued float */
            long idx = ((long)gz * dim_y + gy) * dim_x + gx;
            float dist = fminf(1.0f, diff / trunc_margin);
            float w_old = weight[idx];
            float w_new = w_old + 1.0f;
            weight[idx] = w_new;
            tsdf[idx] = (tsdf[idx] * w_old + dist) / w_new;
        }
    }
}

/* depth-fusion/utils.hpp:89-109 (multiply_matrix): row-major 4x4 product,
 * each entry summed left to right over k. */
ORC_API void orc_multiply_matrix(const float *a, const float *b, float *out)
{
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            float s = a[r * 4 + 0] * b[0 * 4 + c];
            s = s + a[r * 4 + 1] * b[1 * 4 + c];
            s = s + a[r * 4 + 2] * b[2 * 4 + c];
            s = s + a[r * 4 + 3] * b[3 * 4 + c];
            out[r * 4 + c] = s;
        }
}

/* Product of three entries, evaluated (a*b)*c. */
static inline float p3(const float *m, int a, int b, int c) { return m[a] * m[b] * m[c]; }

/* depth-fusion/utils.hpp:112-238 (invert_matrix): cofactor (adjugate)
 * inverse in float, six signed triple products per entry accumulated left to
 * right, det by first-row expansion, reciprocal taken in double then
 * narrowed.  Returns 0 for a singular matrix. */
ORC_API int orc_invert_matrix(const float *m, float *out)
{
    float v[16];
    v[0] = p3(m, 5, 10, 15) - p3(m, 5, 11, 14) - p3(m, 9, 6, 15) + p3(m, 9, 7, 14) + p3(m, 13, 6, 11) - p3(m, 13, 7, 10);
    v[4] = -p3(m, 4, 10, 15) + p3(m, 4, 11, 14) + p3(m, 8, 6, 15) - p3(m, 8, 7, 14) - p3(m, 12, 6, 11) + p3(m, 12, 7, 10);
    v[8] = p3(m, 4, 9, 15) - p3(m, 4, 11, 13) - p3(m, 8, 5, 15) + p3(m, 8, 7, 13) + p3(m, 12, 5, 11) - p3(m, 12, 7, 9);
    v[12] = -p3(m, 4, 9, 14) + p3(m, 4, 10, 13) + p3(m, 8, 5, 14) - p3(m, 8, 6, 13) - p3(m, 12, 5, 10) + p3(m, 12, 6, 9);
    v[1] = -p3(m, 1, 10, 15) + p3(m, 1, 11, 14) + p3(m, 9, 2, 15) - p3(m, 9, 3, 14) - p3(m, 13, 2, 11) + p3(m, 13, 3, 10);
    v[5] = p3(m, 0, 10, 15) - p3(m, 0, 11, 14) - p3(m, 8, 2, 15) + p3(m, 8, 3, 14) + p3(m, 12, 2, 11) - p3(m, 12, 3, 10);
    v[9] = -p3(m, 0, 9, 15) + p3(m, 0, 11, 13) + p3(m, 8, 1, 15) - p3(m, 8, 3, 13) - p3(m, 12, 1, 11) + p3(m, 12, 3, 9);
    v[13] = p3(m, 0, 9, 14) - p3(m, 0, 10, 13) - p3(m, 8, 1, 14) + p3(m, 8, 2, 13) + p3(m, 12, 1, 10) - p3(m, 12, 2, 9);
    v[2] = p3(m, 1, 6, 15) - p3(m, 1, 7, 14) - p3(m, 5, 2, 15) + p3(m, 5, 3, 14) + p3(m, 13, 2, 7) - p3(m, 13, 3, 6);
    v[6] = -p3(m, 0, 6, 15) + p3(m, 0, 7, 14) + p3(m, 4, 2, 15) - p3(m, 4, 3, 14) - p3(m, 12, 2, 7) + p3(m, 12, 3, 6);
    v[10] = p3(m, 0, 5, 15) - p3(m, 0, 7, 13) - p3(m, 4, 1, 15) + p3(m, 4, 3, 13) + p3(m, 12, 1, 7) - p3(m, 12, 3, 5);
    v[14] = -p3(m, 0, 5, 14) + p3(m, 0, 6, 13) + p3(m, 4, 1, 14) - p3(m, 4, 2, 13) - p3(m, 12, 1, 6) + p3(m, 12, 2, 5);
    v[3] = -p3(m, 1, 6, 11) + p3(m, 1, 7, 10) + p3(m, 5, 2, 11) - p3(m, 5, 3, 10) - p3(m, 9, 2, 7) + p3(m, 9, 3, 6);
    v[7] = p3(m, 0, 6, 11) - p3(m, 0, 7, 10) - p3(m, 4, 2, 11) + p3(m, 4, 3, 10) + p3(m, 8, 2, 7) - p3(m, 8, 3, 6);
    v[11] = -p3(m, 0, 5, 11) + p3(m, 0, 7, 9) + p3(m, 4, 1, 11) - p3(m, 4, 3, 9) - p3(m, 8, 1, 7) + p3(m, 8, 3, 5);
    v[15] = p3(m, 0, 5, 10) - p3(m, 0, 6, 9) - p3(m, 4, 1, 10) + p3(m, 4, 2, 9) + p3(m, 8, 1, 6) - p3(m, 8, 2, 5);
    float det = m[0] * v[0] + m[1] * v[4] + m[2] * v[8] + m[3] * v[12];
    if (det == 0)
        return 0;
    det = (float)(1.0 / (double)det);
    for (int i = 0; i < 16; ++i)
        out[i] = v[i] * det;
    return 1;
}

/* depth-fusion/utils.hpp:74-86 (ReadDepth, after decode): millimetres ->
 * metres as (float)u16 / 1000.0f, anything above 6 m becomes 0. */
ORC_API void orc_depth_from_mm(const uint16_t *mm, long n, float *out)
{
    for (long i = 0; i < n; ++i) {
        float d = (float)mm[i] / 1000.0f;
        out[i] = d > 6.0f ? 0.0f : d;
    }
}

/* ------------------------------------------------------------------------ */
/* Point cloud -> TDF                                                       */
/* ------------------------------------------------------------------------ */

/* core/demo.cu:97-122: bounding box, dims = round(extent/vs) + 1 + 2*pad,
 * origin = min - pad*vs + vs/2 (all float). */
ORC_API void orc_grid_params(const float *pts, long n, float voxel_size, int pad, int *dims,
                             float *origin)
{
    float lo[3] = {pts[0], pts[1], pts[2]};
    float hi[3] = {pts[0], pts[1], pts[2]};
    for (long i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
            float v = pts[i * 3 + a];
            lo[a] = v < lo[a] ? v : lo[a];
            hi[a] = v > hi[a] ? v : hi[a];
        }
    for (int a = 0; a < 3; ++a) {
        dims[a] = (int)roundf((hi[a] - lo[a]) / voxel_size) + 1 + pad * 2;
        origin[a] = lo[a] - pad * voxel_size + voxel_size / 2;
    }
}

/* core/demo.cu:127-135: occupancy grid (float 0/1) from rounded voxel
 * coordinates.  Points that land outside the grid are skipped (the
 * reference's bbox rule keeps every point inside). */
ORC_API void orc_occupancy(const float *pts, long n, const int *dims, const float *origin,
                           float voxel_size, float *occ)
{
    long nvox = (long)dims[0] * dims[1] * dims[2];
    memset(occ, 0, sizeof(float) * nvox);
    for (long i = 0; i < n; ++i) {
        int g[3];
        for (int a = 0; a < 3; ++a)
            g[a] = (int)roundf((pts[i * 3 + a] - origin[a]) / voxel_size);
        if (g[0] < 0 || g[1] < 0 || g[2] < 0 || g[0] >= dims[0] || g[1] >= dims[1] || g[2] >= dims[2])
            continue;
        occ[((long)g[2] * dims[1] + g[1]) * dims[0] + g[0]] = 1.0f;
    }
}

/* Search of core/demo.cu:44-56 around (px, py, pz): keep the largest
 * 1 - sqrtf(d2)/r over occupied voxels of the clipped cube, starting from
 * the zero the reference grid is cleared to (demo.cu:145-149). */
static float tdf_search(const float *occ, int dx, int dy, int dz, int px, int py, int pz, int r)
{
    float best = 0.0f;
    int x0 = px - r > 0 ? px - r : 0, x1 = px + r + 1 < dx ? px + r + 1 : dx;
    int y0 = py - r > 0 ? py - r : 0, y1 = py + r + 1 < dy ? py + r + 1 : dy;
    int z0 = pz - r > 0 ? pz - r : 0, z1 = pz + r + 1 < dz ? pz + r + 1 : dz;
    for (int ix = x0; ix < x1; ++ix)
        for (int iy = y0; iy < y1; ++iy)
            for (int iz = z0; iz < z1; ++iz) {
                long j = ((long)iz * dy + iy) * dx + ix;
                if (occ[j] > 0) {
                    float xd = (float)(px - ix), yd = (float)(py - iy), zd = (float)(pz - iz);
                    float dist = sqrtf(xd * xd + yd * yd + zd * zd) / (float)r;
                    if ((1.0f - dist) > best)
                        best = 1.0f - dist;
                }
            }
    return best;
}

/* core/demo.cu:22-57 (ComputeTDF) over the whole grid, with the reference's
 * float32 decode of voxel_idx into (x, y, z) (demo.cu:32-34), which is wrong
 * for part of the voxels once dx*dy*dz exceeds 2^24.  The occupancy test uses
 * the true index (demo.cu:38).  The voxel_idx == N case (demo.cu:29) is not
 * reproduced. */
ORC_API void orc_compute_tdf_ref(const float *occ, float *tdf, int dx, int dy, int dz,
                                 float voxel_size, float trunc_margin)
{
    long n = (long)dx * dy * dz;
    int r = (int)roundf(trunc_margin / voxel_size);
#pragma omp parallel for schedule(dynamic, 4096)
    for (long v = 0; v < n; ++v) {
        int vi = (int)v;
        int gz = (int)floorf((float)vi / ((float)dx * (float)dy));
        int gy = (int)floorf(((float)vi - ((float)gz * (float)dx * (float)dy)) / (float)dx);
        int gx = (int)((float)vi - ((float)gz * (float)dx * (float)dy) - ((float)gy * (float)dx));
        if (occ[v] > 0) {
            tdf[v] = 1.0f;
            continue;
        }
        tdf[v] = tdf_search(occ, dx, dy, dz, gx, gy, gz, r);
    }
}

/* Same kernel with the exact integer decode of voxel_idx.  Identical to
 * orc_compute_tdf_ref whenever dx*dy*dz <= 2^24. */
ORC_API void orc_compute_tdf_exact(const float *occ, float *tdf, int dx, int dy, int dz,
                                   float voxel_size, float trunc_margin)
{
    long n = (long)dx * dy * dz;
    int r = (int)roundf(trunc_margin / voxel_size);
#pragma omp parallel for schedule(dynamic, 4096)
    for (long v = 0; v < n; ++v) {
        int gx = (int)(v % dx), gy = (int)((v / dx) % dy), gz = (int)(v / ((long)dx * dy));
        if (occ[v] > 0) {
            tdf[v] = 1.0f;
            continue;
        }
        tdf[v] = tdf_search(occ, dx, dy, dz, gx, gy, gz, r);
    }
}

/* The float32 decode of core/demo.cu:32-34 on its own (for tests that count
 * mis-decoded voxels). */
ORC_API void orc_float_decode(long v, int dx, int dy, int *xyz)
{
    int vi = (int)v;
    int gz = (int)floorf((float)vi / ((float)dx * (float)dy));
    int gy = (int)floorf(((float)vi - ((float)gz * (float)dx * (float)dy)) / (float)dx);
    int gx = (int)((float)vi - ((float)gz * (float)dx * (float)dy) - ((float)gy * (float)dx));
    xyz[0] = gx;
    xyz[1] = gy;
    xyz[2] = gz;
}

/* core/demo.cu:174-176: keypoint grid coordinate round((p - o)/vs). */
ORC_API void orc_keypoint_grid(const float *kp, long n, const float *origin, float voxel_size,
                               int *grid)
{
    for (long i = 0; i < n; ++i)
        for (int a = 0; a < 3; ++a)
            grid[i * 3 + a] = (int)roundf((kp[i * 3 + a] - origin[a]) / voxel_size);
}

/* core/demo.cu:195-218: 30^3 patch around each keypoint, z-major, x fastest,
 * spanning k-15 .. k+14.  The reference reads outside the grid when k < 15
 * (undefined); here out-of-grid voxels read as 0. */
ORC_API void orc_extract_patches(const float *tdf, int dx, int dy, int dz, const int *kp_grid,
                                 long n, int size, float *out)
{
    int h = size / 2;
#pragma omp parallel for schedule(static)
    for (long p = 0; p < n; ++p) {
        const int *k = kp_grid + p * 3;
        float *o = out + p * (long)size * size * size;
        long li = 0;
        for (int z = k[2] - h; z < k[2] - h + size; ++z)
            for (int y = k[1] - h; y < k[1] - h + size; ++y)
                for (int x = k[0] - h; x < k[0] - h + size; ++x) {
                    int in = x >= 0 && y >= 0 && z >= 0 && x < dx && y < dy && z < dz;
                    o[li++] = in ? tdf[((long)z * dy + y) * dx + x] : 0.0f;
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Descriptor network (cuDNN semantics of core/marvin.hpp:4635-4709,        */
/* 5483-5493, 5578-5588): NCDHW, cross-correlation, stride 1, no padding.   */
/* ------------------------------------------------------------------------ */

/* out[b][co][z][y][x] = bias[co] + sum_{ci,kz,ky,kx} w[co][ci][kz][ky][kx] *
 * in[b][ci][z+kz][y+ky][x+kx], accumulated in float in (ci, kz, ky, kx)
 * order.  in: [nb][ci][d][d][d]; out: [nb][co][d-k+1]^3.  relu != 0 applies
 * max(x, 0) afterwards (ActivationLayer, in place in the reference net). */
ORC_API void orc_conv3d(const float *in, long nb, int cin, int d, const float *w,
                        const float *bias, int cout, int k, int relu, float *out)
{
    int od = d - k + 1;
    long in_vol = (long)d * d * d, out_vol = (long)od * od * od;
    long nwork = nb * cout * od;
#pragma omp parallel for schedule(dynamic, 1)
    for (long job = 0; job < nwork; ++job) {
        long b = job / ((long)cout * od);
        int co = (int)((job / od) % cout);
        int oz = (int)(job % od);
        float acc[1024];
        for (int oy = 0; oy < od; ++oy) {
            for (int ox = 0; ox < od; ++ox)
                acc[ox] = bias ? bias[co] : 0.0f;
            for (int ci = 0; ci < cin; ++ci) {
                const float *src = in + (b * cin + ci) * in_vol;
                const float *wk = w + ((long)co * cin + ci) * k * k * k;
                for (int kz = 0; kz < k; ++kz)
                    for (int ky = 0; ky < k; ++ky) {
                        const float *row = src + ((long)(oz + kz) * d + (oy + ky)) * d;
                        for (int kx = 0; kx < k; ++kx) {
                            float wv = wk[(kz * k + ky) * k + kx];
                            for (int ox = 0; ox < od; ++ox)
                                acc[ox] = acc[ox] + wv * row[ox + kx];
                        }
                    }
            }
            float *dst = out + ((b * cout + co) * out_vol) + ((long)oz * od + oy) * od;
            for (int ox = 0; ox < od; ++ox)
                dst[ox] = relu ? (acc[ox] > 0.0f ? acc[ox] : 0.0f) : acc[ox];
        }
    }
}

/* Max pooling, window w, stride s, no padding; output extent
 * 1 + ceil((d - w)/s) as in core/marvin.hpp:5560-5562, windows clipped to
 * the input. */
ORC_API void orc_maxpool3d(const float *in, long nb, int c, int d, int w, int s, float *out)
{
    int od = 1 + (int)ceilf((float)(d - w) / (float)s);
    long in_vol = (long)d * d * d, out_vol = (long)od * od * od;
#pragma omp parallel for schedule(static)
    for (long bc = 0; bc < nb * c; ++bc) {
        const float *src = in + bc * in_vol;
        float *dst = out + bc * out_vol;
        for (int z = 0; z < od; ++z)
            for (int y = 0; y < od; ++y)
                for (int x = 0; x < od; ++x) {
                    float m = -INFINITY;
                    for (int kz = 0; kz < w; ++kz)
                        for (int ky = 0; ky < w; ++ky)
                            for (int kx = 0; kx < w; ++kx) {
                                int iz = z * s + kz, iy = y * s + ky, ix = x * s + kx;
                                if (iz >= d || iy >= d || ix >= d)
                                    continue;
                                float v = src[((long)iz * d + iy) * d + ix];
                                m = v > m ? v : m;
                            }
                    dst[((long)z * od + y) * od + x] = m;
                }
    }
}

/* orc_compute_tdf_ref at a list of voxel indices only (for grids too large to
 * brute-force whole): out[i] = reference value at voxel idx[i]. */
ORC_API void orc_compute_tdf_ref_at(const float *occ, int dx, int dy, int dz, float voxel_size,
                                    float trunc_margin, const long *idx, long n, float *out)
{
    int r = (int)roundf(trunc_margin / voxel_size);
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < n; ++i) {
        long v = idx[i];
        int vi = (int)v;
        int gz = (int)floorf((float)vi / ((float)dx * (float)dy));
        int gy = (int)floorf(((float)vi - ((float)gz * (float)dx * (float)dy)) / (float)dx);
        int gx = (int)((float)vi - ((float)gz * (float)dx * (float)dy) - ((float)gy * (float)dx));
        out[i] = occ[v] > 0 ? 1.0f : tdf_search(occ, dx, dy, dz, gx, gy, gz, r);
    }
}

/* orc_compute_tdf_exact over the voxel index range [v0, v1) only (for a
 * bounded CPU-baseline sample); tdf is the whole grid. */
ORC_API void orc_compute_tdf_exact_range(const float *occ, float *tdf, int dx, int dy, int dz,
                                         float voxel_size, float trunc_margin, long v0, long v1)
{
    int r = (int)roundf(trunc_margin / voxel_size);
#pragma omp parallel for schedule(dynamic, 4096)
    for (long v = v0; v < v1; ++v) {
        int gx = (int)(v % dx), gy = (int)((v / dx) % dy), gz = (int)(v / ((long)dx * dy));
        tdf[v] = occ[v] > 0 ? 1.0f : tdf_search(occ, dx, dy, dz, gx, gy, gz, r);
    }
}
