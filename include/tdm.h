/*
 * tdm.h -- C ABI of lib3dmatch_amd.so, the MI355X (gfx950) implementation of
 * the 3DMatch-toolbox hot path: TSDF depth fusion, point cloud -> TDF, patch
 * extraction and the 3DMatch 30^3 -> 512 descriptor network.
 *
 * Conventions
 *  - Every entry point returns an int status (TDM_OK == 0).  Nothing calls
 *    exit(); tdm_last_error() gives a message for the calling thread.
 *  - Pointers documented as "device" are HIP device pointers on the device
 *    the context belongs to; "host" pointers are ordinary host memory.
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *    Work is enqueued on it; entry points that must read a result back on
 *    the host (marked "synchronizes") wait for the stream.
 *  - Grids are float32, x fastest: index = (z*dim_y + y)*dim_x + x
 *    (depth-fusion/demo.cu:61, core/demo.cu:47).
 *  - A tdm_ctx owns per-device scratch memory.  One context per host thread;
 *    ctx may be NULL in every call that takes one, meaning a library-owned
 *    context for the current HIP device.
 *
 * Every compute entry point is backed by hand-written HIP kernels; there is
 * no CPU fallback.  Without a usable GPU they return TDM_ERR_NO_DEVICE.
 */
#ifndef TDM_H
#define TDM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDM_VERSION 1

enum tdm_status {
    TDM_OK = 0,
    TDM_ERR_INVALID_ARG = 1,
    TDM_ERR_IO = 2,
    TDM_ERR_FORMAT = 3,
    TDM_ERR_HIP = 4,
    TDM_ERR_NOMEM = 5,
    TDM_ERR_UNSUPPORTED = 6,
    TDM_ERR_NO_DEVICE = 7,
};

/* Descriptor-network arithmetic. */
enum tdm_precision {
    TDM_PREC_FP32 = 0, /* float storage, float MFMA (reference DATATYPE 1)      */
    TDM_PREC_BF16 = 1, /* bf16 weights/activations, fp32 accumulation on MFMA   */
    TDM_PREC_FP16 = 2, /* fp16 weights/activations, fp32 accumulation: the half
                          storage / float compute of the reference's default
                          build (DATATYPE 0, core/marvin.hpp:11-27)          */
    TDM_PREC_FP8 = 3,  /* e4m3 weights (one scale per output channel) and
                          activations (one scale per layer, from
                          tdm_net_calibrate), fp32 accumulation; the first
                          convolution computes in fp16, the last layer's
                          output (the descriptor) stays float               */
};

/* How ComputeTDF turns a linear voxel index into (x, y, z). */
enum tdm_decode {
    TDM_DECODE_EXACT = 0,     /* integer decode                              */
    TDM_DECODE_REFERENCE = 1, /* float32 decode of core/demo.cu:32-34, which
                                 differs from EXACT only on grids > 2^24 voxels */
};

typedef struct tdm_ctx tdm_ctx;
typedef struct tdm_net tdm_net;

/* ---- library ------------------------------------------------------------ */
int tdm_version(void);
const char *tdm_status_string(int status);
const char *tdm_last_error(void);
/* Frees memory returned by tdm_read_* helpers. */
void tdm_free(void *p);

/* Context on HIP device `device` (replaces the demos' implicit default
 * device, core/marvin.hpp:7435 cudaSetDevice). */
int tdm_ctx_create(int device, tdm_ctx **out);
int tdm_ctx_destroy(tdm_ctx *ctx);

/* ---- TSDF fusion ---------------------------------------------------------- */

/* Fills tsdf with 1 and weight with 0 (depth-fusion/demo.cu:131-135).
 * tsdf, weight: device, n floats each. */
int tdm_tsdf_init(tdm_ctx *ctx, float *tsdf, float *weight, size_t n, void *stream);

/* One depth frame into the grid; replaces the Integrate kernel and its launch
 * (depth-fusion/demo.cu:18-68, 179-182).  Same arguments and meaning:
 *   cam_K     device, 9 floats, row-major intrinsics
 *   cam2base  device, 16 floats, row-major camera-to-base pose
 *   depth_im  device, im_height*im_width floats, metres (0 = invalid)
 *   tsdf, weight  device grids of dim_x*dim_y*dim_z floats, updated in place.
 * Results are bit-identical to the reference arithmetic (no FMA contraction,
 * IEEE division and square root).  Unlike the reference launch (one thread
 * per y, so dim_y <= 1024) any grid size is accepted. */
int tdm_integrate(tdm_ctx *ctx, const float *cam_K, const float *cam2base, const float *depth_im,
                  int im_height, int im_width, int dim_x, int dim_y, int dim_z, float origin_x,
                  float origin_y, float origin_z, float voxel_size, float trunc_margin,
                  float *tsdf, float *weight, void *stream);

/* num_frames frames in one pass over the grid: cam2base is num_frames*16
 * floats, depth_ims num_frames*im_height*im_width floats (both device).
 * Each voxel applies the frames in order, so the grids are bit-identical to
 * num_frames calls of tdm_integrate (the frame loop of
 * depth-fusion/demo.cu:156-183); each voxel is read and written once. */
int tdm_integrate_frames(tdm_ctx *ctx, const float *cam_K, const float *cam2base,
                         const float *depth_ims, int num_frames, int im_height, int im_width,
                         int dim_x, int dim_y, int dim_z, float origin_x, float origin_y,
                         float origin_z, float voxel_size, float trunc_margin, float *tsdf,
                         float *weight, void *stream);

/* num_scenes independent fusions in one launch (the frame loop of
 * depth-fusion/demo.cu:156-183 for a batch of scenes sharing intrinsics,
 * image size and grid extent): scene s uses poses cam2base[s*F*16 ..],
 * frames depth_ims[s*F*H*W ..], origin origins[3s .. 3s+2] and the grids
 * tsdf/weight[s*N ..] (N = dim_x*dim_y*dim_z).  All device.  Each scene's
 * grids are bit-identical to tdm_integrate_frames on that scene. */
int tdm_integrate_scenes(tdm_ctx *ctx, const float *cam_K, const float *cam2base,
                         const float *depth_ims, int num_scenes, int num_frames, int im_height,
                         int im_width, int dim_x, int dim_y, int dim_z, const float *origins,
                         float voxel_size, float trunc_margin, float *tsdf, float *weight,
                         void *stream);

/* Surface points of a fused grid on the GPU: what
 * SaveVoxelGrid2SurfacePointCloud (depth-fusion/utils.hpp:14-56) writes to
 * tsdf.ply -- every voxel with |tsdf| < tsdf_thresh and weight >
 * weight_thresh (0.2 and 1 at depth-fusion/demo.cu:192-194), in linear
 * index order, as origin + (float)index * voxel_size per axis.  tsdf,
 * weight: device grids; pts: device, capacity*3 floats (the first
 * `capacity` points are written); *num_points (host) receives the total.
 * Synchronizes `stream`. */
int tdm_surface_points(tdm_ctx *ctx, const float *tsdf, const float *weight, const int dims[3],
                       const float origin[3], float voxel_size, float tsdf_thresh,
                       float weight_thresh, float *pts, size_t capacity, size_t *num_points,
                       void *stream);

/* Frame fusion with overlapped I/O (the frame loop of
 * depth-fusion/demo.cu:156-183 as a pipeline): decoder threads read the
 * 16-bit depth PNGs into a ring of pinned buffers while earlier frames are
 * copied (as u16), converted to metres on the GPU (depth-fusion/utils.hpp:
 * 80-84) and integrated, in submission order on the fusion's own stream, so
 * the grids are bit-identical to the serial loop.  The fusion owns a
 * TSDF/weight grid pair initialised as tdm_tsdf_init does.  threads: decoder
 * threads (0 = 4).  cam_K, cam2base, dims, origin: host. */
typedef struct tdm_fusion tdm_fusion;
int tdm_fusion_create(tdm_ctx *ctx, const float cam_K[9], int im_height, int im_width,
                      const int dims[3], const float origin[3], float voxel_size,
                      float trunc_margin, int threads, tdm_fusion **out);
/* Queues one frame: the PNG path and its camera-to-base pose (copied). */
int tdm_fusion_add_png(tdm_fusion *f, const char *depth_png, const float cam2base[16]);
/* Waits until every queued frame is integrated; reports the first error. */
int tdm_fusion_wait(tdm_fusion *f);
/* The fusion's device grids (valid until tdm_fusion_destroy). */
int tdm_fusion_grids(tdm_fusion *f, float **tsdf, float **weight);
/* Waits, then copies the grids into caller buffers (device) on `stream`. */
int tdm_fusion_copy_grids(tdm_fusion *f, float *tsdf, float *weight, void *stream);
int tdm_fusion_destroy(tdm_fusion *f);

/* u16 millimetre depth (device, n values) -> metres with values above 6 m set
 * to 0 (depth-fusion/utils.hpp:74-86 after the PNG decode). */
int tdm_depth_from_mm(tdm_ctx *ctx, const uint16_t *depth_mm, size_t n, float *depth_m,
                      void *stream);

/* ---- point cloud -> TDF ---------------------------------------------------- */

/* Grid size and origin for a cloud (core/demo.cu:97-122):
 * dims = round(extent/voxel_size) + 1 + 2*padding, origin = min - padding*vs
 * + vs/2.  pts: device, num_pts*3 floats.  Synchronizes `stream`. */
int tdm_grid_params(tdm_ctx *ctx, const float *pts, int num_pts, float voxel_size, int padding,
                    int dims_out[3], float origin_out[3], void *stream);

/* Float occupancy grid (1 where a point rounds to the voxel, else 0), the
 * input of ComputeTDF (core/demo.cu:127-135).  occ: device, dims product. */
int tdm_occupancy(tdm_ctx *ctx, const float *pts, int num_pts, const int dims[3],
                  const float origin[3], float voxel_size, float *occ, void *stream);

/* ComputeTDF over the whole grid (core/demo.cu:22-57 and its launch loop
 * :151-158): tdf = 1 on occupied voxels, else max(0, 1 - d/r) over occupied
 * voxels of the (2r+1)^3 cube, r = round(trunc_margin/voxel_size), d the
 * voxel-centre distance.  occ, tdf: device.  Bit-exact (values come from a
 * table of 1 - sqrtf(d^2)/r built on the host).  decode: enum tdm_decode.
 * The origin is not used by the reference kernel and is not taken here. */
int tdm_compute_tdf(tdm_ctx *ctx, const float *occ, float *tdf, int dim_x, int dim_y, int dim_z,
                    float voxel_size, float trunc_margin, int decode, void *stream);

/* Points straight to TDF, with no float occupancy grid: what
 * core/demo.cu:127-158 computes, for dims/origin from tdm_grid_params.
 * pts, tdf: device. */
int tdm_tdf_from_points(tdm_ctx *ctx, const float *pts, int num_pts, const int dims[3],
                        const float origin[3], float voxel_size, float trunc_margin, int decode,
                        float *tdf, void *stream);

/* No counterpart in the reference.  tdm_tdf_from_points for callers that
 * cannot allocate device memory (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc): it calls no hipMalloc, hipFree or
 * allocating routine and takes no context.  Its bit grid and tile flags live
 * in a static device array of TDM_TDF_NOALLOC_SCRATCH_BYTES per GPU (held
 * until the process ends); a grid needing more (about 16 bytes per 128
 * voxels) fails with TDM_ERR_UNSUPPORTED, as does reference decode above
 * 2^24 voxels (below it both decodes are the same computation).  Every call
 * on one GPU shares that array: calls that overlap on a GPU (other streams,
 * other threads) race and must be ordered by the caller.  Results are
 * bit-identical to tdm_tdf_from_points. */
#define TDM_TDF_NOALLOC_SCRATCH_BYTES (16 * 1024 * 1024)
int tdm_tdf_from_points_noalloc(const float *pts, int num_pts, const int dims[3],
                                const float origin[3], float voxel_size, float trunc_margin,
                                int decode, float *tdf, void *stream);

/* Accurate TDF (training/match.hpp:13-51 Get1NNSearchDist,
 * deprecated/pointCloud2AccTDF.m): tdf = 1 - min(d, trunc_margin) /
 * trunc_margin with d the Euclidean distance from each voxel centre
 * (origin + index * voxel_size) to the nearest point, in float: d^2 summed
 * as (dx^2 + dy^2) + dz^2, sqrtf correctly rounded.  Every point counts,
 * inside the grid or not, as the reference searches all the points it is
 * given.  pts, tdf: device.  Reach up to 31.5 voxels. */
int tdm_tdf_accurate(tdm_ctx *ctx, const float *pts, int num_pts, const int dims[3],
                     const float origin[3], float voxel_size, float trunc_margin, float *tdf,
                     void *stream);

/* Many clouds in one launch per step: the TDF of core/demo.cu:127-158 for a
 * batch of fragments of different sizes, as a plan (made once, on the host)
 * and runs (four launches for the whole batch: zero every cloud's output
 * grid, clear the bit grids, scatter every cloud's points, one tile kernel
 * over every grid's tiles; in reference decode two more per cloud above 2^24
 * voxels -- the only grids whose float32 decode differs -- copy its exact TDF
 * to plan-owned scratch, allocated at plan creation, and gather each voxel's
 * value at its decoded index).  A run
 * allocates nothing, sets no memory with memset and never waits on the host,
 * so it can be captured into a hipGraph and replayed; it reads the point
 * buffers named in the plan at run time.  Results are bit-identical to
 * tdm_tdf_from_points per cloud. */
typedef struct tdm_cloud_ref {
    const float *pts;  /* device, num_pts*3 floats                         */
    float *tdf;        /* device, dim_x*dim_y*dim_z floats (output)        */
    int num_pts;
    int dim_x, dim_y, dim_z;
    float origin_x, origin_y, origin_z;
    int reserved;
} tdm_cloud_ref;
typedef struct tdm_tdf_plan tdm_tdf_plan;
/* clouds: host array of num entries (copied). */
int tdm_tdf_plan_create(tdm_ctx *ctx, const tdm_cloud_ref *clouds, int num, float voxel_size,
                        float trunc_margin, int decode, tdm_tdf_plan **out);
int tdm_tdf_plan_run(tdm_tdf_plan *plan, void *stream);
int tdm_tdf_plan_destroy(tdm_tdf_plan *plan);

/* Keypoint grid coordinates round((p - origin)/voxel_size) (core/demo.cu:
 * 174-176).  kp: host num*3 floats; grid_out: host num*3 ints. */
int tdm_keypoints_to_grid(const float *kp, int num, const float origin[3], float voxel_size,
                          int *grid_out);

/* 30^3 (patch_size^3) TDF patches around keypoints, z-major, x fastest,
 * spanning k-15 .. k+14 (core/demo.cu:195-218).  Voxels outside the grid read
 * as 0 (the reference reads out of bounds there).  tdf: device grid;
 * kp_grid: device num*3 ints; patches: device num*patch_size^3 floats. */
int tdm_extract_patches(tdm_ctx *ctx, const float *tdf, int dim_x, int dim_y, int dim_z,
                        const int *kp_grid, int num, int patch_size, float *patches, void *stream);

/* ---- descriptor network ------------------------------------------------------ */

/* Builds the net of a Marvin JSON file (core/marvin.hpp:7530-7542 Net(json)
 * + Malloc, :7632-7651) for the given precision.  Supported layers:
 * PlaceHolderData (data dims from "dim", [N,1,30,30,30]), Convolution (3-D,
 * stride 1, no padding), Activation ReLU, Pooling max.  max_batch bounds the
 * patches of one internal pass (0 = default 256).  Building the net and
 * setting weights need no GPU; the device (ctx's, or the current one for a
 * NULL ctx) is bound, and weights are uploaded, at the first forward. */
int tdm_net_create(tdm_ctx *ctx, const char *json_path, int precision, int max_batch,
                   tdm_net **out);
/* Reads a .marvin file and sets <layer>.weight / <layer>.bias by name
 * (core/marvin.hpp:7605-7615, 3321-3377).  Float, half and double tensors
 * are accepted and converted.  A tensor whose element count does not match
 * is skipped, as the reference does; a layer left without weights makes
 * tdm_net_forward fail (the reference runs on uninitialized memory). */
int tdm_net_load_weights(tdm_net *net, const char *marvin_path);
/* Sets one layer's weights from host float arrays: weight [O][I][k][k][k],
 * bias [O]. */
int tdm_net_set_layer(tdm_net *net, const char *layer_name, const float *weight,
                      size_t weight_count, const float *bias, size_t bias_count);
int tdm_net_destroy(tdm_net *net);
/* Descriptor width (num_output of the last layer, 512). */
int tdm_net_output_dim(const tdm_net *net);
/* Number of convolution layers and their names, in order. */
int tdm_net_num_layers(const tdm_net *net);
const char *tdm_net_layer_name(const tdm_net *net, int index);
/* Shape of convolution layer `index`: info[0..7] = input channels, output
 * channels, window, input extent, output extent, ReLU (0/1), max pool after
 * it (0/1), extent after the pool.  The counterpart of the response shapes
 * Net::Malloc prints (core/marvin.hpp:2988-2996). */
int tdm_net_layer_info(const tdm_net *net, int index, int info[8]);

/* Net::forward (core/marvin.hpp:7653-7679) on `num` patches.
 * data: device, num*patch^3 floats ([num,1,30,30,30] as the "data"
 * response); feat: device, num*output_dim floats (the "feat" response).
 * Any num >= 1 (the reference needs multiples of the JSON batch size). */
int tdm_net_forward(tdm_net *net, const float *data, int num, float *feat, void *stream);

/* Patch gather fused into the first convolution: descriptors straight from
 * a TDF grid and keypoint grid coordinates (core/demo.cu:195-231 without the
 * host gather and the per-batch copies).  tdf, kp_grid, feat: device. */
int tdm_net_forward_grid(tdm_net *net, const float *tdf, int dim_x, int dim_y, int dim_z,
                         const int *kp_grid, int num, float *feat, void *stream);

/* Activation scales of an fp8 net (no counterpart in the reference, which
 * has no 8-bit type): runs `ref` (an fp32/fp16/bf16 net with the same layers
 * and weights) on `num` patches (device, [num,1,30,30,30]) and sets each
 * layer's output scale to max|activation| / 448, so the calibration set's
 * largest activation maps to the largest e4m3 value.  An fp8 net refuses to
 * run until its scales are set.  Synchronizes `stream`. */
int tdm_net_calibrate(tdm_net *net, tdm_net *ref, const float *data, int num, void *stream);
/* Read or set one layer's output activation scale directly (e.g. to reuse a
 * calibration). */
int tdm_net_act_scale(const tdm_net *net, int index, float *scale);
int tdm_net_set_act_scale(tdm_net *net, int index, float scale);

/* Tile autotuner (no counterpart in the reference, whose convolutions are
 * cuDNN's, core/marvin.hpp:4635-4709): times every tile variant of each
 * convolution's implicit-GEMM kernel on passes of `num` patches (capped at
 * max_batch) and keeps the fastest per layer.  All variants give
 * bit-identical results, so tuning changes speed only.  Overwrites the net's
 * activation buffers; synchronizes `stream`. */
int tdm_net_autotune(tdm_net *net, int num, void *stream);
/* The tile variant of layer `index` (-1: default; -2: bad index), and a way
 * to set it (-1 restores the default). */
int tdm_net_layer_variant(const tdm_net *net, int index);
int tdm_net_set_layer_variant(tdm_net *net, int index, int variant);

/* One TDF grid of a multi-grid pass: device pointer and extents. */
typedef struct tdm_grid_ref {
    const float *tdf;
    int dim_x, dim_y, dim_z;
    int reserved;
} tdm_grid_ref;

/* Descriptors for keypoints spread over several TDF grids in one pass (the
 * per-fragment loop of core/demo.cu:195-231 batched across fragments):
 * grids is a device table; keypoint b (grid coords kp_grid[3b..3b+2]) reads
 * grid grid_index[b].  All pointers device. */
int tdm_net_forward_grids(tdm_net *net, const tdm_grid_ref *grids, const int *kp_grid,
                          const int *grid_index, int num, float *feat, void *stream);

/* Output of convolution layer `layer_index` (after its ReLU, and after the
 * pool that follows it, if any) for `num` patches, as float NCDHW into out
 * (device).  For per-layer error measurements. */
int tdm_net_forward_layer(tdm_net *net, const float *data, int num, int layer_index, float *out,
                          size_t out_count, void *stream);

/* Per-layer timing, the counterpart of Marvin's debug_mode tic/toc per layer
 * (core/marvin.hpp:437-456, 7656-7676): while enabled, every forward records
 * HIP events on its stream around each layer's kernels.  Enabling or
 * disabling resets the statistics. */
int tdm_net_set_profiling(tdm_net *net, int enable);
/* Accumulated device time (ms), number of timed passes and patches of one
 * layer since profiling was enabled.  Waits for the recorded events. */
int tdm_net_layer_stats(tdm_net *net, int layer_index, double *total_ms, long *launches,
                        long *patches);

/* ---- registration: matching and RANSAC ------------------------------------ */

/* Nearest neighbour of every descriptor of A among B and of every descriptor
 * of B among A, under Euclidean distance: the two knnsearch calls on
 * KDTreeSearcher of core/demo.m:49-52 and register2Fragments.m:49-52
 * (evaluation/geometric-registration).  Exact: the result is the argmin of
 * the exact squared distance, ties to the lower index (0-based).  desc_a,
 * desc_b: device, row-major num*dim floats, 16-byte aligned, dim a multiple
 * of 4 up to 2048.  Outputs (device, any may be NULL but not both nn):
 * nn_ab[num_a], nn_ba[num_b] indices; dist_ab/dist_ba the Euclidean
 * distances (knnsearch's second output, fp64 sums rounded to float);
 * refined[2] the number of queries of A and of B whose fp32 answer was not
 * certified and that were searched again in fp64. */
int tdm_match_nn(tdm_ctx *ctx, const float *desc_a, int num_a, const float *desc_b, int num_b,
                 int dim, int *nn_ab, float *dist_ab, int *nn_ba, float *dist_ba, int *refined,
                 void *stream);

/* Mutually nearest pairs (core/demo.m:53-55): j of B is kept when
 * nn_ab[nn_ba[j]] == j; pairs (device, capacity num_b*2 ints) receives
 * (nn_ba[j], j) in ascending j, num_pairs (device, 1 int) their count. */
int tdm_mutual_pairs(tdm_ctx *ctx, const int *nn_ab, int num_a, const int *nn_ba, int num_b,
                     int *pairs, int *num_pairs, void *stream);

/* ransacfitRt(x, threshold, 0) (core/external/sfm/ransacfitRt.m:20-64 with
 * core/external/peter/ransac.m:113-231): the rigid [R|t] with x1 = R*x2 + t
 * over pairs of kp1 (fragment 1) and kp2 (fragment 2), all fits fp64.
 *   kp1, kp2   device, 3 floats per keypoint
 *   pairs      device (i, j) index pairs into kp1/kp2, or NULL for (k, k)
 *   num_pairs  device count of pairs (e.g. from tdm_mutual_pairs), or NULL
 *   max_pairs  the host bound on that count (the count when num_pairs is NULL)
 *   threshold  inlier distance (0.05 in the reference's callers)
 *   max_trials the trial cap (50000 in ransac.m); p = 0.99 as there
 *   seed       the trial samples come from splitmix64(seed, trial), not from
 *              MATLAB's RandStream
 * Outputs (device): rt 12 doubles, row-major 3x4; inliers (capacity
 * max_pairs) the pair indices within threshold of the kept hypothesis, in
 * ascending order; stats 4 ints = {inlier count, kept trial (-1 when no
 * sampling ran), trials run, 1 if fewer than 3 pairs (rt = identity, the
 * fallback of register2Fragments.m:61-64) else 0}; trial_counts (NULL or
 * max_trials+1 ints) every trial's inlier count.  Three pairs are fitted
 * directly (ransacfitRt.m:22-26). */
int tdm_ransac_rt(tdm_ctx *ctx, const float *kp1, const float *kp2, const int *pairs,
                  const int *num_pairs, int max_pairs, double threshold, long max_trials,
                  uint64_t seed, double *rt, int *inliers, int *stats, int *trial_counts,
                  void *stream);

/* estimateRt (core/external/sfm/estimateRt.m, estimateRigidTransform.m:
 * 43-72): least-squares [R|t] with x1 = R*x2 + t over num >= 3 pairs, by
 * Horn's quaternion method in fp64.  x1, x2: device, num*3 floats; rt:
 * device, 12 doubles, row-major 3x4. */
int tdm_estimate_rt(tdm_ctx *ctx, const float *x1, const float *x2, int num, double *rt,
                    void *stream);

/* Alignment ratios of evaluation/geometric-registration/register2Fragments.m
 * :68-69, 90-97: fragment 2 moved by rt (row-major 3x4, x1 = R x2 + t,
 * applied in float with R and t rounded to float, as MATLAB evaluates the
 * double matrix times the single cloud), then ratio[0] = the fraction of
 * fragment-1 points with a moved fragment-2 point closer than radius (0.05
 * there), ratio[1] = the fraction of moved fragment-2 points with a
 * fragment-1 point closer than radius.  Distances are sqrtf((dx*dx + dy*dy)
 * + dz*dz) in float.  pts1, pts2: device xyz floats; ratio: host.
 * Synchronizes `stream`. */
int tdm_alignment_ratio(tdm_ctx *ctx, const float *pts1, int num1, const float *pts2, int num2,
                        const double rt[12], float radius, double ratio[2], void *stream);

/* ---- registration serving ---------------------------------------------------------
 *
 * A request registers fragment 2 onto fragment 1, the pipeline of
 * evaluation/geometric-registration/register2Fragments.m:1-65 (with the
 * descriptors of core/demo.cu): point cloud -> TDF -> descriptors at
 * keypoints -> mutual nearest neighbours -> ransacfitRt.  The server keeps
 * requests' TDF grids, descriptors and matches resident in one device
 * arena of `memory_budget` bytes and admits a queued request when its
 * buffers fit and a slot is free.  Admitted requests run concurrently on
 * their own streams; whenever the descriptor network is free, the keypoints
 * of every request whose TDF grids are ready go into one multi-grid pass
 * (continuous batching), so requests that arrive while others are on the
 * GPU join the next pass.  Results do not depend on what else is in flight:
 * every kernel is deterministic and per-patch results do not depend on the
 * pass they run in.  One scheduler thread issues all GPU work. */
typedef struct tdm_server tdm_server;

typedef struct tdm_server_config {
    size_t memory_budget;     /* device bytes for requests' buffers           */
    int max_inflight;         /* requests admitted at once (streams)          */
    int keypoints;            /* keypoints per fragment when not given        */
    int max_pass_keypoints;   /* keypoints per descriptor pass                */
    float voxel_size;         /* TDF voxel (core/demo.cu:97: 0.01)            */
    float trunc_margin;       /* TDF truncation (core/demo.cu:98: 5 voxels)   */
    int padding;              /* grid padding in voxels (core/demo.cu:99: 15) */
    double ransac_threshold;  /* register2Fragments.m:59: 0.05               */
    long ransac_max_trials;   /* ransac.m:121: 50000                          */
} tdm_server_config;

/* The defaults above; memory_budget 4 GiB, 3 slots (3 request streams and
 * the net's stream fill the 4 hardware queues HIP gives a process by
 * default), 5000 keypoints, passes
 * of up to 20000 keypoints. */
void tdm_server_config_init(tdm_server_config *cfg);

typedef struct tdm_reg_request {
    const float *points1;   /* host, num_points1*3 floats (fragment 1)      */
    int num_points1;
    const float *points2;   /* host, num_points2*3 floats (fragment 2)      */
    int num_points2;
    const int *keypoints1;  /* host point indices, or NULL: chosen from seed */
    const int *keypoints2;
    int num_keypoints;      /* when indices are given (0 = config value)     */
    uint64_t seed;          /* keypoint choice and RANSAC samples            */
} tdm_reg_request;

typedef struct tdm_reg_result {
    double rt[12];       /* row-major [R|t] with x1 = R x2 + t                */
    int status;          /* TDM_OK or the error of this request               */
    int num_matches;     /* mutually nearest keypoint pairs                   */
    int num_inliers;     /* of the kept RANSAC model                          */
    int trials;          /* RANSAC trials run                                 */
    int fallback;        /* 1: fewer than 3 matches, rt = identity            */
    int pass_keypoints;  /* keypoints in the descriptor pass it ran in        */
    double queue_ms;     /* submit -> admitted                                */
    double latency_ms;   /* submit -> result ready                            */
} tdm_reg_result;

/* Server on HIP device `device` using `net` (which it then drives
 * exclusively until destroyed). */
int tdm_server_create(tdm_net *net, int device, const tdm_server_config *cfg,
                      tdm_server **out);
/* Queues a request; the point arrays are copied before return.  Fails with
 * TDM_ERR_NOMEM if the request alone exceeds the memory budget. */
int tdm_server_submit(tdm_server *srv, const tdm_reg_request *req, uint64_t *ticket);
/* Blocks until the request is done and returns its result (once). */
int tdm_server_wait(tdm_server *srv, uint64_t ticket, tdm_reg_result *out);
/* stats[0] peak arena bytes in use, [1] budget, [2] most requests in flight
 * at once, [3] descriptor passes, [4] requests completed, [5] most requests
 * in one pass. */
int tdm_server_stats(tdm_server *srv, long long stats[6]);
/* Finishes the queued and in-flight requests, then frees everything. */
int tdm_server_destroy(tdm_server *srv);

/* ---- file formats (host) --------------------------------------------------- */

/* Binary PLY as read by core/demo.cu:69-93: 7 header lines, the vertex count
 * on line 3, then num*3 little-endian float32.  *pts is freed with tdm_free. */
int tdm_read_ply_points(const char *path, float **pts, int *num);
/* Writes that layout (the header of depth-fusion/utils.hpp:33-41). */
int tdm_write_ply_points(const char *path, const float *pts, int num);

/* <prefix>.keypts.bin: float N then N*3 floats (core/demo.cu:238-244). */
int tdm_write_keypts_bin(const char *path, const float *kp, int num);
int tdm_read_keypts_bin(const char *path, float **kp, int *num);
/* <prefix>.desc.3dmatch.bin: float N, float D, then N*D floats
 * (core/demo.cu:248-255). */
int tdm_write_desc_bin(const char *path, const float *desc, int num, int dim);
int tdm_read_desc_bin(const char *path, float **desc, int *num, int *dim);

/* tsdf.bin / tdf.bin: 8 floats [dx, dy, dz, ox, oy, oz, voxel_size,
 * trunc_margin] then the grid (depth-fusion/demo.cu:196-213,
 * core/demo.cu:257-274). */
int tdm_write_grid_bin(const char *path, const int dims[3], const float origin[3],
                       float voxel_size, float trunc_margin, const float *grid);
int tdm_read_grid_bin(const char *path, int dims[3], float origin[3], float *voxel_size,
                      float *trunc_margin, float **grid);

/* Patch volume file of get3DMatchDescriptors (evaluation/keypoint-matching/
 * get3DMatchDescriptors.cu:11-24): float N, 3 float dims, N*prod(dims)
 * floats.  And feat.bin (:35-64): float N, 4 float dims, N*prod floats. */
int tdm_read_patch_volumes(const char *path, float **data, int *num, int dims[3]);
int tdm_write_feat_bin(const char *path, const float *feat, int num, int dim);

/* M x N whitespace-separated floats (depth-fusion/utils.hpp:60-70). */
int tdm_load_matrix(const char *path, int rows, int cols, float *out);
/* 16-bit grayscale PNG, millimetres -> metres, >6 m -> 0
 * (depth-fusion/utils.hpp:74-86).  depth: host, height*width floats. */
int tdm_read_depth_png(const char *path, int height, int width, float *depth);
/* Raw u16 values of a 16-bit grayscale PNG. *mm freed with tdm_free. */
int tdm_read_png_u16(const char *path, uint16_t **mm, int *height, int *width);
/* depth-fusion/utils.hpp:89-109 and :112-238 (float arithmetic kept). */
int tdm_multiply_matrix(const float a[16], const float b[16], float out[16]);
int tdm_invert_matrix(const float m[16], float out[16]);
/* Surface points |tsdf| < tsdf_thresh && weight > weight_thresh as a binary
 * PLY at origin + idx*voxel_size (depth-fusion/utils.hpp:14-56).  Host grids.
 * Returns the point count in *num_written when non-NULL. */
int tdm_save_surface_ply(const char *path, const int dims[3], float voxel_size,
                         const float origin[3], const float *tsdf, const float *weight,
                         float tsdf_thresh, float weight_thresh, int *num_written);

/* .marvin tensor files (core/marvin.hpp:2581-2611, 2764-2798): a sequence of
 * {u8 type, u32 sizeof, i32 name_len, name, i32 ndims, i32 dims[], data}.
 * Writes float tensors. */
int tdm_marvin_write(const char *path, int num_tensors, const char *const *names,
                     const int *ndims, const int *const *dims, const float *const *data);
/* Number of tensors and total float elements in a .marvin file. */
int tdm_marvin_count(const char *path, int *num_tensors, size_t *num_values);

#ifdef __cplusplus
}
#endif
#endif /* TDM_H */
