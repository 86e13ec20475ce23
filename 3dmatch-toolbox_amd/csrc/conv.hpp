// Internal interface between the network driver (net.cpp) and the HIP
// convolution kernels (conv.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace tdm {

enum class DType { F32 = 0, BF16 = 1, F16 = 2, F8 = 3 };

inline size_t dtype_size(DType t) { return t == DType::F32 ? 4 : t == DType::F8 ? 1 : 2; }

// One 3-D convolution layer in implicit-GEMM form over NDHWC activations:
// rows = output voxels of `batch` cubes, cols = output channels, reduction =
// taps x input channels (laid out tap-major, channel-minor: weights packed
// [N][taps][C]; the staged kernels sum 128-byte channel units outermost).
struct ConvDesc {
    int C = 0;      // input channels (multiple of 32)
    int N = 0;      // output channels (multiple of 64)
    int D = 0;      // input extent
    int OD = 0;     // output extent (before pooling)
    int ksize = 3;  // window (3 or 1)
    bool relu = true;
    bool pool = false;  // fused 2x2x2 stride-2 max pool (OD even)
};

// Whether a layer runs on the LDS-staged kernel, and how many reduction
// steps (the unit split-K divides) its kernel has.
bool conv_uses_lds(DType dt, const ConvDesc &d);
int conv_reduction_steps(DType dt, const ConvDesc &d);

// Launches the convolution for `batch` cubes.  out is NDHWC with the output
// type: float when out_f32, else the storage type `dt`.  `splits` > 1 splits
// the reduction and needs `partial` (splits*rows*N floats); no pooling then.
// `variant` picks the LDS kernel's tile shape (-1: the default); every
// variant gives bit-identical results.
// F8 (e4m3 activations and weights, fp32 accumulation): the result of
// output channel n is acc * colscale[n] + bias[n] (colscale = weight scale x
// input activation scale), and fp8 outputs store value * oscale.  Other types
// take colscale = NULL and oscale = 1.
int launch_conv(DType dt, bool out_f32, const ConvDesc &d, int batch, const void *in,
                const void *w, const float *bias, void *out, int splits, float *partial,
                int variant, hipStream_t s, const float *colscale = nullptr, float oscale = 1.0f);

// Tile variants of the LDS kernel (0 .. kConvVariants-1): whether one fits
// a layer (its N is a multiple of the tile's BN), the default, and the tile.
// Variants 14-16 are the pooled tap-group kernels (16-bit pooled 3x3x3
// layers).
constexpr int kConvVariants = 17;
bool conv_variant_valid(DType dt, int v, const ConvDesc &d);
int conv_default_variant(DType dt, const ConvDesc &d);
void conv_variant_tile(int v, int *bm, int *bn);

// One TDF grid of a multi-grid pass (device table entry; the layout of
// tdm_grid_ref).
struct GridRef {
    const float *tdf;
    int dims[3];
    int reserved;
};

// Where the first (single-channel) convolution reads its input.
struct VolumeSource {
    const float *base = nullptr;     // patches [batch][P^3] or a TDF grid
    const int *kp = nullptr;         // grid modes: keypoint grid coords [batch][3]
    int dims[3] = {0, 0, 0};         // single-grid mode: grid dims
    const GridRef *grids = nullptr;  // multi-grid mode: device table of grids
    const int *gid = nullptr;        // multi-grid mode: grid of each patch [batch]
    int P = 30;                      // patch extent
};

// First convolution: 1 input channel, N output channels (multiple of 64),
// window 3, fused bias + ReLU; output NDHWC of type dt (float if out_f32).
// w is packed [N][KP] with KP = 28 (F32) or 32 (BF16) taps, zero padded.
// For F8 the first convolution computes in fp16 (weights packed as F16) and
// stores fp8 outputs scaled by oscale.
int launch_conv1(DType dt, bool out_f32, const VolumeSource &src, int N, bool relu, int batch,
                 const void *w, const float *bias, void *out, hipStream_t s,
                 float oscale = 1.0f);

// Taps per output channel in the packed weights of the first convolution.
inline int conv1_taps_padded(DType t) { return t == DType::F32 ? 28 : 32; }

// fp8 e4m3 helpers: per-row quantization of packed float weights (row r
// scaled by inv_scale[r]) on the device, and the largest |x| of n floats as
// float bits (atomicMax into *amax, which must start at 0).
int launch_quantize_fp8(const float *in, long rows, long per_row, const float *inv_scale,
                        void *out, hipStream_t s);
int launch_absmax(const float *x, long n, unsigned *amax, hipStream_t s);
// Largest finite e4m3 value.
constexpr float kFp8Max = 448.0f;

// NDHWC (type dt or float) -> NCDHW float, for per-layer outputs.
// F8 inputs are multiplied by `scale` (the activation scale).
int launch_to_ncdhw(DType dt, bool in_f32, const void *in, int batch, int C, int E, float *out,
                    hipStream_t s, float scale = 1.0f);

// Max pool window 2 stride 2 over NDHWC when it cannot be fused:
// out extent 1 + ceil((E - 2)/2), windows clipped.
int launch_maxpool(DType dt, const void *in, int batch, int C, int E, void *out, hipStream_t s);

// Host float -> bf16 / IEEE half bits, round to nearest even (half: overflow
// to infinity, gradual underflow; the rounding of cpu_float2half,
// core/marvin.hpp:467-521).
uint16_t float_to_bf16(float f);
uint16_t float_to_f16(float f);

}  // namespace tdm
