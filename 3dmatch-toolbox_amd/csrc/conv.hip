// 3-D convolutions of the 3DMatch descriptor network on gfx950 MFMA.
//
// Each convolution (core/marvin.hpp:4635-4709, cuDNN cross-correlation,
// stride 1, no padding) is an implicit GEMM over NDHWC activations:
//   rows  m = output voxels (patch, z, y, x)
//   cols  n = output channels
//   k     = (tap, input channel), laid out tap-major, so for a fixed tap the
//           k-run of a row is the contiguous channel vector of one input
//           voxel (the staged kernels sum it in the order of KWalk).
// Weights are packed [n][tap][c], so both operands are k-contiguous and a lane
// loads its MFMA fragment with one 16-byte load:
//   fp32  v_mfma_f32_16x16x4_f32: a 16-byte load holds 4 consecutive k; the
//         four MFMA k-steps of a chunk consume them in turn;
//   bf16  v_mfma_f32_16x16x32_bf16: a 16-byte load is the lane's 8 k.
// Within a 16-row block lane l supplies row l%16 and k-group l/16; any
// assignment of actual k to (k-group, step) is correct as long as the A and
// B fragments use the same one, which they do by construction.
//
// A wave owns a 64x64 output tile (4x4 MFMA blocks, 64 accumulator VGPRs);
// four waves form a workgroup stacked along m and/or n.  The loop over
// (tap, channel chunk) prefetches the next chunk's fragments while the MFMAs
// of the current one run.  Bias and ReLU (core/marvin.hpp:4655-4662,
// 5483-5493) are fused into the epilogue; for conv2 the 2x2x2/2 max pool
// (core/marvin.hpp:5578-5588) is fused too, by ordering rows so that each
// aligned group of 8 rows is one pool window (rows 4g..4g+3 of a 16-row block
// sit in one lane; the partner group is 16 lanes away).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "common.hpp"
#include "conv.hpp"

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t bf16_t;
typedef _Float16 f16_t;
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef unsigned char fp8_t;  // e4m3, the encoding of gfx950's fp8 conversions and MFMA

template <typename T>
struct Kc;
template <>
struct Kc<f16_t> {
    static constexpr int value = 32;
};
template <>
struct Kc<float> {
    static constexpr int value = 16;  // k per 16-byte lane load x 4 k-groups
};
template <>
struct Kc<bf16_t> {
    static constexpr int value = 32;
};

__device__ inline uint4 ld16(const void *p) { return *reinterpret_cast<const uint4 *>(p); }

__device__ inline float comp(const uint4 &v, int s) {
    unsigned u = s == 0 ? v.x : s == 1 ? v.y : s == 2 ? v.z : v.w;
    return __uint_as_float(u);
}

template <typename T>
struct Mfma;

// One 16-byte fragment per lane for each of IB row blocks and JB column
// blocks: fp32 runs 4 k-steps (one per dword), bf16 one 32-k step.
template <>
struct Mfma<float> {
    template <int IB, int JB>
    __device__ static inline void chunk(floatx4 (&acc)[IB][JB], const uint4 (&a)[IB],
                                        const uint4 (&b)[JB]) {
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < IB; ++i)
#pragma unroll
                for (int j = 0; j < JB; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(comp(a[i], s), comp(b[j], s),
                                                                     acc[i][j], 0, 0, 0);
    }
};

template <>
struct Mfma<bf16_t> {
    template <int IB, int JB>
    __device__ static inline void chunk(floatx4 (&acc)[IB][JB], const uint4 (&a)[IB],
                                        const uint4 (&b)[JB]) {
#pragma unroll
        for (int i = 0; i < IB; ++i)
#pragma unroll
            for (int j = 0; j < JB; ++j) {
                bf16x8 av, bv;
                __builtin_memcpy(&av, &a[i], 16);
                __builtin_memcpy(&bv, &b[j], 16);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[i][j], 0, 0, 0);
            }
    }
};

// fp8: a 16-byte fragment is two 8-byte k-groups, one v_mfma_f32_16x16x32_fp8
// step each (lane group g supplies 8 consecutive k of a 32-k step).
template <>
struct Mfma<fp8_t> {
    template <int IB, int JB>
    __device__ static inline void chunk(floatx4 (&acc)[IB][JB], const uint4 (&a)[IB],
                                        const uint4 (&b)[JB]) {
#pragma unroll
        for (int hh = 0; hh < 2; ++hh)
#pragma unroll
            for (int i = 0; i < IB; ++i)
#pragma unroll
                for (int j = 0; j < JB; ++j) {
                    const long av = hh == 0 ? (long)(((unsigned long)a[i].y << 32) | a[i].x)
                                            : (long)(((unsigned long)a[i].w << 32) | a[i].z);
                    const long bv = hh == 0 ? (long)(((unsigned long)b[j].y << 32) | b[j].x)
                                            : (long)(((unsigned long)b[j].w << 32) | b[j].z);
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_fp8_fp8(av, bv, acc[i][j], 0, 0, 0);
                }
    }
};

template <>
struct Mfma<f16_t> {
    template <int IB, int JB>
    __device__ static inline void chunk(floatx4 (&acc)[IB][JB], const uint4 (&a)[IB],
                                        const uint4 (&b)[JB]) {
#pragma unroll
        for (int i = 0; i < IB; ++i)
#pragma unroll
            for (int j = 0; j < JB; ++j) {
                halfx8 av, bv;
                __builtin_memcpy(&av, &a[i], 16);
                __builtin_memcpy(&bv, &b[j], 16);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(av, bv, acc[i][j], 0, 0, 0);
            }
    }
};

// Round to nearest even: gfx950 converts in one instruction
// (v_cvt_pk_bf16_f32), bit-identical to the integer rounding of
// float_to_bf16 on every non-NaN input (checked over all 2^32 patterns on
// the device); a NaN stays a NaN.
__device__ inline bf16_t to_bf16(float f) {
    const __bf16 h = (__bf16)f;
    bf16_t u;
    __builtin_memcpy(&u, &h, 2);
    return u;
}

__device__ inline float from_bf16(bf16_t h) { return __uint_as_float((unsigned)h << 16); }

template <typename OT>
__device__ inline void store(OT *p, float v);
template <>
__device__ inline void store<float>(float *p, float v) {
    *p = v;
}
template <>
__device__ inline void store<bf16_t>(bf16_t *p, float v) {
    *p = to_bf16(v);
}
template <>
__device__ inline void store<f16_t>(f16_t *p, float v) {
    *p = (f16_t)v;  // round to nearest even, overflow to inf
}
// Saturating round-to-nearest-even conversion to e4m3 (the caller has
// applied the activation scale).
__device__ inline fp8_t to_fp8(float v) {
    v = fminf(fmaxf(v, -tdm::kFp8Max), tdm::kFp8Max);
    return (fp8_t)(__builtin_amdgcn_cvt_pk_fp8_f32(v, v, 0, false) & 0xff);
}
__device__ inline float from_fp8(fp8_t b) { return __builtin_amdgcn_cvt_f32_fp8((int)b, 0); }
template <>
__device__ inline void store<fp8_t>(fp8_t *p, float v) {
    *p = to_fp8(v);
}

// Four consecutive values in one store (p aligned to 4 values).
template <typename OT>
__device__ inline void store4(OT *p, const float (&v)[4]) {
    if constexpr (sizeof(OT) == 4) {
        *reinterpret_cast<float4 *>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else {
        OT o[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) store<OT>(&o[r], v[r]);
        if constexpr (sizeof(OT) == 2) {
            uint2 u;
            __builtin_memcpy(&u, o, 8);
            *reinterpret_cast<uint2 *>(p) = u;
        } else {
            unsigned u;
            __builtin_memcpy(&u, o, 4);
            *reinterpret_cast<unsigned *>(p) = u;
        }
    }
}

// 16-bit storage bits of a float for the MFMA operand types.
template <typename T>
__device__ inline unsigned short bits16(float v);
template <>
__device__ inline unsigned short bits16<bf16_t>(float v) {
    return to_bf16(v);
}
template <>
__device__ inline unsigned short bits16<f16_t>(float v) {
    f16_t h = (f16_t)v;
    unsigned short u;
    __builtin_memcpy(&u, &h, 2);
    return u;
}

template <typename T>
constexpr tdm::DType dtype_of() {
    return sizeof(T) == 4   ? tdm::DType::F32
           : sizeof(T) == 1 ? tdm::DType::F8
                            : (__is_same(T, f16_t) ? tdm::DType::F16 : tdm::DType::BF16);
}

struct ConvArgs {
    const void *in;
    const void *w;
    const float *bias;
    void *out;
    float *partial;
    long M;  // rows
    int N, C, D, OD, ksize, taps, splits;
    int relu;
    const float *colscale;  // fp8: per output channel dequantization scale (else NULL)
    float oscale;           // fp8 outputs: multiplier before conversion (else 1)
    // Row decode when M < 2^31 (small_rows): n / E = umulhi(n, rdiv_mul) >>
    // rdiv_sh for n < 2^31, E the extent rows are counted in (OD, or OD/2
    // with a fused pool; rdiv_mul = 0: E == 1); else 64-bit divisions.
    unsigned rdiv_mul;
    int rdiv_sh;
    int small_rows;
    // the same for E + 1 (the pooled group kernel's slots per window line)
    unsigned sdiv_mul;
    int sdiv_sh;
};

// Multiply-high division by d for n < 2^31 (Granlund and Montgomery's
// round-up method: l = ceil(log2 d), m = ceil(2^(31 + l) / d) < 2^32, then
// n / d = (n * m) >> (31 + l); exact because m * d - 2^(31 + l) < d <= 2^l).
void fast_div(int d, unsigned *mul, int *sh) {
    if (d == 1) {
        *mul = 0u;
        *sh = 0;
        return;
    }
    int l = 0;
    while ((1L << l) < d) ++l;
    const unsigned long long num = 1ull << (31 + l);
    *mul = (unsigned)((num + (unsigned long long)d - 1) / (unsigned long long)d);
    *sh = l - 1;
}

__device__ inline unsigned row_div(unsigned n, const ConvArgs &p) {
    return p.rdiv_mul ? __umulhi(n, p.rdiv_mul) >> p.rdiv_sh : n;
}
__device__ inline unsigned slot_div(unsigned n, const ConvArgs &p) {
    return __umulhi(n, p.sdiv_mul) >> p.sdiv_sh;
}

// Workgroup -> tile order that keeps neighbouring tiles on one XCD.  The
// dispatcher hands workgroup i (linear id, x fastest) to XCD i % 8, so
// consecutive tiles -- which share input voxels through the 3x3x3 halo --
// would land in eight different L2s.  Here XCD x gets one contiguous run of
// the tile sequence instead.
constexpr unsigned kXcds = 8;
__device__ inline void xcd_tile(unsigned &tx, unsigned &ty, unsigned &tz) {
    const unsigned nx = gridDim.x, ny = gridDim.y;
    const unsigned total = nx * ny * gridDim.z;
    const unsigned bid = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
    const unsigned per = total / kXcds, rem = total % kXcds;
    const unsigned x = bid % kXcds, idx = bid / kXcds;
    const unsigned t = x * per + min(x, rem) + idx;
    tx = t % nx;
    ty = (t / nx) % ny;
    tz = t / (nx * ny);
}

// Input voxel index of output row m: rows are (b, z, y, x) in order, or, with
// POOL, (b, pz, py, px, window) with the 2x2x2 window innermost.
template <bool POOL>
__device__ inline long row_voxel(long m, const ConvArgs &p) {
    const int D = p.D, OD = p.OD;
    int ox, oy, oz;
    long b;
    if (p.small_rows) {
        // rows fit 31 bits: the three divisions by the extent are multiply-highs
        const unsigned E = POOL ? (unsigned)(OD >> 1) : (unsigned)OD;
        const unsigned q0 = POOL ? (unsigned)(m >> 3) : (unsigned)m;
        const unsigned q1 = row_div(q0, p), q2 = row_div(q1, p), q3 = row_div(q2, p);
        const int cx = (int)(q0 - q1 * E), cy = (int)(q1 - q2 * E), cz = (int)(q2 - q3 * E);
        b = q3;
        if (POOL) {
            const int w = (int)(m & 7);
            ox = 2 * cx + (w & 1);
            oy = 2 * cy + ((w >> 1) & 1);
            oz = 2 * cz + (w >> 2);
        } else {
            ox = cx;
            oy = cy;
            oz = cz;
        }
    } else if (POOL) {
        const int w = (int)(m & 7);
        long q = m >> 3;
        const int PD = OD >> 1;
        const int px = (int)(q % PD);
        q /= PD;
        const int py = (int)(q % PD);
        q /= PD;
        const int pz = (int)(q % PD);
        b = q / PD;
        ox = 2 * px + (w & 1);
        oy = 2 * py + ((w >> 1) & 1);
        oz = 2 * pz + (w >> 2);
    } else {
        ox = (int)(m % OD);
        long q = m / OD;
        oy = (int)(q % OD);
        q /= OD;
        oz = (int)(q % OD);
        b = q / OD;
    }
    return ((b * D + oz) * D + oy) * (long)D + ox;
}

template <typename T, typename OT, int WAVES_N, bool POOL, bool SPLIT>
__global__ __launch_bounds__(256) void conv_kernel(ConvArgs p) {
    constexpr int KC = Kc<T>::value;
    constexpr int WAVES_M = 4 / WAVES_N;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = wave / WAVES_N, wn = wave % WAVES_N;
    unsigned tx, ty, tz;
    xcd_tile(tx, ty, tz);
    const long m0 = ((long)tx * WAVES_M + wm) * 64;
    const int n0 = (ty * WAVES_N + wn) * 64;
    if (m0 >= p.M) return;  // no barriers below
    const int r16 = lane & 15, g = lane >> 4;
    const int lanek = g * (KC / 4);
    const T *__restrict__ in = static_cast<const T *>(p.in);
    const T *__restrict__ w = static_cast<const T *>(p.w);
    const long K = (long)p.taps * p.C;

    const T *ap[4];
    const T *bp[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        long m = min(m0 + 16 * i + r16, p.M - 1);
        ap[i] = in + row_voxel<POOL>(m, p) * p.C + lanek;
        bp[i] = w + (long)(n0 + 16 * i + r16) * K + lanek;
    }

    const int cpt = p.C / KC;
    const int nq = p.taps * cpt;
    int q0 = 0, q1 = nq;
    if (SPLIT) {
        q0 = (int)((long)nq * tz / p.splits);
        q1 = (int)((long)nq * (tz + 1) / p.splits);
    }
    const int D = p.D, C = p.C, ks = p.ksize;
    auto offsets = [&](int q, long &aoff, long &boff) {
        const int tap = q / cpt;
        const int c0 = (q - tap * cpt) * KC;
        int kz = 0, ky = 0, kx = 0;
        if (ks == 3) {
            kz = tap / 9;
            ky = (tap / 3) % 3;
            kx = tap % 3;
        }
        aoff = (long)((kz * D + ky) * D + kx) * C + c0;
        boff = (long)tap * C + c0;
    };

    floatx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    if (q0 < q1) {
        uint4 a[4], b[4];
        long aoff, boff;
        offsets(q0, aoff, boff);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a[i] = ld16(ap[i] + aoff);
            b[i] = ld16(bp[i] + boff);
        }
        for (int q = q0; q < q1; ++q) {
            uint4 an[4], bn[4];
            const bool more = q + 1 < q1;
            if (more) {
                offsets(q + 1, aoff, boff);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    an[i] = ld16(ap[i] + aoff);
                    bn[i] = ld16(bp[i] + boff);
                }
            }
            Mfma<T>::chunk(acc, a, b);
            if (more) {
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    a[i] = an[i];
                    b[i] = bn[i];
                }
            }
        }
    }

    // Epilogue.  Lane holds rows m0+16i+4g+r (r = 0..3) of column n0+16j+r16.
    if (SPLIT) {
        float *__restrict__ part = p.partial + (long)tz * p.M * p.N;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int col = n0 + 16 * j + r16;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const long row = m0 + 16 * i + 4 * g + r;
                    if (row < p.M) part[row * p.N + col] = acc[i][j][r];
                }
            }
        return;
    }
    OT *__restrict__ out = static_cast<OT *>(p.out);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int col = n0 + 16 * j + r16;
        const float bv = p.bias[col];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = acc[i][j][r] + bv;
                if (p.relu) v[r] = v[r] > 0.0f ? v[r] : 0.0f;
            }
            if (POOL) {
                float mx = fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3]));
                mx = fmaxf(mx, __shfl_xor(mx, 16));
                const long row = m0 + 16 * i + 4 * g;
                if ((g & 1) == 0 && row < p.M) store<OT>(out + (row >> 3) * p.N + col, mx);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const long row = m0 + 16 * i + 4 * g + r;
                    if (row < p.M) store<OT>(out + row * p.N + col, v[r]);
                }
            }
        }
    }
}

// Epilogue of a wave tile of IB x JB 16x16 blocks at (m0, n0): the lane holds
// rows m0+16i+4g+r (r = 0..3) of column n0+16j+r16.  Split-K writes float
// partials; otherwise bias, ReLU, the fused 2x2x2 pool and the store.
template <typename OT, int IB, int JB, bool POOL, bool SPLIT>
__device__ inline void conv_epilogue(const floatx4 (&acc)[IB][JB], const ConvArgs &p, long m0,
                                     int n0, int lane, unsigned split) {
    const int r16 = lane & 15, g = lane >> 4;
    if (SPLIT) {
        float *__restrict__ part = p.partial + (long)split * p.M * p.N;
#pragma unroll
        for (int i = 0; i < IB; ++i)
#pragma unroll
            for (int j = 0; j < JB; ++j) {
                const int col = n0 + 16 * j + r16;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const long row = m0 + 16 * i + 4 * g + r;
                    if (row < p.M) part[row * p.N + col] = acc[i][j][r];
                }
            }
        return;
    }
    OT *__restrict__ out = static_cast<OT *>(p.out);
    constexpr bool Q8 = sizeof(OT) == 1;
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        const int col = n0 + 16 * j + r16;
        const float bv = p.bias[col];
        const float cs = p.colscale ? p.colscale[col] : 1.0f;
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = p.colscale ? acc[i][j][r] * cs + bv : acc[i][j][r] + bv;
                if (p.relu) v[r] = v[r] > 0.0f ? v[r] : 0.0f;
                if (Q8) v[r] *= p.oscale;
            }
            if (POOL) {
                float mx = fmaxf(fmaxf(v[0], v[1]), fmaxf(v[2], v[3]));
                mx = fmaxf(mx, __shfl_xor(mx, 16));
                const long row = m0 + 16 * i + 4 * g;
                if ((g & 1) == 0 && row < p.M) store<OT>(out + (row >> 3) * p.N + col, mx);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const long row = m0 + 16 * i + 4 * g + r;
                    if (row < p.M) store<OT>(out + row * p.N + col, v[r]);
                }
            }
        }
    }
}

// LDS-staged implicit GEMM.  A workgroup computes a BM x BN tile with four
// waves in a WMW x (4/WMW) grid, each wave owning a WTM x WTN sub-tile
// (WTM/16 x WTN/16 MFMA blocks).  Each stage is 128 bytes of k per row -- one
// tap and 32 (fp32) or 64 (16-bit) input channels: the 256 threads copy the
// stage's BM + BN rows from global memory into registers one stage ahead,
// then into the other half of a double-buffered LDS tile (rows padded to 144
// bytes against bank conflicts); the waves read their fragments from LDS.
// Larger wave tiles read fewer LDS bytes per MFMA; larger workgroup tiles
// fetch each A row for more columns.  Every variant feeds each output the
// same k in the same order (lane group g reads bytes 16g + 64h of a stage),
// so all variants give bit-identical results and can be picked by timing.
// A stage is 128 or 64 bytes of k per row.  Both stage sizes walk k in the
// same order (64-byte halves in sequence).
//
// LDS rows are unpadded and the 16-byte pieces of a row are XOR-swizzled:
// piece c of row r sits at piece c ^ swz(r).  A fragment read has lane l take
// piece g (+ 4h) of row l % 16; a staging store has consecutive threads take
// consecutive pieces of consecutive rows.  With the swizzle, the 16 lanes of
// a read land on 16 different 16-byte bank slots of a 256-byte bank line, and
// any 8 consecutive lanes on 8 different slots of a 128-byte line, for every
// g, h and block; with the earlier 16-byte row padding (80-byte pitch) the
// counters showed ~5 cycles of bank conflict per LDS instruction.
template <int STAGE>
__device__ inline int swz(int r) {
    return STAGE == 64 ? ((r >> 1) ^ (r >> 3)) & 3 : ((r >> 1) ^ (r << 2)) & 7;
}

// The order of the reduction, shared by every staged variant so that all of
// them sum each output's k in the same order.  k runs over units of 128 bytes
// of input channels (64 bytes when the channel vector is not a multiple of
// 128 bytes), the unit outermost, then the tap, then the stage's half of the
// unit: a tile reads one channel unit of its input rows for all 27 taps
// before the next, so the tiles in flight on an XCD touch 128 bytes of each
// input voxel at a time rather than the whole channel vector, and their
// working set stays within the XCD's 4 MB L2 (with the whole vector, tap
// after tap, fp32 layers re-read their input from beyond L2 ~17 times over).
// Split-K boundaries fall on unit edges.
struct KWalk {
    int taps = 1, ub = 128, half_n = 1, tap = 0, unit = 0, half = 0, csz = 4, stage = 128;
    // returns the number of stages
    __device__ int init(int C, int esz, int STAGE, int ntaps) {
        taps = ntaps;
        csz = esz;
        stage = STAGE;
        ub = (C * esz) % 128 == 0 ? 128 : 64;
        half_n = ub / STAGE;
        return ntaps * (C * esz / ub) * half_n;
    }
    __device__ void split(int ns, unsigned tz, int splits, int *s0, int *s1) const {
        const int nc = ns / half_n;  // units
        *s0 = (int)((long)nc * tz / splits) * half_n;
        *s1 = tz + 1 == (unsigned)splits ? ns : (int)((long)nc * (tz + 1) / splits) * half_n;
    }
    __device__ void seek(int s) {
        const int u = s / half_n;
        half = s - u * half_n;
        unit = u / taps;
        tap = u - unit * taps;
    }
    __device__ int c0() const { return (unit * ub + half * stage) / csz; }
    __device__ void next() {
        if (++half == half_n) {
            half = 0;
            if (++tap == taps) {
                tap = 0;
                ++unit;
            }
        }
    }
};

// Byte offsets of a stage's input rows and weight rows from the tile's bases,
// advanced stage by stage in the order of KWalk with scalar adds only: the
// input offset of (unit, tap (kz, ky, kx), half) is kz*sz + ky*sy + kx*sx +
// unit*ub + half*stage, the weight offset tap*tb + unit*ub + half*stage.
// Offsets wrap modulo 2^32 in between; every stage's own offset is below 4 GB.
struct CopyWalk {
    int half = 0, kx = 0, ky = 0, kz = 0, half_n = 1, ks = 3;
    unsigned a = 0, b = 0, sx = 0, sy = 0, sz = 0, tb = 0, ub = 128, stage = 128, wrap_b = 0;
    __device__ void init(const KWalk &kw, int C, int D, int esz, int ksize) {
        half_n = kw.half_n;
        ks = ksize;
        stage = (unsigned)kw.stage;
        ub = (unsigned)kw.ub;
        sx = (unsigned)(C * esz);
        sy = sx * (unsigned)D;
        sz = sy * (unsigned)D;
        tb = sx;
        wrap_b = ub - (unsigned)kw.taps * tb;
        half = kw.half;
        if (ks == 3) {
            kz = kw.tap / 9;
            ky = (kw.tap / 3) % 3;
            kx = kw.tap % 3;
        }
        a = kz * sz + ky * sy + kx * sx + (unsigned)kw.unit * ub + (unsigned)half * stage;
        b = (unsigned)kw.tap * tb + (unsigned)kw.unit * ub + (unsigned)half * stage;
    }
    __device__ void next() {
        if (++half < half_n) {
            a += stage;
            b += stage;
            return;
        }
        half = 0;
        a -= (unsigned)(half_n - 1) * stage;
        b -= (unsigned)(half_n - 1) * stage;
        b += tb;
        if (++kx < ks) {
            a += sx;
            return;
        }
        kx = 0;
        a -= (unsigned)(ks - 1) * sx;
        if (++ky < ks) {
            a += sy;
            return;
        }
        ky = 0;
        a -= (unsigned)(ks - 1) * sy;
        if (++kz < ks) {
            a += sz;
            return;
        }
        kz = 0;
        a -= (unsigned)(ks - 1) * sz;
        a += ub;
        b += wrap_b;
    }
};

template <int WMW, int WTM, int WTN, int STAGE = 128>
struct Tile {
    static constexpr int WNW = 4 / WMW;
    static constexpr int BM = WMW * WTM, BN = WNW * WTN;
    static constexpr int ROW = STAGE;
    static constexpr int LDS = 2 * (BM + BN) * ROW;
    static constexpr int OCC = LDS <= 80 * 1024 ? 2 : 1;  // workgroups per CU
};

template <typename T, typename OT, int WMW, int WTM, int WTN, int STAGE, bool POOL, bool SPLIT>
__global__ __launch_bounds__(256, (Tile<WMW, WTM, WTN, STAGE>::OCC)) void conv_lds_kernel(ConvArgs p) {
    using TL = Tile<WMW, WTM, WTN, STAGE>;
    constexpr int BM = TL::BM, BN = TL::BN, WNW = TL::WNW;
    constexpr int kLdsRow = TL::ROW;
    constexpr int PCS = STAGE / 16;                   // 16-byte pieces per staged row
    constexpr int HALVES = STAGE / 64;
    constexpr int IB = WTM / 16, JB = WTN / 16;
    constexpr int A_PT = BM * PCS / 256;              // 16-byte pieces per thread
    constexpr int B_PT = BN * PCS / 256;
    __shared__ __attribute__((aligned(16))) unsigned char lds[2][(BM + BN) * kLdsRow];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave / WNW, wn = wave % WNW;
    unsigned tx, ty, tz;
    xcd_tile(tx, ty, tz);
    const long mb = (long)tx * BM;
    const int nb = ty * BN;
    const int C = p.C, D = p.D, ks = p.ksize;
    const long K = (long)p.taps * C;
    const unsigned char *__restrict__ in = static_cast<const unsigned char *>(p.in);
    const unsigned char *__restrict__ w = static_cast<const unsigned char *>(p.w);

    const unsigned char *a_src[A_PT];
    int a_dst[A_PT];
#pragma unroll
    for (int j = 0; j < A_PT; ++j) {
        const int idx = tid + 256 * j, row = idx / PCS, pc = idx % PCS;
        const long m = min(mb + row, p.M - 1);
        a_src[j] = in + (row_voxel<POOL>(m, p) * C) * (long)sizeof(T) + pc * 16;
        a_dst[j] = row * kLdsRow + ((pc ^ swz<STAGE>(row)) << 4);
    }
    const unsigned char *b_src[B_PT];
    int b_dst[B_PT];
#pragma unroll
    for (int j = 0; j < B_PT; ++j) {
        const int idx = tid + 256 * j, col = idx / PCS, pc = idx % PCS;
        b_src[j] = w + ((long)(nb + col) * K) * (long)sizeof(T) + pc * 16;
        b_dst[j] = (BM + col) * kLdsRow + ((pc ^ swz<STAGE>(BM + col)) << 4);
    }

    KWalk kw;
    const int ns = kw.init(C, (int)sizeof(T), STAGE, p.taps);
    int s0 = 0, s1 = ns;
    if (SPLIT) kw.split(ns, tz, p.splits, &s0, &s1);
    // Stages are fetched strictly in order s0, s0+1, ...: the tap and channel
    // offset advance with each fetch instead of being divided out of s.
    kw.seek(s0);
    // Two register sets: the loads of stage s+2 are issued before the MFMAs
    // of stage s, and stage s+1 (loaded one stage earlier) goes to LDS after
    // them, so a global load has two stages of compute to arrive.
    uint4 ra0[A_PT], rb0[B_PT], ra1[A_PT], rb1[B_PT];
    auto fetch = [&](uint4 (&ra)[A_PT], uint4 (&rb)[B_PT]) {
        const int tap = kw.tap, c0 = kw.c0();
        int kz = 0, ky = 0, kx = 0;
        if (ks == 3) {
            kz = tap / 9;
            ky = (tap / 3) % 3;
            kx = tap % 3;
        }
        const long aoff = ((long)((kz * D + ky) * D + kx) * C + c0) * (long)sizeof(T);
        const long boff = ((long)tap * C + c0) * (long)sizeof(T);
#pragma unroll
        for (int j = 0; j < A_PT; ++j) ra[j] = ld16(a_src[j] + aoff);
#pragma unroll
        for (int j = 0; j < B_PT; ++j) rb[j] = ld16(b_src[j] + boff);
        kw.next();
    };
    auto stash = [&](int buf, const uint4 (&ra)[A_PT], const uint4 (&rb)[B_PT]) {
#pragma unroll
        for (int j = 0; j < A_PT; ++j) *reinterpret_cast<uint4 *>(&lds[buf][a_dst[j]]) = ra[j];
#pragma unroll
        for (int j = 0; j < B_PT; ++j) *reinterpret_cast<uint4 *>(&lds[buf][b_dst[j]]) = rb[j];
    };

    floatx4 acc[IB][JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int r16 = lane & 15, g = lane >> 4;
    // Per-lane byte offsets of the fragment pieces within a stage buffer.
    int a_off[IB][HALVES], b_off[JB][HALVES];
#pragma unroll
    for (int i = 0; i < IB; ++i) {
        const int row = wm * WTM + 16 * i + r16;
#pragma unroll
        for (int h = 0; h < HALVES; ++h)
            a_off[i][h] = row * kLdsRow + (((g + 4 * h) ^ swz<STAGE>(row)) << 4);
    }
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        const int row = BM + wn * WTN + 16 * j + r16;
#pragma unroll
        for (int h = 0; h < HALVES; ++h)
            b_off[j][h] = row * kLdsRow + (((g + 4 * h) ^ swz<STAGE>(row)) << 4);
    }
    auto compute = [&](int buf) {
        const unsigned char *L = lds[buf];
#pragma unroll
        for (int h = 0; h < HALVES; ++h) {
            uint4 a[IB], b[JB];
#pragma unroll
            for (int i = 0; i < IB; ++i) a[i] = *reinterpret_cast<const uint4 *>(L + a_off[i][h]);
#pragma unroll
            for (int j = 0; j < JB; ++j) b[j] = *reinterpret_cast<const uint4 *>(L + b_off[j][h]);
            Mfma<T>::chunk(acc, a, b);
        }
    };
    // stage s from LDS; `held` = stage s+1 in registers, `next` receives s+2
    auto step = [&](int s, const uint4 (&ha)[A_PT], const uint4 (&hb)[B_PT], uint4 (&na)[A_PT],
                    uint4 (&nb)[B_PT]) {
        const int buf = (s - s0) & 1;
        if (s + 2 < s1) fetch(na, nb);
        compute(buf);
        if (s + 1 < s1) stash(buf ^ 1, ha, hb);
        __syncthreads();
    };
    if (s0 < s1) {
        fetch(ra0, rb0);
        stash(0, ra0, rb0);
        if (s0 + 1 < s1) fetch(ra1, rb1);
        __syncthreads();
        for (int s = s0; s < s1; s += 2) {
            step(s, ra1, rb1, ra0, rb0);
            if (s + 1 < s1) step(s + 1, ra0, rb0, ra1, rb1);
        }
    }
    conv_epilogue<OT, IB, JB, POOL, SPLIT>(acc, p, mb + wm * WTM, nb + wn * WTN, lane, tz);
}

// One wave copies 64 16-byte pieces into consecutive LDS slots at `lds`: lane
// l reads base + soff + voff(l) with a raw buffer load (the base in the
// descriptor, the stage's offset in a scalar, the lane's in a 32-bit vector
// register).  Descriptor word 3 0x00020000 is the gfx9 raw-buffer setting (32-bit
// data format); 2^32 - 1 records leave every offset below 4 GB unclamped.
__device__ inline void copy16_to_lds(const unsigned char *base, unsigned soff, unsigned voff,
                                     unsigned char *lds) {
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, -1, 0x00020000),
        (__attribute__((address_space(3))) void *)lds, 16, voff, soff, 0, 0);
}

// s_waitcnt with vmcnt(n) and no wait on the other counters (gfx9 encoding:
// vmcnt bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int N>
__device__ inline void wait_vmcnt() {
    static_assert(N >= 0 && N < 64, "vmcnt");
    __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}

// Direct-to-LDS staging.  The same tiles, swizzled layout and k order as
// conv_lds_kernel, but each stage is copied from global memory into LDS by
// global_load_lds_dwordx4: one wave instruction writes 64 consecutive 16-byte
// slots, each lane fetching the global piece that belongs at its slot, so no
// data passes through VGPRs and no ds_write is issued (the staged form spends
// about a third of its LDS instructions on those writes).  A ring of NST
// stage buffers keeps NST-1 stages of loads in flight: at each stage a wave
// waits for its own loads of that stage (vmcnt), the barrier makes everyone's
// visible and frees the buffer read at the previous stage, which then
// receives the loads of stage s+NST-1.  The buffers are separate LDS arrays
// indexed at compile time (the loop is unrolled NST times), so the compiler
// can see that the reads of one stage do not depend on the copies in flight
// into another.  The launch bound of two workgroups per CU holds the
// accumulators in VGPRs (without it the compiler moves them to AGPRs, which
// measured 6-8 % slower on the 16-bit layers).
template <typename T, typename OT, int WMW, int WTM, int WTN, int STAGE, int NST, bool POOL,
          bool SPLIT>
__global__ __launch_bounds__(256, 2) void conv_dma_kernel(ConvArgs p) {
    using TL = Tile<WMW, WTM, WTN, STAGE>;
    constexpr int BM = TL::BM, BN = TL::BN, WNW = TL::WNW;
    constexpr int PCS = STAGE / 16, HALVES = STAGE / 64;
    constexpr int IB = WTM / 16, JB = WTN / 16;
    constexpr int SLOTS = (BM + BN) * PCS;  // 16-byte slots per stage
    constexpr int J = SLOTS / 256;          // load instructions per wave per stage
    // JA loads of each wave fill A rows, the rest B rows (the A slots are a
    // whole number of four-wave load rounds)
    constexpr int JA = BM * PCS / 256;
    static_assert(SLOTS % 256 == 0 && (64 / PCS) <= 16 && BM % 16 == 0 && (BM * PCS) % 256 == 0,
                  "stage layout");
    static_assert(NST == 2 || NST == 3, "ring depth");
    static_assert((NST - 2) * J < 64, "vmcnt range");
    constexpr int SB = (BM + BN) * STAGE;
    __shared__ __attribute__((aligned(16))) unsigned char L0[SB];
    __shared__ __attribute__((aligned(16))) unsigned char L1[SB];
    __shared__ __attribute__((aligned(16))) unsigned char L2[NST > 2 ? SB : 16];

    // the wave index is uniform: read it into a scalar so that everything
    // derived from it (LDS slots, A or B role of a load) stays scalar
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wm = wave / WNW, wn = wave % WNW;
    unsigned tx, ty, tz;
    xcd_tile(tx, ty, tz);
    const long mb = (long)tx * BM;
    const int nb = ty * BN;
    const int C = p.C, D = p.D, ks = p.ksize;
    const long K = (long)p.taps * C;
    // Scalar bases: the input voxel of the tile's first row (the smallest of
    // its rows: row_voxel grows with m, and the first row of a pooled tile
    // is window element 0) and the tile's first weight row; each lane keeps
    // a 32-bit byte offset from them (a tile spans a few patches of input
    // and at most N*K weights, far below 4 GB).
    const long vb0 = row_voxel<POOL>(mb, p);
    const unsigned char *__restrict__ a_base =
        static_cast<const unsigned char *>(p.in) + vb0 * C * (long)sizeof(T);
    const unsigned char *__restrict__ b_base =
        static_cast<const unsigned char *>(p.w) + (long)nb * K * (long)sizeof(T);

    // load j of this wave fills slots (4j + wave)*64 + lane of a stage
    // buffer: A rows for j < JA, B rows after
    unsigned voff[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int slot = (4 * j + wave) * 64 + lane;
        const int row = slot / PCS, pl = slot % PCS;
        const int ps = pl ^ swz<STAGE>(row);
        if (j < JA) {
            const long m = min(mb + row, p.M - 1);
            voff[j] = (unsigned)((row_voxel<POOL>(m, p) - vb0) * C * (long)sizeof(T)) + ps * 16;
        } else {
            voff[j] = (unsigned)((long)(row - BM) * K * (long)sizeof(T)) + ps * 16;
        }
    }
    const int r16 = lane & 15, g = lane >> 4;
    int a_off[IB][HALVES], b_off[JB][HALVES];
#pragma unroll
    for (int i = 0; i < IB; ++i) {
        const int row = wm * WTM + 16 * i + r16;
#pragma unroll
        for (int h = 0; h < HALVES; ++h)
            a_off[i][h] = row * STAGE + (((g + 4 * h) ^ swz<STAGE>(row)) << 4);
    }
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        const int row = BM + wn * WTN + 16 * j + r16;
#pragma unroll
        for (int h = 0; h < HALVES; ++h)
            b_off[j][h] = row * STAGE + (((g + 4 * h) ^ swz<STAGE>(row)) << 4);
    }

    KWalk kw;
    const int ns = kw.init(C, (int)sizeof(T), STAGE, p.taps);
    int s0 = 0, s1 = ns;
    if (SPLIT) kw.split(ns, tz, p.splits, &s0, &s1);
    kw.seek(s0);
    CopyWalk cw;  // the next stage to copy
    cw.init(kw, C, D, (int)sizeof(T), ks);
    auto issue = [&](unsigned char *buf) {
#pragma unroll
        for (int j = 0; j < J; ++j)
            copy16_to_lds(j < JA ? a_base : b_base, j < JA ? cw.a : cw.b, voff[j],
                          buf + (4 * j + wave) * 1024);
        cw.next();
    };

    floatx4 acc[IB][JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    auto compute = [&](const unsigned char *buf) {
#pragma unroll
        for (int h = 0; h < HALVES; ++h) {
            uint4 a[IB], b[JB];
#pragma unroll
            for (int i = 0; i < IB; ++i) a[i] = *reinterpret_cast<const uint4 *>(buf + a_off[i][h]);
#pragma unroll
            for (int j = 0; j < JB; ++j) b[j] = *reinterpret_cast<const uint4 *>(buf + b_off[j][h]);
            Mfma<T>::chunk(acc, a, b);
        }
    };
    // stage s in buffer `cur`; the copies of stage s+NST-1 go to `nxt` (the
    // buffer of stage s-1)
    auto step = [&](int s, unsigned char *cur, unsigned char *nxt) {
        if (NST == 3 && s + 1 < s1) wait_vmcnt<(NST - 2) * J>();
        else wait_vmcnt<0>();
        __builtin_amdgcn_s_barrier();
        if (s + NST - 1 < s1) issue(nxt);
        compute(cur);
    };
    if (s0 < s1) {
        issue(L0);
        if (NST == 3 && s0 + 1 < s1) issue(L1);
        for (int s = s0; s < s1; s += NST) {
            if (NST == 3) {
                step(s, L0, L2);
                if (s + 1 < s1) step(s + 1, L1, L0);
                if (s + 2 < s1) step(s + 2, L2, L1);
            } else {
                step(s, L0, L1);
                if (s + 1 < s1) step(s + 1, L1, L0);
            }
        }
    }
    conv_epilogue<OT, IB, JB, POOL, SPLIT>(acc, p, mb + wm * WTM, nb + wn * WTN, lane, tz);
}

// Pooled 3x3x3 convolution (16-bit, 128-byte channel units) staging the
// input rows of a (kz, ky) tap group once for its three kx taps.
//
// A tile is kPoolWin consecutive pool windows (8 rows each, window element
// (dz, dy, dx) innermost, as row_voxel orders them) by 64 columns; four waves
// along m own 48 rows each.  The windows of one window line (fixed batch
// item, pz, py) take consecutive "slots", and each line one spare slot
// after its last window, so slot s = window + window / PD.  For a (kz, ky)
// group the tile stages, per (dz, dy) plane, two rows per slot: the input
// voxels at x = 2 px + xo (xo = 0, 1) of the slot (the spare slot reads x =
// OD, OD + 1, inside the input).  Row (window, dz, dy, dx) at tap kx then
// reads staged row plane*64 + 2 (slot - slot0) + dx + kx -- the kx = 0 row
// shifted by kx -- so the group moves 4 x 56 rows of input into LDS instead
// of 3 x 192.  Weights keep their own two-buffer ring, one tap per step.
// Each output still sums (unit, tap, half) in KWalk's order with the same
// MFMA sequence, so the results are bit-identical to the other variants.
//
// Staged rows are XOR-swizzled by (plane, row parity): the 16 rows a
// fragment read touches are pairs of adjacent rows in the four planes, and
// (plane, parity) keeps every 8-lane group on distinct 16-byte bank slots for
// every shift kx.
constexpr int kPoolWin = 24;           // windows per tile
constexpr int kPoolPlane = 64;         // LDS rows per (dz, dy) plane
constexpr int kPoolStaged = 56;        // rows loaded per plane (two per slot)
__device__ inline int swz_plane(int r) { return (((r >> 6) << 1) | (r & 1)) & 7; }

template <typename T, typename OT>
__global__ __launch_bounds__(256, 2) void conv_pool3_kernel(ConvArgs p) {
    static_assert(sizeof(T) == 2, "16-bit operands");
    constexpr int BM = 8 * kPoolWin, BN = 64, WTM = BM / 4, IB = WTM / 16, JB = BN / 16;
    constexpr int ASZ = 4 * kPoolPlane * 128, BSZ = BN * 128;
    constexpr int JA = 4 * kPoolStaged * 8 / 256;  // A loads per wave per group (7)
    constexpr int JW = BN * 8 / 256;               // weight loads per wave per step (2)
    constexpr int PARTS = kPoolStaged * 8 / 64;    // 64-slot loads per plane (7)
    static_assert(WTM % 16 == 0 && 4 * PARTS == 4 * JA && JA == PARTS, "pooled tile layout");
    __shared__ __attribute__((aligned(16))) unsigned char A0[ASZ];
    __shared__ __attribute__((aligned(16))) unsigned char A1[ASZ];
    __shared__ __attribute__((aligned(16))) unsigned char B0[BSZ];
    __shared__ __attribute__((aligned(16))) unsigned char B1[BSZ];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    unsigned tx, ty, tz;
    xcd_tile(tx, ty, tz);
    const long mb = (long)tx * BM;
    const int nb = ty * BN;
    const int C = p.C, D = p.D, PD = p.OD >> 1;
    const unsigned csz = (unsigned)C * (unsigned)sizeof(T);
    const long K = (long)p.taps * C;

    // the tile's windows wi0 .. wl and the slots they read, s0 .. shi
    const unsigned nwin = (unsigned)(p.M >> 3), wi0 = (unsigned)(mb >> 3);
    const unsigned wl = min(wi0 + (unsigned)kPoolWin, nwin) - 1;
    const unsigned s0 = wi0 + row_div(wi0, p), shi = wl + row_div(wl, p) + 1;
    // The slots s0 .. shi lie on at most three window lines (24 windows, at
    // least 12 per line): decode the lines' first voxels once (uniform).
    // input voxel of slot s, plane (dz, dy) and x offset xo
    auto slot_voxel = [&](unsigned s, int plane, int xo) -> long {
        const unsigned line = slot_div(s, p);
        const int px = (int)(s - line * (unsigned)(PD + 1));
        const unsigned q1 = row_div(line, p), q2 = row_div(q1, p);
        const int py = (int)(line - q1 * (unsigned)PD), pz = (int)(q1 - q2 * (unsigned)PD);
        return (((long)q2 * D + 2 * pz + (plane >> 1)) * D + 2 * py + (plane & 1)) * (long)D +
               2 * px + xo;
    };
    const long vb0 = slot_voxel(s0, 0, 0);  // the smallest voxel the tile stages
    const unsigned char *__restrict__ a_base =
        static_cast<const unsigned char *>(p.in) + vb0 * (long)csz;
    const unsigned char *__restrict__ b_base =
        static_cast<const unsigned char *>(p.w) + (long)nb * K * (long)sizeof(T);

    // A load j of this wave fills part (L % PARTS) of plane L / PARTS, L = 4j + wave
    unsigned avoff[JA];
#pragma unroll
    for (int j = 0; j < JA; ++j) {
        const int L = 4 * j + wave, plane = L / PARTS, part = L % PARTS;
        const int q = (part * 64 + lane) >> 3, pl = lane & 7, r = plane * kPoolPlane + q;
        const unsigned s = min(s0 + (unsigned)(q >> 1), shi);
        avoff[j] = (unsigned)((slot_voxel(s, plane, q & 1) - vb0) * (long)csz) +
                   ((pl ^ swz_plane(r)) << 4);
    }
    unsigned bvoff[JW];
#pragma unroll
    for (int j = 0; j < JW; ++j) {
        const int slot = (4 * j + wave) * 64 + lane, row = slot >> 3, pl = slot & 7;
        bvoff[j] = (unsigned)((long)row * K * (long)sizeof(T)) + ((pl ^ swz<128>(row)) << 4);
    }
    const int r16 = lane & 15, g = lane >> 4;
    int aoff[IB][3][2], boff[JB][2];
#pragma unroll
    for (int i = 0; i < IB; ++i) {
        const long m = min(mb + wave * WTM + 16 * i + r16, p.M - 1);
        const unsigned wi = (unsigned)(m >> 3);
        const int w = (int)(m & 7);
        const int r0 = (w >> 1) * kPoolPlane + 2 * (int)(wi + row_div(wi, p) - s0) + (w & 1);
#pragma unroll
        for (int kx = 0; kx < 3; ++kx)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                aoff[i][kx][h] = (r0 + kx) * 128 + (((g + 4 * h) ^ swz_plane(r0 + kx)) << 4);
    }
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        const int row = 16 * j + r16;
#pragma unroll
        for (int h = 0; h < 2; ++h) boff[j][h] = row * 128 + (((g + 4 * h) ^ swz<128>(row)) << 4);
    }

    // copy walks: the A group (unit, kz, ky) and the weight step (unit, tap)
    const int units = (int)(csz / 128u), ng = 9 * units, nsteps = 27 * units;
    const unsigned dy_b = (unsigned)D * csz, dz_b = dy_b * (unsigned)D;
    unsigned a_off = 0, b_off = 0;
    int a_kz = 0, a_ky = 0, b_tap = 0;
    auto issue_a = [&](unsigned char *buf) {
#pragma unroll
        for (int j = 0; j < JA; ++j) {
            const int L = 4 * j + wave;
            copy16_to_lds(a_base, a_off, avoff[j], buf + (L / PARTS) * (kPoolPlane * 128) +
                                                       (L % PARTS) * 1024);
        }
        if (++a_ky < 3) {
            a_off += dy_b;
        } else {
            a_ky = 0;
            a_off -= 2 * dy_b;
            if (++a_kz < 3) {
                a_off += dz_b;
            } else {
                a_kz = 0;
                a_off = a_off - 2 * dz_b + 128u;
            }
        }
    };
    auto issue_b = [&](unsigned char *buf) {
#pragma unroll
        for (int j = 0; j < JW; ++j)
            copy16_to_lds(b_base, b_off, bvoff[j], buf + (4 * j + wave) * 1024);
        if (++b_tap < 27) {
            b_off += csz;
        } else {
            b_tap = 0;
            b_off = b_off - 26 * csz + 128u;
        }
    };

    floatx4 acc[IB][JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    auto compute = [&](const unsigned char *abuf, const unsigned char *bbuf, const int kx) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint4 a[IB], b[JB];
#pragma unroll
            for (int i = 0; i < IB; ++i) a[i] = *reinterpret_cast<const uint4 *>(abuf + aoff[i][kx][h]);
#pragma unroll
            for (int j = 0; j < JB; ++j) b[j] = *reinterpret_cast<const uint4 *>(bbuf + boff[j][h]);
            Mfma<T>::chunk(acc, a, b);
        }
    };
    // Step t = 3 gi + kx of group gi: the weights of step t were copied at
    // step t - 1 (before that step's group copy), the group's rows at the
    // group's first step - 3.  At kx = 1 the next group's copies, issued after
    // this step's weights, may still be in flight.
    auto group = [&](const int gi, unsigned char *acur, unsigned char *anxt, unsigned char *bev,
                     unsigned char *bod) {
#pragma unroll
        for (int kx = 0; kx < 3; ++kx) {
            const int t = 3 * gi + kx;
            // weight buffer of step t: even steps bev, odd bod (gi parity folded in by the caller)
            unsigned char *bcur = (kx & 1) ? bod : bev, *bnxt = (kx & 1) ? bev : bod;
            if (kx == 1 && gi + 1 < ng) wait_vmcnt<JA>();
            else wait_vmcnt<0>();
            __builtin_amdgcn_s_barrier();
            if (t + 1 < nsteps) issue_b(bnxt);
            if (kx == 0 && gi + 1 < ng) issue_a(anxt);
            compute(acur, bcur, kx);
        }
    };
    issue_b(B0);
    issue_a(A0);
    for (int gi = 0; gi < ng; gi += 2) {
        group(gi, A0, A1, B0, B1);                    // steps 3gi.. start even
        if (gi + 1 < ng) group(gi + 1, A1, A0, B1, B0);  // steps 3gi+3.. start odd
    }
    conv_epilogue<OT, IB, JB, true, false>(acc, p, mb + wave * WTM, nb, lane, 0u);
}

// conv_pool3_kernel with resident workgroups (two per CU) that walk a list of
// tiles, for layers of one 128-byte channel unit (9 tap groups, 27 steps): a
// tile's last group issues the next tile's first input rows and weights into
// the buffers it has just freed, and the next tile's copy offsets are worked
// out while this tile's MFMAs run, so a tile no longer starts by waiting for
// its first copies.  XCD x (workgroups w with w % 8 == x) walks one contiguous
// run of the tile sequence, as xcd_tile arranges for the one-shot kernels.
// Tiles, staging, reduction order and epilogue are conv_pool3_kernel's
// (bit-identical).
template <typename T, typename OT>
__global__ __launch_bounds__(256, 2) void conv_pool3p_kernel(ConvArgs p, int ntx, int ntiles) {
    static_assert(sizeof(T) == 2, "16-bit operands");
    constexpr int BM = 8 * kPoolWin, BN = 64, WTM = BM / 4, IB = WTM / 16, JB = BN / 16;
    constexpr int ASZ = 4 * kPoolPlane * 128, BSZ = BN * 128;
    constexpr int JA = 4 * kPoolStaged * 8 / 256, JW = BN * 8 / 256, PARTS = kPoolStaged * 8 / 64;
    constexpr int NG = 9, NSTEPS = 27;  // one channel unit
    __shared__ __attribute__((aligned(16))) unsigned char A0[ASZ];
    __shared__ __attribute__((aligned(16))) unsigned char A1[ASZ];
    __shared__ __attribute__((aligned(16))) unsigned char B0[BSZ];
    __shared__ __attribute__((aligned(16))) unsigned char B1[BSZ];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int C = p.C, D = p.D, PD = p.OD >> 1;
    const unsigned csz = (unsigned)C * (unsigned)sizeof(T);
    const long K = (long)p.taps * C;
    const unsigned dy_b = (unsigned)D * csz, dz_b = dy_b * (unsigned)D;
    const int r16 = lane & 15, g = lane >> 4;

    // this workgroup's tiles: first + k * stride within its XCD's run
    const int xcd = (int)(blockIdx.x % kXcds), per_xcd = (int)(gridDim.x / kXcds);
    const int run = (ntiles + (int)kXcds - 1) / (int)kXcds;
    const int run_end = min((xcd + 1) * run, ntiles);
    const int first = xcd * run + (int)(blockIdx.x / kXcds);
    if (first >= run_end) return;  // whole workgroup

    struct Tl {
        long mb;
        int nb;
        const unsigned char *a_base, *b_base;
        unsigned avoff[JA];
        int aoff[IB][3][2];
    };
    auto setup = [&](int t, Tl &s) {
        const int tx = t % ntx, ty = t / ntx;
        s.mb = (long)tx * BM;
        s.nb = ty * BN;
        const unsigned nwin = (unsigned)(p.M >> 3), wi0 = (unsigned)(s.mb >> 3);
        const unsigned wl = min(wi0 + (unsigned)kPoolWin, nwin) - 1;
        const unsigned s0 = wi0 + row_div(wi0, p), shi = wl + row_div(wl, p) + 1;
        auto slot_voxel = [&](unsigned sl, int plane, int xo) -> long {
            const unsigned line = slot_div(sl, p);
            const int px = (int)(sl - line * (unsigned)(PD + 1));
            const unsigned q1 = row_div(line, p), q2 = row_div(q1, p);
            const int py = (int)(line - q1 * (unsigned)PD), pz = (int)(q1 - q2 * (unsigned)PD);
            return (((long)q2 * D + 2 * pz + (plane >> 1)) * D + 2 * py + (plane & 1)) * (long)D +
                   2 * px + xo;
        };
        const long vb0 = slot_voxel(s0, 0, 0);
        s.a_base = static_cast<const unsigned char *>(p.in) + vb0 * (long)csz;
        s.b_base = static_cast<const unsigned char *>(p.w) + (long)s.nb * K * (long)sizeof(T);
#pragma unroll
        for (int j = 0; j < JA; ++j) {
            const int L = 4 * j + wave, plane = L / PARTS, part = L % PARTS;
            const int q = (part * 64 + lane) >> 3, pl = lane & 7, r = plane * kPoolPlane + q;
            const unsigned sl = min(s0 + (unsigned)(q >> 1), shi);
            s.avoff[j] = (unsigned)((slot_voxel(sl, plane, q & 1) - vb0) * (long)csz) +
                         ((pl ^ swz_plane(r)) << 4);
        }
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            const long m = min(s.mb + wave * WTM + 16 * i + r16, p.M - 1);
            const unsigned wi = (unsigned)(m >> 3);
            const int w = (int)(m & 7);
            const int r0 = (w >> 1) * kPoolPlane + 2 * (int)(wi + row_div(wi, p) - s0) + (w & 1);
#pragma unroll
            for (int kx = 0; kx < 3; ++kx)
#pragma unroll
                for (int h = 0; h < 2; ++h)
                    s.aoff[i][kx][h] = (r0 + kx) * 128 + (((g + 4 * h) ^ swz_plane(r0 + kx)) << 4);
        }
    };
    unsigned bvoff[JW];
#pragma unroll
    for (int j = 0; j < JW; ++j) {
        const int slot = (4 * j + wave) * 64 + lane, row = slot >> 3, pl = slot & 7;
        bvoff[j] = (unsigned)((long)row * K * (long)sizeof(T)) + ((pl ^ swz<128>(row)) << 4);
    }
    int boff[JB][2];
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        const int row = 16 * j + r16;
#pragma unroll
        for (int h = 0; h < 2; ++h) boff[j][h] = row * 128 + (((g + 4 * h) ^ swz<128>(row)) << 4);
    }
    // group g's input rows (kz = g / 3, ky = g % 3) and step t's weights (tap t)
    auto issue_a = [&](unsigned char *buf, const Tl &s, int gr) {
        const unsigned off = (unsigned)(gr / 3) * dz_b + (unsigned)(gr % 3) * dy_b;
#pragma unroll
        for (int j = 0; j < JA; ++j) {
            const int L = 4 * j + wave;
            copy16_to_lds(s.a_base, off, s.avoff[j],
                          buf + (L / PARTS) * (kPoolPlane * 128) + (L % PARTS) * 1024);
        }
    };
    auto issue_b = [&](unsigned char *buf, const Tl &s, int st) {
#pragma unroll
        for (int j = 0; j < JW; ++j)
            copy16_to_lds(s.b_base, (unsigned)st * csz, bvoff[j], buf + (4 * j + wave) * 1024);
    };

    Tl cur, nxt;
    floatx4 acc[IB][JB];
    auto compute = [&](const unsigned char *abuf, const unsigned char *bbuf, const int kx) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint4 a[IB], b[JB];
#pragma unroll
            for (int i = 0; i < IB; ++i)
                a[i] = *reinterpret_cast<const uint4 *>(abuf + cur.aoff[i][kx][h]);
#pragma unroll
            for (int j = 0; j < JB; ++j) b[j] = *reinterpret_cast<const uint4 *>(bbuf + boff[j][h]);
            Mfma<T>::chunk(acc, a, b);
        }
    };
    // one tile; its groups alternate between (aev, bev) and (aod, bod) from
    // group 0 (steps alike), and the next tile's group 0 goes to the pair its
    // group 8 does not use
    auto tile = [&](unsigned char *aev, unsigned char *aod, unsigned char *bev, unsigned char *bod,
                    const bool has_next, const int tn) {
#pragma unroll
        for (int gi = 0; gi < NG; ++gi) {
            unsigned char *acur = (gi & 1) ? aod : aev, *anxt = (gi & 1) ? aev : aod;
#pragma unroll
            for (int kx = 0; kx < 3; ++kx) {
                const int t = 3 * gi + kx;
                unsigned char *bcur = (t & 1) ? bod : bev, *bnxt = (t & 1) ? bev : bod;
                // at kx = 1 the group copy issued at kx = 0 (after this
                // step's weights) may still be in flight
                if (kx == 1 && (gi + 1 < NG || has_next)) wait_vmcnt<JA>();
                else wait_vmcnt<0>();
                __builtin_amdgcn_s_barrier();
                if (t + 1 < NSTEPS) issue_b(bnxt, cur, t + 1);
                else if (has_next) issue_b(bnxt, nxt, 0);
                if (kx == 0) {
                    if (gi + 1 < NG) issue_a(anxt, cur, gi + 1);
                    else if (has_next) issue_a(anxt, nxt, 0);
                }
                compute(acur, bcur, kx);
                // the next tile's offsets, behind the first step's MFMAs
                if (t == 0 && has_next) setup(tn, nxt);
            }
        }
    };
    setup(first, cur);
    issue_b(B0, cur, 0);
    issue_a(A0, cur, 0);
    for (int t = first; t < run_end;) {
        // two tiles per iteration: the buffer pairs swap roles each tile
#pragma unroll
        for (int par = 0; par < 2; ++par) {
            const int tn = t + per_xcd;
            const bool has_next = tn < run_end;
#pragma unroll
            for (int i = 0; i < IB; ++i)
#pragma unroll
                for (int j = 0; j < JB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
            if (par == 0) tile(A0, A1, B0, B1, has_next, tn);
            else tile(A1, A0, B1, B0, has_next, tn);
            conv_epilogue<OT, IB, JB, true, false>(acc, p, cur.mb + wave * WTM, cur.nb, lane, 0u);
            t = tn;
            if (!has_next) break;
            cur = nxt;
        }
    }
}

// The pooled tap-group kernel with one workgroup of eight waves per CU and a
// whole (kz, ky) group -- its input rows and the weights of its three kx taps
// -- per stage buffer: 48 windows x 64 columns per tile, eight waves of 48
// rows, so one barrier publishes three taps of work (against one per tap)
// and the copies of the next group have a whole group of MFMAs to land.  The
// two-buffer ring takes the CU's whole 160 KB of LDS (2 x 56 KB of input rows,
// 4 planes x 112 rows, and 2 x 24 KB of weights).  Same row order, staging,
// swizzle and reduction order as conv_pool3_kernel (bit-identical).
constexpr int kPoolWinG = 48;     // windows per tile
constexpr int kPoolPlaneG = 112;  // LDS rows per (dz, dy) plane, all loaded

template <typename T, typename OT>
__global__ __launch_bounds__(512, 1) void conv_pool3g_kernel(ConvArgs p) {
    static_assert(sizeof(T) == 2, "16-bit operands");
    constexpr int WAVES = 8, BM = 8 * kPoolWinG, BN = 64, WTM = BM / WAVES, IB = WTM / 16,
                  JB = BN / 16;
    constexpr int ASZ = 4 * kPoolPlaneG * 128, BSZ = 3 * BN * 128;
    constexpr int PARTS = kPoolPlaneG * 8 / 64;   // 64-slot loads per plane (14)
    constexpr int JA = 4 * PARTS / WAVES;         // input-row loads per wave per group (7)
    constexpr int BPARTS = BN * 8 / 64;           // 64-slot loads per tap of weights (8)
    constexpr int JW = 3 * BPARTS / WAVES;        // weight loads per wave per group (3)
    static_assert(WTM % 16 == 0 && (4 * PARTS) % WAVES == 0 && (3 * BPARTS) % WAVES == 0,
                  "pooled group tile layout");
    __shared__ __attribute__((aligned(16))) unsigned char A0[ASZ];
    __shared__ __attribute__((aligned(16))) unsigned char A1[ASZ];
    __shared__ __attribute__((aligned(16))) unsigned char B0[BSZ];
    __shared__ __attribute__((aligned(16))) unsigned char B1[BSZ];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    unsigned tx, ty, tz;
    xcd_tile(tx, ty, tz);
    const long mb = (long)tx * BM;
    const int nb = ty * BN;
    const int C = p.C, D = p.D, PD = p.OD >> 1;
    const unsigned csz = (unsigned)C * (unsigned)sizeof(T);
    const long K = (long)p.taps * C;

    const unsigned nwin = (unsigned)(p.M >> 3), wi0 = (unsigned)(mb >> 3);
    const unsigned wl = min(wi0 + (unsigned)kPoolWinG, nwin) - 1;
    const unsigned s0 = wi0 + row_div(wi0, p), shi = wl + row_div(wl, p) + 1;
    auto slot_voxel = [&](unsigned s, int plane, int xo) -> long {
        const unsigned line = slot_div(s, p);
        const int px = (int)(s - line * (unsigned)(PD + 1));
        const unsigned q1 = row_div(line, p), q2 = row_div(q1, p);
        const int py = (int)(line - q1 * (unsigned)PD), pz = (int)(q1 - q2 * (unsigned)PD);
        return (((long)q2 * D + 2 * pz + (plane >> 1)) * D + 2 * py + (plane & 1)) * (long)D +
               2 * px + xo;
    };
    const long vb0 = slot_voxel(s0, 0, 0);
    const unsigned char *__restrict__ a_base =
        static_cast<const unsigned char *>(p.in) + vb0 * (long)csz;
    const unsigned char *__restrict__ b_base =
        static_cast<const unsigned char *>(p.w) + (long)nb * K * (long)sizeof(T);
    auto swz_g = [](int plane, int r) { return ((plane << 1) | (r & 1)) & 7; };

    unsigned avoff[JA];
#pragma unroll
    for (int j = 0; j < JA; ++j) {
        const int L = WAVES * j + wave, plane = L / PARTS, part = L % PARTS;
        const int q = (part * 64 + lane) >> 3, pl = lane & 7;
        const unsigned s = min(s0 + (unsigned)(q >> 1), shi);
        avoff[j] = (unsigned)((slot_voxel(s, plane, q & 1) - vb0) * (long)csz) +
                   ((pl ^ swz_g(plane, q)) << 4);
    }
    unsigned bvoff[JW];
#pragma unroll
    for (int j = 0; j < JW; ++j) {
        const int L = WAVES * j + wave, kx = L / BPARTS, part = L % BPARTS;
        const int slot = part * 64 + lane, row = slot >> 3, pl = slot & 7;
        bvoff[j] = (unsigned)((long)row * K * (long)sizeof(T)) + (unsigned)kx * csz +
                   ((pl ^ swz<128>(row)) << 4);
    }
    const int r16 = lane & 15, g = lane >> 4;
    int aoff[IB][3][2], boff[JB][2];
#pragma unroll
    for (int i = 0; i < IB; ++i) {
        const long m = min(mb + wave * WTM + 16 * i + r16, p.M - 1);
        const unsigned wi = (unsigned)(m >> 3);
        const int w = (int)(m & 7), plane = w >> 1;
        const int q0 = 2 * (int)(wi + row_div(wi, p) - s0) + (w & 1);
#pragma unroll
        for (int kx = 0; kx < 3; ++kx)
#pragma unroll
            for (int h = 0; h < 2; ++h)
                aoff[i][kx][h] = (plane * kPoolPlaneG + q0 + kx) * 128 +
                                 (((g + 4 * h) ^ swz_g(plane, q0 + kx)) << 4);
    }
#pragma unroll
    for (int j = 0; j < JB; ++j) {
        const int row = 16 * j + r16;
#pragma unroll
        for (int h = 0; h < 2; ++h) boff[j][h] = row * 128 + (((g + 4 * h) ^ swz<128>(row)) << 4);
    }

    // group walk (unit, kz, ky): input offset kz*dz + ky*dy + unit*128, weight
    // offset (3 (3 kz + ky)) taps + unit*128
    const int units = (int)(csz / 128u), ng = 9 * units;
    const unsigned dy_b = (unsigned)D * csz, dz_b = dy_b * (unsigned)D;
    unsigned a_off = 0, b_off = 0;
    int a_kz = 0, a_ky = 0;
    auto issue = [&](unsigned char *abuf, unsigned char *bbuf) {
#pragma unroll
        for (int j = 0; j < JA; ++j) {
            const int L = WAVES * j + wave;
            copy16_to_lds(a_base, a_off, avoff[j],
                          abuf + (L / PARTS) * (kPoolPlaneG * 128) + (L % PARTS) * 1024);
        }
#pragma unroll
        for (int j = 0; j < JW; ++j) {
            const int L = WAVES * j + wave;
            copy16_to_lds(b_base, b_off, bvoff[j],
                          bbuf + (L / BPARTS) * (BN * 128) + (L % BPARTS) * 1024);
        }
        if (++a_ky < 3) {
            a_off += dy_b;
            b_off += 3 * csz;
        } else {
            a_ky = 0;
            a_off -= 2 * dy_b;
            if (++a_kz < 3) {
                a_off += dz_b;
                b_off += 3 * csz;
            } else {
                a_kz = 0;
                a_off = a_off - 2 * dz_b + 128u;
                b_off = b_off - 24 * csz + 128u;
            }
        }
    };

    floatx4 acc[IB][JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    auto compute = [&](const unsigned char *abuf, const unsigned char *bbuf) {
#pragma unroll
        for (int kx = 0; kx < 3; ++kx)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                uint4 a[IB], b[JB];
#pragma unroll
                for (int i = 0; i < IB; ++i)
                    a[i] = *reinterpret_cast<const uint4 *>(abuf + aoff[i][kx][h]);
#pragma unroll
                for (int j = 0; j < JB; ++j)
                    b[j] = *reinterpret_cast<const uint4 *>(bbuf + kx * (BN * 128) + boff[j][h]);
                Mfma<T>::chunk(acc, a, b);
            }
    };
    // group gi in buffers gi & 1; the copies of group gi + 1 go to the other
    // pair, last read by group gi - 1 (before this group's barrier)
    auto group = [&](const int gi, unsigned char *acur, unsigned char *bcur, unsigned char *anxt,
                     unsigned char *bnxt) {
        wait_vmcnt<0>();
        __builtin_amdgcn_s_barrier();
        if (gi + 1 < ng) issue(anxt, bnxt);
        compute(acur, bcur);
    };
    issue(A0, B0);
    for (int gi = 0; gi < ng; gi += 2) {
        group(gi, A0, B0, A1, B1);
        if (gi + 1 < ng) group(gi + 1, A1, B1, A0, B0);
    }
    conv_epilogue<OT, IB, JB, true, false>(acc, p, mb + wave * WTM, nb, lane, 0u);
}

// Sum of split-K partials + bias (+ReLU), converted to the output type.
template <typename OT>
__global__ void split_reduce_kernel(const float *__restrict__ part, int splits, long MN, int N,
                                    const float *__restrict__ bias, int relu, OT *__restrict__ out,
                                    const float *__restrict__ colscale, float oscale) {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < MN;
         i += (long)gridDim.x * blockDim.x) {
        float s = part[i];
        for (int k = 1; k < splits; ++k) s += part[(long)k * MN + i];
        s = colscale ? s * colscale[i % N] + bias[i % N] : s + bias[i % N];
        if (relu) s = s > 0.0f ? s : 0.0f;
        if (sizeof(OT) == 1) s *= oscale;
        store<OT>(out + i, s);
    }
}

// ---- first convolution (one input channel) -----------------------------------------

struct Conv1Args {
    const float *base;
    const int *kp;
    const tdm::GridRef *grids;  // multi-grid mode: table of grids, indexed by gid
    const int *gid;
    int dx, dy, dz;
    int P;  // patch extent; output extent P - 2
    const void *w;
    const float *bias;
    void *out;
    long M;
    int N;
    int relu;
    float oscale;  // fp8 outputs: multiplier before conversion
    // n / (P - 2) = umulhi(n, od_mul) >> od_sh for n < 2^31 (od_mul = 0:
    // P - 2 == 1); used when M < 2^31, else the row decode divides in 64 bits
    unsigned od_mul;
    int od_sh;
};

__device__ inline unsigned div_od(unsigned n, const Conv1Args &a) {
    return a.od_mul ? __umulhi(n, a.od_mul) >> a.od_sh : n;
}

// MODE 0: patches [batch][P^3]; 1: keypoints in one grid; 2: keypoints in
// a table of grids (gid[b] picks patch b's grid).
// Four workgroups per CU (as many as the LDS staging of the output allows;
// the register budget that follows keeps the gathers' latency covered:
// 1.42 -> 1.19 ms per 1024-patch fp16 pass); three with fp32 operands, whose
// registers do not fit four without spilling.
template <typename T, typename OT, int MODE>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 3 : 4) void conv1_kernel(Conv1Args a) {
    constexpr bool GRID = MODE != 0;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long m0 = ((long)blockIdx.x * 4 + wave) * 64;
    const int n0 = blockIdx.y * 64;
    // No early exit: the epilogue has a workgroup barrier.  Rows past M are
    // clamped for loads and never stored.
    const int r16 = lane & 15, g = lane >> 4;
    const int OD = a.P - 2;
    constexpr bool F32 = sizeof(T) == 4;
    constexpr int KP = F32 ? 28 : 32;

    // Per row block: the source voxel of tap (0,0,0) (pointer), its grid
    // coordinates, the grid's extents and pitches, and whether the whole
    // 3x3x3 neighbourhood is inside the grid (then taps need no bounds test).
    const float *vb[4];
    int gx[4], gy[4], gz[4];
    bool inb[4];
    int ex[4], ey[4], ez[4], sy[4];
    long sz[4];
    const bool small = a.M < (1L << 31);  // rows fit 31 bits: multiply-high division
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        long m = min(m0 + 16 * i + r16, a.M - 1);
        int px, py, pz;
        long b;
        if (small) {
            const unsigned u = (unsigned)m;
            const unsigned q1 = div_od(u, a), q2 = div_od(q1, a), q3 = div_od(q2, a);
            px = (int)(u - q1 * (unsigned)OD);
            py = (int)(q1 - q2 * (unsigned)OD);
            pz = (int)(q2 - q3 * (unsigned)OD);
            b = q3;
        } else {
            px = (int)(m % OD);
            long q = m / OD;
            py = (int)(q % OD);
            q /= OD;
            pz = (int)(q % OD);
            b = q / OD;
        }
        if (GRID) {
            const float *base = a.base;
            ex[i] = a.dx;
            ey[i] = a.dy;
            ez[i] = a.dz;
            if (MODE == 2) {
                const tdm::GridRef &gr = a.grids[a.gid[b]];
                base = gr.tdf;
                ex[i] = gr.dims[0];
                ey[i] = gr.dims[1];
                ez[i] = gr.dims[2];
            }
            sy[i] = ex[i];
            sz[i] = (long)ex[i] * ey[i];
            const int h = a.P / 2;
            gx[i] = a.kp[b * 3 + 0] - h + px;
            gy[i] = a.kp[b * 3 + 1] - h + py;
            gz[i] = a.kp[b * 3 + 2] - h + pz;
            inb[i] = gx[i] >= 0 && gy[i] >= 0 && gz[i] >= 0 && gx[i] + 2 < ex[i] &&
                     gy[i] + 2 < ey[i] && gz[i] + 2 < ez[i];
            vb[i] = base + ((long)gz[i] * sz[i] + (long)gy[i] * sy[i] + gx[i]);
        } else {
            gx[i] = px;
            gy[i] = py;
            gz[i] = pz;
            inb[i] = true;
            ex[i] = ey[i] = ez[i] = a.P;
            sy[i] = a.P;
            sz[i] = (long)a.P * a.P;
            vb[i] = a.base + (b * (long)a.P * a.P * a.P + (long)pz * sz[i] + (long)py * sy[i] + px);
        }
    }
    // Input value of the tap (kz, ky, kx) for row block i.  Outside the grid
    // reads 0; the address is formed only for taps inside it.
    auto tap_value = [&](int i, int kz, int ky, int kx) -> float {
        if (!GRID || inb[i]) return vb[i][kz * sz[i] + ky * sy[i] + kx];
        const int x = gx[i] + kx, y = gy[i] + ky, z = gz[i] + kz;
        if (x < 0 || y < 0 || z < 0 || x >= ex[i] || y >= ey[i] || z >= ez[i]) return 0.0f;
        return vb[i][kz * sz[i] + ky * sy[i] + kx];
    };
    // A row block whose rows all have their 3x3x3 neighbourhood inside the
    // grid (a wave-wide vote, so the branch is uniform) reads its taps without
    // bounds tests, at 32-bit tap offsets.  Patch and single-grid modes share
    // one pair of pitches, so a lane's offsets are computed once (the host
    // checks that they fit); with a table of grids each row's own grid gives
    // them, and a block takes this path only if its grids' slices are below
    // 2^30 voxels.
    constexpr bool UNI = MODE != 2;
    const int usy = MODE == 0 ? a.P : a.dx, usz = MODE == 0 ? a.P * a.P : a.dx * a.dy;
    auto tap_off = [&](int t) { return (t / 9) * usz + ((t / 3) % 3) * usy + t % 3; };
    auto block_in = [&](int i) {
        return !GRID || __all(inb[i] && (UNI || sz[i] < (1L << 30)));
    };
    auto row_tap_off = [&](int i, int kz, int ky, int kx) {
        return kz * (int)sz[i] + ky * sy[i] + kx;
    };
    floatx4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    const T *__restrict__ w = static_cast<const T *>(a.w);
    if constexpr (F32) {
        // 7 k-steps of 4 taps: lane group g supplies tap 4s + g.  The tap
        // number depends on the lane, so its (kz, ky, kx) is decoded once
        // here, not once per row block.
        bool fast[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) fast[i] = block_in(i);
#pragma unroll
        for (int s = 0; s < 7; ++s) {
            const int t = 4 * s + g;
            const int kz = t / 9, ky = (t / 3) % 3, kx = t % 3;
            const int to = t < 27 ? tap_off(t) : 0;
            float av[4], bv[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                if (fast[i]) {
                    const float v = vb[i][UNI ? to : t < 27 ? row_tap_off(i, kz, ky, kx) : 0];
                    av[i] = t < 27 ? v : 0.0f;
                } else {
                    av[i] = t < 27 ? tap_value(i, kz, ky, kx) : 0.0f;
                }
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float *wf = reinterpret_cast<const float *>(w);
                bv[j] = wf[(long)(n0 + 16 * j + r16) * KP + t];
            }
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    acc[j][i] = __builtin_amdgcn_mfma_f32_16x16x4f32(bv[j], av[i], acc[j][i], 0, 0, 0);
        }
    } else {
        // One k-step of 32 taps: lane group g supplies taps 8g .. 8g+7.  The
        // tap numbers depend on the lane, so their (kz, ky, kx) are decoded
        // once here, not once per row block.
        int tk[8][3];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int t = 8 * g + e;
            tk[e][0] = t / 9;
            tk[e][1] = (t / 3) % 3;
            tk[e][2] = t % 3;
        }
        int to[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) to[e] = 8 * g + e < 27 ? tap_off(8 * g + e) : 0;
        uint4 af[4], bf[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            unsigned short hv[8];
            if (block_in(i)) {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int te = 8 * g + e;
                    const float v = vb[i][UNI ? to[e]
                                          : te < 27 ? row_tap_off(i, tk[e][0], tk[e][1], tk[e][2])
                                                    : 0];
                    hv[e] = bits16<T>(8 * g + e < 27 ? v : 0.0f);
                }
            } else {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int t = 8 * g + e;
                    hv[e] = bits16<T>(t < 27 ? tap_value(i, tk[e][0], tk[e][1], tk[e][2]) : 0.0f);
                }
            }
            __builtin_memcpy(&af[i], hv, 16);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) bf[j] = ld16(w + (long)(n0 + 16 * j + r16) * KP + 8 * g);
        Mfma<T>::chunk(acc, bf, af);  // acc[j][i]: the weights are the first operand
    }

    // Epilogue through LDS.  With the weights as the MFMAs' first operand,
    // acc[j][i] gives the lane four consecutive channels (16j + 4g ..) of row
    // 16i + r16, which go to LDS in one 4-value store; staging the wave's
    // 64x64 tile then lets every lane write whole 16-byte pieces of contiguous
    // output rows (conv1 is bound by its output writes: 28^3 x 64 values per
    // patch).  Float outputs go in two 32-column halves, so every output type
    // stages 36 KB per workgroup (four workgroups per CU).
    constexpr int HALVES = sizeof(OT) == 4 ? 2 : 1;
    constexpr int COLS = 64 / HALVES;
    constexpr int PITCH = COLS + 16 / (int)sizeof(OT);  // padded row, in OT elements
    __shared__ __attribute__((aligned(16))) OT tile[4][64 * PITCH];
    OT *__restrict__ t = tile[wave];
    OT *__restrict__ out = static_cast<OT *>(a.out);
    constexpr int PIECES = COLS * (int)sizeof(OT) / 16;  // 16-byte pieces per half row
    constexpr int PER = 16 / (int)sizeof(OT);            // elements per piece
#pragma unroll
    for (int h = 0; h < HALVES; ++h) {
#pragma unroll
        for (int j = h * (4 / HALVES); j < (h + 1) * (4 / HALVES); ++j) {
            const int c0 = 16 * j + 4 * g;
            const float4 bv = *reinterpret_cast<const float4 *>(a.bias + n0 + c0);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v[4] = {acc[j][i][0] + bv.x, acc[j][i][1] + bv.y, acc[j][i][2] + bv.z,
                              acc[j][i][3] + bv.w};
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    if (a.relu) v[r] = v[r] > 0.0f ? v[r] : 0.0f;
                    if (sizeof(OT) == 1) v[r] *= a.oscale;
                }
                store4<OT>(t + (16 * i + r16) * PITCH + c0 - h * COLS, v);
            }
        }
        __syncthreads();
#pragma unroll
        for (int it = 0; it < 64 * PIECES / 64; ++it) {
            const int idx = it * 64 + lane;
            const int rl = idx / PIECES, pc = idx % PIECES;
            const long row = m0 + rl;
            if (row < a.M)
                *reinterpret_cast<uint4 *>(out + row * a.N + n0 + h * COLS + pc * PER) =
                    *reinterpret_cast<const uint4 *>(t + rl * PITCH + pc * PER);
        }
        if (h + 1 < HALVES) __syncthreads();  // the half is copied out before the next overwrites it
    }
}

// ---- layout helpers -------------------------------------------------------------------

template <typename T>
__device__ inline float load_as_float(const T *p);
template <>
__device__ inline float load_as_float<float>(const float *p) {
    return *p;
}
template <>
__device__ inline float load_as_float<bf16_t>(const bf16_t *p) {
    return from_bf16(*p);
}
template <>
__device__ inline float load_as_float<f16_t>(const f16_t *p) {
    return (float)*p;
}

template <>
__device__ inline float load_as_float<fp8_t>(const fp8_t *p) {
    return from_fp8(*p);
}

template <typename T>
__global__ void to_ncdhw_kernel(const T *__restrict__ in, long batch, int C, int E,
                                float *__restrict__ out, float scale) {
    const long vol = (long)E * E * E;
    const long total = batch * C * vol;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x) {
        const long v = i % vol;
        const long bc = i / vol;
        const int c = (int)(bc % C);
        const long b = bc / C;
        const float x = load_as_float<T>(in + (b * vol + v) * C + c);
        out[i] = sizeof(T) == 1 ? x * scale : x;
    }
}

template <typename T>
__global__ void maxpool_kernel(const T *__restrict__ in, long batch, int C, int E, int OE,
                               T *__restrict__ out) {
    const long total = batch * OE * OE * OE * (long)C;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x) {
        const int c = (int)(i % C);
        long q = i / C;
        const int x = (int)(q % OE);
        q /= OE;
        const int y = (int)(q % OE);
        q /= OE;
        const int z = (int)(q % OE);
        const long b = q / OE;
        float m = -INFINITY;
        for (int kz = 0; kz < 2; ++kz)
            for (int ky = 0; ky < 2; ++ky)
                for (int kx = 0; kx < 2; ++kx) {
                    const int iz = 2 * z + kz, iy = 2 * y + ky, ix = 2 * x + kx;
                    if (iz >= E || iy >= E || ix >= E) continue;
                    m = fmaxf(m, load_as_float<T>(in + (((b * E + iz) * E + iy) * E + ix) * C + c));
                }
        store<T>(out + i, m);
    }
}

// Compute units of the current device (cached per device).
int device_cus() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    if (cached[dev] <= 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0)
            n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

template <typename T, typename OT>
int conv_dispatch(const tdm::ConvDesc &d, int batch, const void *in, const void *w,
                  const float *bias, void *out, int splits, float *partial, int variant,
                  hipStream_t s, const float *colscale, float oscale) {
    ConvArgs p;
    p.colscale = colscale;
    p.oscale = oscale;
    p.in = in;
    p.w = w;
    p.bias = bias;
    p.out = out;
    p.partial = partial;
    p.N = d.N;
    p.C = d.C;
    p.D = d.D;
    p.OD = d.OD;
    p.ksize = d.ksize;
    p.taps = d.ksize * d.ksize * d.ksize;
    p.splits = splits;
    p.relu = d.relu ? 1 : 0;
    p.M = (long)batch * d.OD * d.OD * d.OD;
    p.small_rows = p.M < (1L << 31) ? 1 : 0;
    fast_div(d.pool ? d.OD / 2 : d.OD, &p.rdiv_mul, &p.rdiv_sh);
    fast_div((d.pool ? d.OD / 2 : d.OD) + 1, &p.sdiv_mul, &p.sdiv_sh);
    const bool lds = tdm::conv_uses_lds(dtype_of<T>(), d);
    if (lds) {
        int v = variant;
        constexpr tdm::DType DT = dtype_of<T>();
        if (v < 0 || !tdm::conv_variant_valid(DT, v, d)) v = tdm::conv_default_variant(DT, d);
#define TDM_CONV_LDS(WMW, WTM, WTN, ST)                                                       \
    do {                                                                                      \
        using TL = Tile<WMW, WTM, WTN, ST>;                                                   \
        dim3 grid((unsigned)((p.M + TL::BM - 1) / TL::BM), d.N / TL::BN, splits);             \
        if (splits > 1)                                                                       \
            conv_lds_kernel<T, OT, WMW, WTM, WTN, ST, false, true><<<grid, 256, 0, s>>>(p);   \
        else if (d.pool)                                                                      \
            conv_lds_kernel<T, OT, WMW, WTM, WTN, ST, true, false><<<grid, 256, 0, s>>>(p);   \
        else                                                                                  \
            conv_lds_kernel<T, OT, WMW, WTM, WTN, ST, false, false><<<grid, 256, 0, s>>>(p);  \
    } while (0)
#define TDM_CONV_DMA(WMW, WTM, WTN, ST, NS)                                                   \
    do {                                                                                      \
        using TL = Tile<WMW, WTM, WTN, ST>;                                                   \
        dim3 grid((unsigned)((p.M + TL::BM - 1) / TL::BM), d.N / TL::BN, splits);             \
        if (splits > 1)                                                                       \
            conv_dma_kernel<T, OT, WMW, WTM, WTN, ST, NS, false, true><<<grid, 256, 0, s>>>(p); \
        else if (d.pool)                                                                      \
            conv_dma_kernel<T, OT, WMW, WTM, WTN, ST, NS, true, false><<<grid, 256, 0, s>>>(p); \
        else                                                                                  \
            conv_dma_kernel<T, OT, WMW, WTM, WTN, ST, NS, false, false><<<grid, 256, 0, s>>>(p); \
    } while (0)
        switch (v) {
        case 9: TDM_CONV_DMA(4, 64, 64, 64, 3); break;
        case 10: TDM_CONV_DMA(2, 64, 64, 64, 3); break;
        case 11: TDM_CONV_DMA(2, 128, 32, 64, 3); break;
        case 12: TDM_CONV_DMA(2, 64, 64, 128, 2); break;
        case 13: TDM_CONV_DMA(4, 64, 64, 128, 2); break;
        case 14:
            // the pooled tap-group kernel runs unsplit pooled layers with
            // 31-bit rows; the same layer split (or wider) runs variant 13
            if constexpr (sizeof(T) == 2) {
                if (splits == 1 && d.pool && p.small_rows) {
                    dim3 grid((unsigned)((p.M + 8 * kPoolWin - 1) / (8 * kPoolWin)), d.N / 64, 1);
                    conv_pool3_kernel<T, OT><<<grid, 256, 0, s>>>(p);
                    break;
                }
            }
            TDM_CONV_DMA(4, 64, 64, 128, 2);
            break;
        case 16:
            // resident pooled tap groups: the same conditions as 14, two
            // workgroups per CU (a multiple of the 8 XCDs)
            if constexpr (sizeof(T) == 2) {
                if (splits == 1 && d.pool && p.small_rows) {
                    const int ntx = (int)((p.M + 8 * kPoolWin - 1) / (8 * kPoolWin));
                    const int ntiles = ntx * (d.N / 64);
                    const unsigned wgs = (unsigned)(2 * device_cus()) / kXcds * kXcds;
                    conv_pool3p_kernel<T, OT><<<wgs, 256, 0, s>>>(p, ntx, ntiles);
                    break;
                }
            }
            TDM_CONV_DMA(4, 64, 64, 128, 2);
            break;
        case 15:
            // eight-wave pooled tap groups: the same conditions as 14
            if constexpr (sizeof(T) == 2) {
                if (splits == 1 && d.pool && p.small_rows) {
                    dim3 grid((unsigned)((p.M + 8 * kPoolWinG - 1) / (8 * kPoolWinG)), d.N / 64, 1);
                    conv_pool3g_kernel<T, OT><<<grid, 512, 0, s>>>(p);
                    break;
                }
            }
            TDM_CONV_DMA(4, 64, 64, 128, 2);
            break;
        case 0: TDM_CONV_LDS(2, 64, 32, 128); break;
        case 1: TDM_CONV_LDS(2, 64, 64, 128); break;
        case 2: TDM_CONV_LDS(4, 64, 64, 128); break;
        case 3: TDM_CONV_LDS(2, 128, 32, 128); break;
        case 4: TDM_CONV_LDS(2, 128, 64, 128); break;
        case 5: TDM_CONV_LDS(1, 64, 64, 128); break;
        case 6: TDM_CONV_LDS(4, 64, 64, 64); break;
        case 7: TDM_CONV_LDS(2, 64, 64, 64); break;
        default: TDM_CONV_LDS(2, 128, 32, 64); break;
        }
#undef TDM_CONV_LDS
#undef TDM_CONV_DMA
    } else if constexpr (sizeof(T) == 1) {
        return tdm::fail(TDM_ERR_UNSUPPORTED, "fp8 convolution with %d input channels", d.C);
    } else {
        const int waves_n = d.N % 256 == 0 ? 4 : d.N % 128 == 0 ? 2 : 1;
        const int waves_m = 4 / waves_n;
        dim3 grid((unsigned)((p.M + 64L * waves_m - 1) / (64L * waves_m)), d.N / (64 * waves_n),
                  splits);
#define TDM_CONV_LAUNCH(WN)                                                                   \
    do {                                                                                      \
        if (splits > 1)                                                                       \
            conv_kernel<T, OT, WN, false, true><<<grid, 256, 0, s>>>(p);                      \
        else if (d.pool)                                                                      \
            conv_kernel<T, OT, WN, true, false><<<grid, 256, 0, s>>>(p);                      \
        else                                                                                  \
            conv_kernel<T, OT, WN, false, false><<<grid, 256, 0, s>>>(p);                     \
    } while (0)
        if (waves_n == 4) TDM_CONV_LAUNCH(4);
        else if (waves_n == 2) TDM_CONV_LAUNCH(2);
        else TDM_CONV_LAUNCH(1);
#undef TDM_CONV_LAUNCH
    }
    TDM_LAUNCHED();
    if (splits > 1) {
        const long MN = p.M * d.N;
        unsigned blocks = (unsigned)std::min<long>((MN + 255) / 256, 16384);
        split_reduce_kernel<OT><<<blocks, 256, 0, s>>>(partial, splits, MN, d.N, bias, p.relu,
                                                       static_cast<OT *>(out), p.colscale,
                                                       p.oscale);
        TDM_LAUNCHED();
    }
    return TDM_OK;
}

}  // namespace

namespace tdm {

// The LDS-staged kernel handles channel counts that are whole stages
// (32 fp32 / 64 bf16 channels); TDM_CONV_IMPL=direct selects the
// register-direct kernel for every layer (for comparisons).
bool conv_uses_lds(DType dt, const ConvDesc &d) {
    static const bool direct = [] {
        const char *e = std::getenv("TDM_CONV_IMPL");
        return e && std::strcmp(e, "direct") == 0;
    }();
    // channels per 64-byte half stage
    const int sk = dt == DType::F32 ? 16 : dt == DType::F8 ? 64 : 32;
    if (dt == DType::F8) return d.C % sk == 0 && d.N % 64 == 0;
    return !direct && d.C % (2 * sk) == 0 && d.N % 64 == 0;
}

// Reduction steps in 128-byte units of k (the unit split-K divides).
int conv_reduction_steps(DType dt, const ConvDesc &d) {
    const int taps = d.ksize * d.ksize * d.ksize;
    if (conv_uses_lds(dt, d)) {
        // units of the staged kernels' reduction order (KWalk)
        const int cb = d.C * (int)dtype_size(dt);
        return std::max(1, taps * (cb / (cb % 128 == 0 ? 128 : 64)));
    }
    return taps * d.C / (dt == DType::F32 ? 16 : 32);
}

// Tile variants: {waves along m, wave tile m, wave tile n, stage bytes};
// 0-8 stage through VGPRs (conv_lds_kernel), 9-13 copy straight into LDS
// (conv_dma_kernel; 9-11 with three stage buffers, 12-13 with two), 14 the
// pooled tap-group kernel (conv_pool3_kernel), 15 its eight-wave form
// (conv_pool3g_kernel), 16 its resident-workgroup form (conv_pool3p_kernel).
namespace {
constexpr int kVariants[kConvVariants][4] = {
    {2, 64, 32, 128},  {2, 64, 64, 128}, {4, 64, 64, 128}, {2, 128, 32, 128}, {2, 128, 64, 128},
    {1, 64, 64, 128},  {4, 64, 64, 64},  {2, 64, 64, 64},  {2, 128, 32, 64},  {4, 64, 64, 64},
    {2, 64, 64, 64},   {2, 128, 32, 64}, {2, 64, 64, 128}, {4, 64, 64, 128}, {4, 48, 64, 128},
    {4, 96, 64, 128},   // 15: eight waves of 48 rows, listed as a 384 x 64 tile
    {4, 48, 64, 128}};  // 16: 14's tiles, resident workgroups
}

bool conv_variant_valid(DType dt, int v, const ConvDesc &d) {
    if (v < 0 || v >= kConvVariants) return false;
    const int bn = (4 / kVariants[v][0]) * kVariants[v][2];
    const int stage_k = kVariants[v][3] / (int)dtype_size(dt);
    if (d.N % bn != 0 || d.C % stage_k != 0) return false;
    // pooled tap groups: 16-bit, a fused pool after a 3x3x3 window, pooled
    // extents of at least 12 (a tile's windows then read at most 27 / 53
    // slots of their window lines), and one 128-byte channel unit (the only
    // form the descriptor net has, and so the only one tested)
    if ((v == 14 || v == 15 || v == 16) && (dtype_size(dt) != 2 || !d.pool || d.ksize != 3 ||
                                 d.OD / 2 < 12 || d.C * (int)dtype_size(dt) != 128))
        return false;
    if (v >= 9) {
        // conv_dma_kernel addresses a tile's input rows and weight rows with
        // 32-bit byte offsets from scalar bases: the input of the cubes a
        // tile's rows can span, and the weights, must stay below 4 GB
        const long bm = (long)kVariants[v][0] * kVariants[v][1];
        const long ocube = (long)d.OD * d.OD * d.OD;
        const long icube = (long)d.D * d.D * d.D * d.C * (long)dtype_size(dt);
        const long K = (long)d.ksize * d.ksize * d.ksize * d.C;
        if ((bm / ocube + 2) * icube >= (1L << 32)) return false;
        if ((long)d.N * K * (long)dtype_size(dt) >= (1L << 32)) return false;
    }
    return true;
}

int conv_default_variant(DType dt, const ConvDesc &d) {
    const int v = d.N % 128 == 0 ? 1 : 0;
    if (conv_variant_valid(dt, v, d)) return v;
    return d.N % 128 == 0 ? 7 : 6;  // 64-byte stages (fp8 with 64 channels)
}

void conv_variant_tile(int v, int *bm, int *bn) {
    *bm = kVariants[v][0] * kVariants[v][1];
    *bn = (4 / kVariants[v][0]) * kVariants[v][2];
}

uint16_t float_to_f16(float f) {
    const _Float16 h = (_Float16)f;  // IEEE conversion, round to nearest even
    uint16_t u;
    std::memcpy(&u, &h, 2);
    return u;
}

uint16_t float_to_bf16(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7F800000u) == 0x7F800000u) return (uint16_t)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

int launch_conv(DType dt, bool out_f32, const ConvDesc &d, int batch, const void *in,
                const void *w, const float *bias, void *out, int splits, float *partial,
                int variant, hipStream_t s, const float *colscale, float oscale) {
    TDM_CHECK(d.C % 32 == 0 && d.N % 64 == 0, TDM_ERR_UNSUPPORTED,
              "convolution with %d->%d channels (need multiples of 32 and 64)", d.C, d.N);
    TDM_CHECK(d.ksize == 3 || d.ksize == 1, TDM_ERR_UNSUPPORTED, "window %d", d.ksize);
    TDM_CHECK(d.OD == d.D - d.ksize + 1 && d.OD > 0, TDM_ERR_INVALID_ARG, "bad extents");
    TDM_CHECK(!d.pool || (d.OD % 2 == 0 && splits <= 1), TDM_ERR_INVALID_ARG, "bad fused pool");
    TDM_CHECK(splits >= 1 && splits <= 65535 && (splits == 1 || partial), TDM_ERR_INVALID_ARG,
              "bad split");
    if (batch <= 0) return TDM_OK;
    if (dt == DType::F8) {
        TDM_CHECK(colscale, TDM_ERR_INVALID_ARG, "fp8 convolution without scales");
        if (out_f32)
            return conv_dispatch<fp8_t, float>(d, batch, in, w, bias, out, splits, partial,
                                               variant, s, colscale, oscale);
        return conv_dispatch<fp8_t, fp8_t>(d, batch, in, w, bias, out, splits, partial, variant,
                                           s, colscale, oscale);
    }
    if (dt == DType::F32) return conv_dispatch<float, float>(d, batch, in, w, bias, out, splits, partial, variant, s, nullptr, 1.0f);
    if (dt == DType::F16) {
        if (out_f32) return conv_dispatch<f16_t, float>(d, batch, in, w, bias, out, splits, partial, variant, s, nullptr, 1.0f);
        return conv_dispatch<f16_t, f16_t>(d, batch, in, w, bias, out, splits, partial, variant, s, nullptr, 1.0f);
    }
    if (out_f32) return conv_dispatch<bf16_t, float>(d, batch, in, w, bias, out, splits, partial, variant, s, nullptr, 1.0f);
    return conv_dispatch<bf16_t, bf16_t>(d, batch, in, w, bias, out, splits, partial, variant, s, nullptr, 1.0f);
}

int launch_conv1(DType dt, bool out_f32, const VolumeSource &src, int N, bool relu, int batch,
                 const void *w, const float *bias, void *out, hipStream_t s, float oscale) {
    TDM_CHECK(N % 64 == 0, TDM_ERR_UNSUPPORTED, "first convolution needs a multiple of 64 outputs");
    TDM_CHECK(src.P >= 3, TDM_ERR_INVALID_ARG, "patch extent %d", src.P);
    // the kernel reads four consecutive biases at a time
    TDM_CHECK(reinterpret_cast<uintptr_t>(bias) % 16 == 0, TDM_ERR_INVALID_ARG,
              "first convolution's bias is not 16-byte aligned");
    if (batch <= 0) return TDM_OK;
    Conv1Args a;
    a.base = src.base;
    a.kp = src.kp;
    a.dx = src.dims[0];
    a.dy = src.dims[1];
    a.dz = src.dims[2];
    a.P = src.P;
    a.w = w;
    a.bias = bias;
    a.out = out;
    a.N = N;
    a.relu = relu ? 1 : 0;
    a.oscale = oscale;
    const long OD = src.P - 2;
    a.M = (long)batch * OD * OD * OD;
    fast_div((int)OD, &a.od_mul, &a.od_sh);
    dim3 grid((unsigned)((a.M + 255) / 256), N / 64);
    a.grids = src.grids;
    a.gid = src.gid;
    const int mode = src.kp == nullptr ? 0 : src.grids == nullptr ? 1 : 2;
    // the kernel forms single-grid tap offsets in 32 bits
    TDM_CHECK(mode != 1 || 2L * a.dx * a.dy + 2L * a.dx + 2 < (1L << 31), TDM_ERR_INVALID_ARG,
              "grid slices of %d x %d voxels are too large", a.dx, a.dy);
#define TDM_C1(T, OT)                                                                         \
    do {                                                                                      \
        if (mode == 2) conv1_kernel<T, OT, 2><<<grid, 256, 0, s>>>(a);                        \
        else if (mode == 1) conv1_kernel<T, OT, 1><<<grid, 256, 0, s>>>(a);                   \
        else conv1_kernel<T, OT, 0><<<grid, 256, 0, s>>>(a);                                  \
    } while (0)
    if (dt == DType::F32) TDM_C1(float, float);
    else if (dt == DType::F8 && out_f32) TDM_C1(f16_t, float);
    else if (dt == DType::F8) TDM_C1(f16_t, fp8_t);
    else if (dt == DType::F16 && out_f32) TDM_C1(f16_t, float);
    else if (dt == DType::F16) TDM_C1(f16_t, f16_t);
    else if (out_f32) TDM_C1(bf16_t, float);
    else TDM_C1(bf16_t, bf16_t);
#undef TDM_C1
    TDM_LAUNCHED();
    return TDM_OK;
}

int launch_to_ncdhw(DType dt, bool in_f32, const void *in, int batch, int C, int E, float *out,
                    hipStream_t s, float scale) {
    const long total = (long)batch * C * E * E * E;
    if (total <= 0) return TDM_OK;
    unsigned blocks = (unsigned)std::min<long>((total + 255) / 256, 65536);
    if (dt == DType::F32 || in_f32)
        to_ncdhw_kernel<float><<<blocks, 256, 0, s>>>(static_cast<const float *>(in), batch, C, E, out, 1.0f);
    else if (dt == DType::F8)
        to_ncdhw_kernel<fp8_t><<<blocks, 256, 0, s>>>(static_cast<const fp8_t *>(in), batch, C, E, out, scale);
    else if (dt == DType::F16)
        to_ncdhw_kernel<f16_t><<<blocks, 256, 0, s>>>(static_cast<const f16_t *>(in), batch, C, E, out, 1.0f);
    else
        to_ncdhw_kernel<bf16_t><<<blocks, 256, 0, s>>>(static_cast<const bf16_t *>(in), batch, C, E, out, 1.0f);
    TDM_LAUNCHED();
    return TDM_OK;
}

// Row r of `in` (per_row floats) times inv_scale[r], converted to e4m3.
__global__ void quantize_fp8_kernel(const float *__restrict__ in, long rows, long per_row,
                                    const float *__restrict__ inv_scale, fp8_t *__restrict__ out) {
    const long total = rows * per_row;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x)
        out[i] = to_fp8(in[i] * inv_scale[i / per_row]);
}

__global__ void absmax_kernel(const float *__restrict__ x, long n, unsigned *__restrict__ amax) {
    float m = 0.0f;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(x[i]));
    for (int o = 32; o >= 1; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(amax, __float_as_uint(m));
}

int launch_quantize_fp8(const float *in, long rows, long per_row, const float *inv_scale,
                        void *out, hipStream_t s) {
    const long total = rows * per_row;
    if (total <= 0) return TDM_OK;
    const unsigned blocks = (unsigned)std::min<long>((total + 255) / 256, 65536);
    quantize_fp8_kernel<<<blocks, 256, 0, s>>>(in, rows, per_row, inv_scale,
                                               static_cast<fp8_t *>(out));
    TDM_LAUNCHED();
    return TDM_OK;
}

int launch_absmax(const float *x, long n, unsigned *amax, hipStream_t s) {
    if (n <= 0) return TDM_OK;
    const unsigned blocks = (unsigned)std::min<long>((n + 255) / 256, 4096);
    absmax_kernel<<<blocks, 256, 0, s>>>(x, n, amax);
    TDM_LAUNCHED();
    return TDM_OK;
}

int launch_maxpool(DType dt, const void *in, int batch, int C, int E, void *out, hipStream_t s) {
    const int OE = 1 + (E - 2 + 1) / 2;  // 1 + ceil((E - 2) / 2)
    const long total = (long)batch * OE * OE * OE * C;
    if (total <= 0) return TDM_OK;
    unsigned blocks = (unsigned)std::min<long>((total + 255) / 256, 65536);
    if (dt == DType::F32)
        maxpool_kernel<float><<<blocks, 256, 0, s>>>(static_cast<const float *>(in), batch, C, E,
                                                     OE, static_cast<float *>(out));
    else if (dt == DType::F16)
        maxpool_kernel<f16_t><<<blocks, 256, 0, s>>>(static_cast<const f16_t *>(in), batch, C, E,
                                                     OE, static_cast<f16_t *>(out));
    else if (dt == DType::F8)  // exact: the max of e4m3 values is one of them
        maxpool_kernel<fp8_t><<<blocks, 256, 0, s>>>(static_cast<const fp8_t *>(in), batch, C, E,
                                                     OE, static_cast<fp8_t *>(out));
    else
        maxpool_kernel<bf16_t><<<blocks, 256, 0, s>>>(static_cast<const bf16_t *>(in), batch, C,
                                                      E, OE, static_cast<bf16_t *>(out));
    TDM_LAUNCHED();
    return TDM_OK;
}

}  // namespace tdm
