// Shared pieces of the drop-in command-line tools: error checks, the
// --precision option and the fp8 calibration step.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>

#include "tdm.h"

#define CHECK(call)                                                                        \
    do {                                                                                   \
        int s_ = (call);                                                                   \
        if (s_ != TDM_OK) {                                                                \
            std::fprintf(stderr, "error: %s: %s\n", tdm_status_string(s_), tdm_last_error()); \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)
#define HIPCHECK(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "HIP error: %s\n", hipGetErrorString(e_));                \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

// fp32: float storage and arithmetic (Marvin's DATATYPE 1); fp16: half
// storage with float accumulation, what the reference's descriptor tools
// compile by default (DATATYPE 0, core/marvin.hpp:11-27,
// evaluation/keypoint-matching/get3DMatchDescriptors.cu:1); bf16; fp8 (e4m3,
// calibrated on the run's own patches).  Anything else is an error.
inline bool parse_precision(const std::string &p, int *out) {
    if (p == "fp32") *out = TDM_PREC_FP32;
    else if (p == "fp16") *out = TDM_PREC_FP16;
    else if (p == "bf16") *out = TDM_PREC_BF16;
    else if (p == "fp8") *out = TDM_PREC_FP8;
    else return false;
    return true;
}

// Patches used to set an fp8 net's activation scales.
constexpr int kCalibrationPatches = 256;

// Sets the activation scales of an fp8 net from an fp32 net with the same
// layers and weights run on `num` patches (device, [num,1,P,P,P]).
inline int calibrate_fp8(tdm_net *net, const std::string &net_json, const std::string &weights,
                         const float *d_patches, int num) {
    tdm_net *ref = nullptr;
    int s = tdm_net_create(nullptr, net_json.c_str(), TDM_PREC_FP32, num, &ref);
    if (s == TDM_OK) s = tdm_net_load_weights(ref, weights.c_str());
    if (s == TDM_OK) s = tdm_net_calibrate(net, ref, d_patches, num, nullptr);
    if (ref) tdm_net_destroy(ref);
    return s;
}

// Input extent of the net's first layer (the patch size, 30).
inline int net_patch_extent(const tdm_net *net) {
    int info[8];
    if (tdm_net_layer_info(net, 0, info) != TDM_OK) return -1;
    return info[3];
}
