// Drop-in for core/demo (core/demo.cu:63-279): point cloud -> TDF grid ->
// random surface keypoints -> 3DMatch descriptors, all on the GPU.
//
//   descriptor_demo <cloud.ply> <out_prefix> [--num-keypoints N] [--seed S]
//       [--keypoint-indices FILE] [--net JSON] [--weights MARVIN]
//       [--precision fp32|fp16|bf16|fp8] [--decode exact|reference]
//       [--tdf-bin PATH]
//       [--batch B]
//
// Writes <out_prefix>.keypts.bin and <out_prefix>.desc.3dmatch.bin in the
// reference layouts.  The reference draws 500 keypoints from a fresh
// std::random_device on every call (core/utils.hpp:15-20); here the draw uses
// the same distribution, U[0, N - 1e-4) truncated to an index, from an
// mt19937 with an explicit seed, or an explicit index file (one index per
// line).  Defaults: net 3dmatch-net-test.json, weights
// 3dmatch-weights-snapshot-137000.marvin (both in the working directory).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <string>
#include <vector>

#include "cli_common.hpp"
#include "tdm.h"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <cloud.ply> <out_prefix> [options]\n", argv[0]);
        return 1;
    }
    const std::string cloud = argv[1], prefix = argv[2];
    int num_keypts = 50 * 10;
    unsigned seed = 0;
    bool have_seed = false;
    std::string idx_file, net_json = "3dmatch-net-test.json",
                      weights = "3dmatch-weights-snapshot-137000.marvin", tdf_bin;
    int precision = TDM_PREC_FP32, decode = TDM_DECODE_EXACT, batch = 0;
    for (int i = 3; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&]() -> std::string { return i + 1 < argc ? argv[++i] : std::string(); };
        if (a == "--num-keypoints") num_keypts = std::atoi(next().c_str());
        else if (a == "--seed") { seed = (unsigned)std::strtoul(next().c_str(), nullptr, 10); have_seed = true; }
        else if (a == "--keypoint-indices") idx_file = next();
        else if (a == "--net") net_json = next();
        else if (a == "--weights") weights = next();
        else if (a == "--tdf-bin") tdf_bin = next();
        else if (a == "--batch") batch = std::atoi(next().c_str());
        else if (a == "--precision") {
            std::string p = next();
            if (!parse_precision(p, &precision)) {
                std::fprintf(stderr, "unknown precision %s (fp32|fp16|bf16|fp8)\n", p.c_str());
                return 1;
            }
        } else if (a == "--decode") {
            std::string d = next();
            if (d == "exact") decode = TDM_DECODE_EXACT;
            else if (d == "reference") decode = TDM_DECODE_REFERENCE;
            else { std::fprintf(stderr, "unknown decode %s\n", d.c_str()); return 1; }
        } else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 1;
        }
    }

    float *pts = nullptr;
    int num_pts = 0;
    CHECK(tdm_read_ply_points(cloud.c_str(), &pts, &num_pts));
    std::printf("Loaded point cloud with %d points!\n", num_pts);

    const float voxel_size = 0.01f;           // core/demo.cu:97-99
    const float trunc_margin = voxel_size * 5;
    const int padding = 15;

    float *d_pts = nullptr;
    HIPCHECK(hipMalloc(&d_pts, sizeof(float) * 3 * (size_t)num_pts));
    HIPCHECK(hipMemcpy(d_pts, pts, sizeof(float) * 3 * (size_t)num_pts, hipMemcpyHostToDevice));
    int dims[3];
    float origin[3];
    CHECK(tdm_grid_params(nullptr, d_pts, num_pts, voxel_size, padding, dims, origin, nullptr));
    std::printf("Size of TDF voxel grid: %d x %d x %d\n", dims[0], dims[1], dims[2]);
    std::printf("Computing TDF voxel grid...\n");
    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    float *d_tdf = nullptr;
    HIPCHECK(hipMalloc(&d_tdf, sizeof(float) * nvox));
    CHECK(tdm_tdf_from_points(nullptr, d_pts, num_pts, dims, origin, voxel_size, trunc_margin, decode,
                              d_tdf, nullptr));

    // Keypoints (core/demo.cu:164-177).
    std::printf("Finding random surface keypoints...\n");
    std::vector<int> idx;
    if (!idx_file.empty()) {
        std::ifstream in(idx_file);
        long v;
        while (in >> v) {
            if (v < 0 || v >= num_pts) {
                std::fprintf(stderr, "keypoint index %ld out of range\n", v);
                return 1;
            }
            idx.push_back((int)v);
        }
        num_keypts = (int)idx.size();
    } else {
        std::mt19937 mt(have_seed ? seed : std::random_device{}());
        std::uniform_real_distribution<double> dist(0.0, (double)((float)num_pts - 0.0001f));
        for (int k = 0; k < num_keypts; ++k) idx.push_back((int)dist(mt));
    }
    std::vector<float> keypts((size_t)num_keypts * 3);
    for (int k = 0; k < num_keypts; ++k)
        for (int a = 0; a < 3; ++a) keypts[(size_t)k * 3 + a] = pts[(size_t)idx[k] * 3 + a];
    std::vector<int> kp_grid((size_t)num_keypts * 3);
    CHECK(tdm_keypoints_to_grid(keypts.data(), num_keypts, origin, voxel_size, kp_grid.data()));

    tdm_net *net = nullptr;
    CHECK(tdm_net_create(nullptr, net_json.c_str(), precision, batch, &net));
    CHECK(tdm_net_load_weights(net, weights.c_str()));
    std::printf("3DMatch network architecture successfully loaded!\n");
    const int desc_size = tdm_net_output_dim(net);
    int *d_kp = nullptr;
    float *d_desc = nullptr;
    HIPCHECK(hipMalloc(&d_kp, sizeof(int) * kp_grid.size() + 4));
    HIPCHECK(hipMalloc(&d_desc, sizeof(float) * (size_t)num_keypts * desc_size + 4));
    HIPCHECK(hipMemcpy(d_kp, kp_grid.data(), sizeof(int) * kp_grid.size(), hipMemcpyHostToDevice));
    if (precision == TDM_PREC_FP8 && num_keypts > 0) {
        const int nc = std::min(num_keypts, kCalibrationPatches), P = net_patch_extent(net);
        float *d_cal = nullptr;
        HIPCHECK(hipMalloc(&d_cal, sizeof(float) * (size_t)nc * P * P * P));
        CHECK(tdm_extract_patches(nullptr, d_tdf, dims[0], dims[1], dims[2], d_kp, nc, P, d_cal,
                                  nullptr));
        CHECK(calibrate_fp8(net, net_json, weights, d_cal, nc));
        (void)hipFree(d_cal);
    }
    std::printf("Computing 3DMatch descriptors for %d keypoints...\n", num_keypts);
    CHECK(tdm_net_forward_grid(net, d_tdf, dims[0], dims[1], dims[2], d_kp, num_keypts, d_desc, nullptr));
    std::vector<float> desc((size_t)num_keypts * desc_size);
    HIPCHECK(hipMemcpy(desc.data(), d_desc, sizeof(float) * desc.size(), hipMemcpyDeviceToHost));

    std::printf("Saving keypoints to disk (keypts.bin)...\n");
    CHECK(tdm_write_keypts_bin((prefix + ".keypts.bin").c_str(), keypts.data(), num_keypts));
    std::printf("Saving 3DMatch descriptors to disk (desc.3dmatch.bin)...\n");
    CHECK(tdm_write_desc_bin((prefix + ".desc.3dmatch.bin").c_str(), desc.data(), num_keypts, desc_size));
    if (!tdf_bin.empty()) {
        std::vector<float> tdf(nvox);
        HIPCHECK(hipMemcpy(tdf.data(), d_tdf, sizeof(float) * nvox, hipMemcpyDeviceToHost));
        CHECK(tdm_write_grid_bin(tdf_bin.c_str(), dims, origin, voxel_size, trunc_margin, tdf.data()));
    }
    tdm_net_destroy(net);
    tdm_free(pts);
    (void)hipFree(d_pts);
    (void)hipFree(d_tdf);
    (void)hipFree(d_kp);
    (void)hipFree(d_desc);
    return 0;
}
