// Drop-in for depth-fusion/demo (depth-fusion/demo.cu:72-216): fuses
// frame-%06d.depth.png / .pose.txt into a TSDF grid on the GPU and writes
// tsdf.ply (surface voxels |tsdf| < 0.2, weight > 1).
//
//   fusion_demo [camK dataPath baseIdx firstIdx nFrames ox oy oz voxelSize truncMargin]
//               [--dims X Y Z] [--tsdf-bin PATH] [--weight-bin PATH] [--fused] [--out PLY]
//
// Positional arguments and defaults are the reference's (500^3 grid at 6 mm,
// origin (-1.5, -1.5, 0.5), truncation 5 voxels).  --fused uploads all frames
// and integrates them in one pass (bit-identical result).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

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

static std::string frame_prefix(const std::string &dir, int idx) {
    char buf[32];
    std::snprintf(buf, sizeof buf, "/frame-%06d", idx);
    return dir + buf;
}

int main(int argc, char **argv) {
    std::string cam_K_file = "../data/sample/depth-fusion-demo/camera-intrinsics.txt";
    std::string data_path = "../data/sample/depth-fusion-demo/rgbd-frames";
    int base_frame_idx = 8, first_frame_idx = 8;
    float num_frames = 50;
    int H = 480, W = 640;
    float ox = -1.5f, oy = -1.5f, oz = 0.5f;
    float voxel_size = 0.006f;
    float trunc_margin = voxel_size * 5;
    int dims[3] = {500, 500, 500};
    std::string tsdf_bin, weight_bin, out_ply = "tsdf.ply";
    bool fused = false;

    std::vector<std::string> pos;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--dims" && i + 3 < argc) {
            for (int k = 0; k < 3; ++k) dims[k] = std::atoi(argv[++i]);
        } else if (a == "--tsdf-bin" && i + 1 < argc) {
            tsdf_bin = argv[++i];
        } else if (a == "--weight-bin" && i + 1 < argc) {
            weight_bin = argv[++i];
        } else if (a == "--out" && i + 1 < argc) {
            out_ply = argv[++i];
        } else if (a == "--fused") {
            fused = true;
        } else {
            pos.push_back(a);
        }
    }
    if (!pos.empty()) {  // depth-fusion/demo.cu:102-113
        if (pos.size() < 10) {
            std::fprintf(stderr, "usage: %s camK dataPath baseIdx firstIdx nFrames ox oy oz voxelSize truncMargin\n", argv[0]);
            return 1;
        }
        cam_K_file = pos[0];
        data_path = pos[1];
        base_frame_idx = std::atoi(pos[2].c_str());
        first_frame_idx = std::atoi(pos[3].c_str());
        num_frames = (float)std::atof(pos[4].c_str());
        ox = (float)std::atof(pos[5].c_str());
        oy = (float)std::atof(pos[6].c_str());
        oz = (float)std::atof(pos[7].c_str());
        voxel_size = (float)std::atof(pos[8].c_str());
        trunc_margin = (float)std::atof(pos[9].c_str());
    }

    float cam_K[9], base2world[16], base2world_inv[16];
    CHECK(tdm_load_matrix(cam_K_file.c_str(), 3, 3, cam_K));
    CHECK(tdm_load_matrix((frame_prefix(data_path, base_frame_idx) + ".pose.txt").c_str(), 4, 4, base2world));
    CHECK(tdm_invert_matrix(base2world, base2world_inv));

    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    float *d_tsdf = nullptr, *d_weight = nullptr, *d_K = nullptr;
    HIPCHECK(hipMalloc(&d_tsdf, nvox * sizeof(float)));
    HIPCHECK(hipMalloc(&d_weight, nvox * sizeof(float)));
    HIPCHECK(hipMalloc(&d_K, sizeof cam_K));
    HIPCHECK(hipMemcpy(d_K, cam_K, sizeof cam_K, hipMemcpyHostToDevice));
    CHECK(tdm_tsdf_init(nullptr, d_tsdf, d_weight, nvox, nullptr));

    const int nf = (int)num_frames;
    const size_t px = (size_t)H * W;
    std::vector<float> depth(px);
    float *h_depth = nullptr, *h_pose = nullptr;  // pinned staging
    const int stage = fused ? std::max(nf, 1) : 1;
    HIPCHECK(hipHostMalloc(&h_depth, px * stage * sizeof(float), hipHostMallocDefault));
    HIPCHECK(hipHostMalloc(&h_pose, 16 * stage * sizeof(float), hipHostMallocDefault));
    float *d_depth = nullptr, *d_pose = nullptr;
    HIPCHECK(hipMalloc(&d_depth, px * stage * sizeof(float)));
    HIPCHECK(hipMalloc(&d_pose, 16 * stage * sizeof(float)));

    for (int f = 0; f < nf; ++f) {  // depth-fusion/demo.cu:156-183
        const std::string pre = frame_prefix(data_path, first_frame_idx + f);
        const std::string depth_file = pre + ".depth.png";
        const int slot = fused ? f : 0;
        CHECK(tdm_read_depth_png(depth_file.c_str(), H, W, h_depth + slot * px));
        float cam2world[16];
        CHECK(tdm_load_matrix((pre + ".pose.txt").c_str(), 4, 4, cam2world));
        CHECK(tdm_multiply_matrix(base2world_inv, cam2world, h_pose + slot * 16));
        std::printf("Fusing: %s\n", depth_file.c_str());
        if (!fused) {
            HIPCHECK(hipMemcpy(d_pose, h_pose, 16 * sizeof(float), hipMemcpyHostToDevice));
            HIPCHECK(hipMemcpy(d_depth, h_depth, px * sizeof(float), hipMemcpyHostToDevice));
            CHECK(tdm_integrate(nullptr, d_K, d_pose, d_depth, H, W, dims[0], dims[1], dims[2], ox, oy,
                                oz, voxel_size, trunc_margin, d_tsdf, d_weight, nullptr));
        }
    }
    if (fused && nf > 0) {
        HIPCHECK(hipMemcpy(d_pose, h_pose, 16 * nf * sizeof(float), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_depth, h_depth, px * nf * sizeof(float), hipMemcpyHostToDevice));
        CHECK(tdm_integrate_frames(nullptr, d_K, d_pose, d_depth, nf, H, W, dims[0], dims[1], dims[2],
                                   ox, oy, oz, voxel_size, trunc_margin, d_tsdf, d_weight, nullptr));
    }

    std::vector<float> tsdf(nvox), weight(nvox);
    HIPCHECK(hipMemcpy(tsdf.data(), d_tsdf, nvox * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(weight.data(), d_weight, nvox * sizeof(float), hipMemcpyDeviceToHost));
    std::printf("Saving surface point cloud (%s)...\n", out_ply.c_str());
    const float origin[3] = {ox, oy, oz};
    int npts = 0;
    CHECK(tdm_save_surface_ply(out_ply.c_str(), dims, voxel_size, origin, tsdf.data(), weight.data(),
                               0.2f, 1.0f, &npts));
    if (!tsdf_bin.empty()) {
        std::printf("Saving TSDF voxel grid values to disk (%s)...\n", tsdf_bin.c_str());
        CHECK(tdm_write_grid_bin(tsdf_bin.c_str(), dims, origin, voxel_size, trunc_margin, tsdf.data()));
    }
    if (!weight_bin.empty())
        CHECK(tdm_write_grid_bin(weight_bin.c_str(), dims, origin, voxel_size, trunc_margin, weight.data()));
    (void)hipHostFree(h_depth);
    (void)hipHostFree(h_pose);
    (void)hipFree(d_depth);
    (void)hipFree(d_pose);
    (void)hipFree(d_tsdf);
    (void)hipFree(d_weight);
    (void)hipFree(d_K);
    return 0;
}
