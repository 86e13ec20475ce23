// Drop-in for depth-fusion/demo (depth-fusion/demo.cu:72-216): fuses
// frame-%06d.depth.png / .pose.txt into a TSDF grid on the GPU and writes
// tsdf.ply (surface voxels |tsdf| < 0.2, weight > 1).
//
//   fusion_demo [camK dataPath baseIdx firstIdx nFrames ox oy oz voxelSize truncMargin]
//               [--dims X Y Z] [--tsdf-bin PATH] [--weight-bin PATH] [--out PLY]
//               [--serial | --fused] [--threads N]
//
// Positional arguments and defaults are the reference's (500^3 grid at 6 mm,
// origin (-1.5, -1.5, 0.5), truncation 5 voxels).  By default the frames go
// through the overlapped pipeline (tdm_fusion_*: PNG decoder threads, pinned
// ring, u16 upload, conversion and Integrate on the GPU in frame order).
// --serial runs the reference's loop shape (decode, copy, integrate, one
// frame at a time); --fused uploads all frames and integrates them in one
// pass.  All three give bit-identical grids.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "cli_common.hpp"
#include "tdm.h"

static std::string frame_prefix(const std::string &dir, int idx) {
    char buf[32];
    std::snprintf(buf, sizeof buf, "/frame-%06d", idx);
    return dir + buf;
}

// tsdf.ply (depth-fusion/demo.cu:186-194) and the optional grid files.
static int write_outputs(const std::string &out_ply, const std::string &tsdf_bin,
                         const std::string &weight_bin, const int dims[3], const float origin[3],
                         float voxel_size, float trunc_margin, const std::vector<float> &tsdf,
                         const std::vector<float> &weight) {
    std::printf("Saving surface point cloud (%s)...\n", out_ply.c_str());
    int npts = 0;
    CHECK(tdm_save_surface_ply(out_ply.c_str(), dims, voxel_size, origin, tsdf.data(), weight.data(),
                               0.2f, 1.0f, &npts));
    if (!tsdf_bin.empty()) {
        std::printf("Saving TSDF voxel grid values to disk (%s)...\n", tsdf_bin.c_str());
        CHECK(tdm_write_grid_bin(tsdf_bin.c_str(), dims, origin, voxel_size, trunc_margin, tsdf.data()));
    }
    if (!weight_bin.empty())
        CHECK(tdm_write_grid_bin(weight_bin.c_str(), dims, origin, voxel_size, trunc_margin,
                                 weight.data()));
    return 0;
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
    bool fused = false, serial = false;
    int threads = 0;

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
        } else if (a == "--serial") {
            serial = true;
        } else if (a == "--threads" && i + 1 < argc) {
            threads = std::atoi(argv[++i]);
        } else if (a.size() > 2 && a.compare(0, 2, "--") == 0) {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 1;
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

    if (fused && serial) {
        std::fprintf(stderr, "--fused and --serial exclude each other\n");
        return 1;
    }
    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    const int nf = (int)num_frames;
    const float origin[3] = {ox, oy, oz};
    std::vector<float> tsdf(nvox), weight(nvox);
    if (!fused && !serial) {  // the frame loop of depth-fusion/demo.cu:156-183, overlapped
        tdm_fusion *fu = nullptr;
        CHECK(tdm_fusion_create(nullptr, cam_K, H, W, dims, origin, voxel_size, trunc_margin, threads,
                                &fu));
        for (int f = 0; f < nf; ++f) {
            const std::string pre = frame_prefix(data_path, first_frame_idx + f);
            const std::string depth_file = pre + ".depth.png";
            float cam2world[16], cam2base[16];
            CHECK(tdm_load_matrix((pre + ".pose.txt").c_str(), 4, 4, cam2world));
            CHECK(tdm_multiply_matrix(base2world_inv, cam2world, cam2base));
            std::printf("Fusing: %s\n", depth_file.c_str());
            CHECK(tdm_fusion_add_png(fu, depth_file.c_str(), cam2base));
        }
        CHECK(tdm_fusion_wait(fu));
        float *g_tsdf = nullptr, *g_weight = nullptr;
        CHECK(tdm_fusion_grids(fu, &g_tsdf, &g_weight));
        HIPCHECK(hipMemcpy(tsdf.data(), g_tsdf, nvox * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHECK(hipMemcpy(weight.data(), g_weight, nvox * sizeof(float), hipMemcpyDeviceToHost));
        CHECK(tdm_fusion_destroy(fu));
        return write_outputs(out_ply, tsdf_bin, weight_bin, dims, origin, voxel_size, trunc_margin,
                             tsdf, weight);
    }

    float *d_tsdf = nullptr, *d_weight = nullptr, *d_K = nullptr;
    HIPCHECK(hipMalloc(&d_tsdf, nvox * sizeof(float)));
    HIPCHECK(hipMalloc(&d_weight, nvox * sizeof(float)));
    HIPCHECK(hipMalloc(&d_K, sizeof cam_K));
    HIPCHECK(hipMemcpy(d_K, cam_K, sizeof cam_K, hipMemcpyHostToDevice));
    CHECK(tdm_tsdf_init(nullptr, d_tsdf, d_weight, nvox, nullptr));

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

    HIPCHECK(hipMemcpy(tsdf.data(), d_tsdf, nvox * sizeof(float), hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(weight.data(), d_weight, nvox * sizeof(float), hipMemcpyDeviceToHost));
    (void)hipHostFree(h_depth);
    (void)hipHostFree(h_pose);
    (void)hipFree(d_depth);
    (void)hipFree(d_pose);
    (void)hipFree(d_tsdf);
    (void)hipFree(d_weight);
    (void)hipFree(d_K);
    return write_outputs(out_ply, tsdf_bin, weight_bin, dims, origin, voxel_size, trunc_margin, tsdf,
                         weight);
}
