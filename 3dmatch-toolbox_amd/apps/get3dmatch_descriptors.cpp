// Drop-in for the evaluation descriptor tools:
//   get3dmatch_descriptors data.bin
//       (evaluation/keypoint-matching/get3DMatchDescriptors.cu): patch volumes
//       [N, dx, dy, dz, N*dx*dy*dz floats] -> feat.bin [N, 512, 1, 1, 1, ...];
//   get3dmatch_descriptors TDF.bin keypoints.bin
//       (evaluation/mesh-correspondence-shape2pose/get3DMatchDescriptors.cu):
//       grid [dx, dy, dz, grid] + keypoints [N, N*3 grid coordinates, rounded]
//       -> <keypoints without .bin>.3dmatch.descriptors.bin [N, 512, ...].
// Options: --net JSON (default tmp.json), --weights MARVIN (default
// 3dmatch-weights-snapshot-137000.marvin), --precision fp32|fp16|bf16|fp8
// (default fp32; fp16 is the reference build's half storage), --batch B.
// Every volume is processed (the reference drops the N mod 50 remainder).
// Volumes must be patch^3 with the net's patch extent (30).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "cli_common.hpp"
#include "tdm.h"

static int read_floats(const char *path, std::vector<float> &v) {
    FILE *f = std::fopen(path, "rb");
    if (!f) return 1;
    std::fseek(f, 0, SEEK_END);
    long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    v.resize((size_t)n / 4);
    size_t got = std::fread(v.data(), 4, v.size(), f);
    std::fclose(f);
    return got == v.size() ? 0 : 1;
}

int main(int argc, char **argv) {
    std::vector<std::string> pos;
    std::string net_json = "tmp.json", weights = "3dmatch-weights-snapshot-137000.marvin";
    int precision = TDM_PREC_FP32, batch = 0;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--net" && i + 1 < argc) net_json = argv[++i];
        else if (a == "--weights" && i + 1 < argc) weights = argv[++i];
        else if (a == "--batch" && i + 1 < argc) batch = std::atoi(argv[++i]);
        else if (a == "--precision" && i + 1 < argc) {
            std::string p = argv[++i];
            if (!parse_precision(p, &precision)) {
                std::fprintf(stderr, "unknown precision %s (fp32|fp16|bf16|fp8)\n", p.c_str());
                return 1;
            }
        } else if (a.size() > 2 && a.compare(0, 2, "--") == 0) {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return 1;
        } else pos.push_back(a);
    }
    if (pos.empty() || pos.size() > 2) {
        std::fprintf(stderr, "usage: %s data.bin | TDF.bin keypoints.bin [options]\n", argv[0]);
        return 1;
    }
    tdm_net *net = nullptr;
    CHECK(tdm_net_create(nullptr, net_json.c_str(), precision, batch, &net));
    CHECK(tdm_net_load_weights(net, weights.c_str()));
    const int D = tdm_net_output_dim(net);

    if (pos.size() == 1) {
        std::printf("%s\n", pos[0].c_str());
        float *data = nullptr;
        int n = 0, dims[3];
        CHECK(tdm_read_patch_volumes(pos[0].c_str(), &data, &n, dims));
        std::printf("%d\n", n);
        const int P = net_patch_extent(net);
        if (dims[0] != P || dims[1] != P || dims[2] != P) {
            std::fprintf(stderr, "%s holds %dx%dx%d volumes; the net takes %d^3 patches\n",
                         pos[0].c_str(), dims[0], dims[1], dims[2], P);
            return 1;
        }
        const size_t per = (size_t)dims[0] * dims[1] * dims[2];
        float *d_data = nullptr, *d_feat = nullptr;
        HIPCHECK(hipMalloc(&d_data, sizeof(float) * per * n + 4));
        HIPCHECK(hipMalloc(&d_feat, sizeof(float) * (size_t)n * D + 4));
        HIPCHECK(hipMemcpy(d_data, data, sizeof(float) * per * n, hipMemcpyHostToDevice));
        if (precision == TDM_PREC_FP8 && n > 0)
            CHECK(calibrate_fp8(net, net_json, weights, d_data, std::min(n, kCalibrationPatches)));
        CHECK(tdm_net_forward(net, d_data, n, d_feat, nullptr));
        std::vector<float> feat((size_t)n * D);
        HIPCHECK(hipMemcpy(feat.data(), d_feat, sizeof(float) * feat.size(), hipMemcpyDeviceToHost));
        CHECK(tdm_write_feat_bin("feat.bin", feat.data(), n, D));
        tdm_free(data);
        (void)hipFree(d_data);
        (void)hipFree(d_feat);
    } else {
        std::vector<float> g, k;
        if (read_floats(pos[0].c_str(), g) || g.size() < 3) {
            std::fprintf(stderr, "cannot read %s\n", pos[0].c_str());
            return 1;
        }
        for (int a = 0; a < 3; ++a)
            if (!(g[a] >= 1.0f && g[a] <= 1e6f)) {
                std::fprintf(stderr, "%s: bad grid dimension %g\n", pos[0].c_str(), g[a]);
                return 1;
            }
        const int dims[3] = {(int)g[0], (int)g[1], (int)g[2]};
        std::printf("%d %d %d\n", dims[0], dims[1], dims[2]);
        const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
        if (g.size() < 3 + nvox) {
            std::fprintf(stderr, "%s is truncated\n", pos[0].c_str());
            return 1;
        }
        if (read_floats(pos[1].c_str(), k) || k.empty()) {
            std::fprintf(stderr, "cannot read %s\n", pos[1].c_str());
            return 1;
        }
        if (!(k[0] >= 0.0f && k[0] <= 1e8f)) {
            std::fprintf(stderr, "%s: bad keypoint count %g\n", pos[1].c_str(), k[0]);
            return 1;
        }
        const int n = (int)k[0];
        if (k.size() < 1 + (size_t)n * 3) {
            std::fprintf(stderr, "%s is truncated\n", pos[1].c_str());
            return 1;
        }
        std::printf("%d\n", n);
        std::vector<int> kp((size_t)n * 3);
        for (size_t i = 0; i < kp.size(); ++i) kp[i] = (int)std::round(k[1 + i]);
        float *d_tdf = nullptr, *d_feat = nullptr;
        int *d_kp = nullptr;
        HIPCHECK(hipMalloc(&d_tdf, sizeof(float) * nvox));
        HIPCHECK(hipMalloc(&d_kp, sizeof(int) * kp.size() + 4));
        HIPCHECK(hipMalloc(&d_feat, sizeof(float) * (size_t)n * D + 4));
        HIPCHECK(hipMemcpy(d_tdf, g.data() + 3, sizeof(float) * nvox, hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(d_kp, kp.data(), sizeof(int) * kp.size(), hipMemcpyHostToDevice));
        if (precision == TDM_PREC_FP8 && n > 0) {
            const int nc = std::min(n, kCalibrationPatches), P = net_patch_extent(net);
            float *d_cal = nullptr;
            HIPCHECK(hipMalloc(&d_cal, sizeof(float) * (size_t)nc * P * P * P));
            CHECK(tdm_extract_patches(nullptr, d_tdf, dims[0], dims[1], dims[2], d_kp, nc, P, d_cal,
                                      nullptr));
            CHECK(calibrate_fp8(net, net_json, weights, d_cal, nc));
            (void)hipFree(d_cal);
        }
        CHECK(tdm_net_forward_grid(net, d_tdf, dims[0], dims[1], dims[2], d_kp, n, d_feat, nullptr));
        std::vector<float> feat((size_t)n * D);
        HIPCHECK(hipMemcpy(feat.data(), d_feat, sizeof(float) * feat.size(), hipMemcpyDeviceToHost));
        std::string out = pos[1].substr(0, pos[1].size() >= 4 ? pos[1].size() - 4 : 0) + ".3dmatch.descriptors.bin";
        CHECK(tdm_write_desc_bin(out.c_str(), feat.data(), n, D));
        (void)hipFree(d_tdf);
        (void)hipFree(d_kp);
        (void)hipFree(d_feat);
    }
    tdm_net_destroy(net);
    return 0;
}
