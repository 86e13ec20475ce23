// Depth-frame fusion with overlapped I/O: the frame loop of
// depth-fusion/demo.cu:156-183 (PNG decode, pose, Integrate) as a pipeline.
// Decoder threads turn frame-%06d.depth.png files into raw u16 millimetres in
// a ring of pinned slots; an issuer thread takes the slots in frame order and
// enqueues, on one stream, the u16 upload (half the bytes of the float
// frame), the millimetre -> metre conversion on the GPU
// (depth-fusion/utils.hpp:80-84) and the Integrate kernel.  Decoding frame
// f+k, copying frame f+1 and integrating frame f overlap; the grid sees the
// frames in submission order, so the result is bit-identical to the serial
// loop.
#include <hip/hip_runtime.h>

#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "pipeline.hpp"

namespace {

struct Slot {
    uint16_t *h_mm = nullptr;  // pinned
    float h_pose[16];
    uint16_t *d_mm = nullptr;
    float *d_depth = nullptr;
    float *d_pose = nullptr;
    hipEvent_t done = nullptr;  // GPU finished with this slot
    long seq = -1;              // frame held (-1: free)
    bool decoded = false;
    int status = TDM_OK;
};

struct Job {
    long seq;
    std::string path;
    float pose[16];
};

}  // namespace

struct tdm_fusion {
    int device = 0;
    int H = 0, W = 0;
    int dims[3] = {0, 0, 0};
    float origin[3] = {0, 0, 0};
    float vs = 0, trunc = 0;
    float *tsdf = nullptr, *weight = nullptr, *d_K = nullptr;
    hipStream_t stream = nullptr;
    std::vector<Slot> slots;

    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> jobs;      // waiting for a decoder
    long submitted = 0;        // frames submitted
    long issued = 0;           // frames enqueued on the GPU
    bool stopping = false;
    int error = TDM_OK;
    std::string error_msg;
    std::vector<std::thread> decoders;
    std::thread issuer;

    void decode_loop();
    void issue_loop();
    void set_error(int st, const std::string &msg) {
        if (error == TDM_OK) {
            error = st;
            error_msg = msg;
        }
    }
};

void tdm_fusion::decode_loop() {
    (void)hipSetDevice(device);
    std::vector<uint16_t> pix;
    for (;;) {
        Job job;
        Slot *slot = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            // the next job whose slot is free (slots are reused in frame order)
            cv.wait(lk, [&] {
                if (stopping && jobs.empty()) return true;
                if (jobs.empty()) return false;
                const Slot &s = slots[jobs.front().seq % slots.size()];
                return s.seq == -1;
            });
            if (jobs.empty()) return;
            job = std::move(jobs.front());
            jobs.pop_front();
            slot = &slots[job.seq % slots.size()];
            slot->seq = job.seq;
            slot->decoded = false;
        }
        int H = 0, W = 0;
        int st = tdm::read_png_u16(job.path.c_str(), pix, H, W);
        std::string msg;
        if (st != TDM_OK) msg = tdm_last_error();
        if (st == TDM_OK && (H != this->H || W != this->W)) {
            st = TDM_ERR_FORMAT;
            msg = job.path + ": unexpected image size";
        }
        if (st == TDM_OK) {
            std::memcpy(slot->h_mm, pix.data(), pix.size() * sizeof(uint16_t));
            std::memcpy(slot->h_pose, job.pose, sizeof job.pose);
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            slot->status = st;
            slot->decoded = true;
            if (st != TDM_OK) set_error(st, msg);
        }
        cv.notify_all();
    }
}

void tdm_fusion::issue_loop() {
    (void)hipSetDevice(device);
    const size_t px = (size_t)H * W;
    for (;;) {
        Slot *slot = nullptr;
        long seq = 0;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] {
                if (issued < submitted) {
                    const Slot &s = slots[issued % slots.size()];
                    return s.seq == issued && s.decoded;
                }
                return stopping;
            });
            if (issued >= submitted) return;
            seq = issued;
            slot = &slots[seq % slots.size()];
        }
        int st = slot->status;
        if (st == TDM_OK) {
            hipError_t e = hipMemcpyAsync(slot->d_mm, slot->h_mm, px * sizeof(uint16_t),
                                          hipMemcpyHostToDevice, stream);
            if (e == hipSuccess)
                e = hipMemcpyAsync(slot->d_pose, slot->h_pose, sizeof slot->h_pose,
                                   hipMemcpyHostToDevice, stream);
            if (e != hipSuccess) st = TDM_ERR_HIP;
            if (st == TDM_OK) st = tdm_depth_from_mm(nullptr, slot->d_mm, px, slot->d_depth, stream);
            if (st == TDM_OK)
                st = tdm_integrate(nullptr, d_K, slot->d_pose, slot->d_depth, H, W, dims[0],
                                   dims[1], dims[2], origin[0], origin[1], origin[2], vs, trunc,
                                   tsdf, weight, stream);
            if (st == TDM_OK && hipEventRecord(slot->done, stream) != hipSuccess) st = TDM_ERR_HIP;
            // the slot's pinned buffer is reused only after its copy has run
            if (st == TDM_OK && hipEventSynchronize(slot->done) != hipSuccess) st = TDM_ERR_HIP;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            if (st != TDM_OK) set_error(st, "frame " + std::to_string(seq) + " failed");
            slot->seq = -1;
            slot->decoded = false;
            ++issued;
        }
        cv.notify_all();
    }
}

namespace {

// Device and pinned resources of a new pipeline (no threads yet).
int fusion_setup(tdm_ctx *ctx, tdm_fusion *f, const float cam_K[9], int im_height, int im_width,
                 const int dims[3], const float origin[3], float voxel_size, float trunc_margin,
                 int threads) {
    f->device = ctx->device;
    f->H = im_height;
    f->W = im_width;
    for (int a = 0; a < 3; ++a) {
        f->dims[a] = dims[a];
        f->origin[a] = origin[a];
    }
    f->vs = voxel_size;
    f->trunc = trunc_margin;
    const size_t nvox = (size_t)dims[0] * dims[1] * dims[2];
    const size_t px = (size_t)im_height * im_width;
    TDM_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
    TDM_HIP(hipMalloc(&f->tsdf, nvox * sizeof(float)));
    TDM_HIP(hipMalloc(&f->weight, nvox * sizeof(float)));
    TDM_HIP(hipMalloc(&f->d_K, 9 * sizeof(float)));
    TDM_HIP(hipMemcpy(f->d_K, cam_K, 9 * sizeof(float), hipMemcpyHostToDevice));
    TDM_TRY(tdm_tsdf_init(ctx, f->tsdf, f->weight, nvox, f->stream));
    const int nthreads = threads > 0 ? threads : 4;
    f->slots.resize(nthreads + 2);
    for (Slot &s : f->slots) {
        TDM_HIP(hipHostMalloc((void **)&s.h_mm, px * sizeof(uint16_t), hipHostMallocDefault));
        TDM_HIP(hipMalloc(&s.d_mm, px * sizeof(uint16_t)));
        TDM_HIP(hipMalloc(&s.d_depth, px * sizeof(float)));
        TDM_HIP(hipMalloc(&s.d_pose, 16 * sizeof(float)));
        TDM_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
    }
    TDM_HIP(hipStreamSynchronize(f->stream));
    return TDM_OK;
}

}  // namespace

extern "C" {

int tdm_fusion_create(tdm_ctx *ctx, const float cam_K[9], int im_height, int im_width,
                      const int dims[3], const float origin[3], float voxel_size,
                      float trunc_margin, int threads, tdm_fusion **out) {
    tdm::clear_error();
    TDM_CHECK(cam_K && dims && origin && out, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_CHECK(im_height > 0 && im_width > 0 && dims[0] > 0 && dims[1] > 0 && dims[2] > 0,
              TDM_ERR_INVALID_ARG, "bad sizes");
    TDM_TRY(tdm::resolve_ctx(ctx, &ctx));
    tdm_fusion *f = new tdm_fusion;
    // a failure part-way releases what was created (tdm_fusion_destroy frees
    // the non-null members and has no threads to join yet)
    const int st = fusion_setup(ctx, f, cam_K, im_height, im_width, dims, origin, voxel_size,
                                trunc_margin, threads);
    if (st != TDM_OK) {
        const std::string msg = tdm_last_error();
        tdm_fusion_destroy(f);
        return tdm::fail(st, "%s", msg.c_str());
    }
    tdm_fusion *raw = f;
    const int nthreads = (int)raw->slots.size() - 2;
    for (int t = 0; t < nthreads; ++t) raw->decoders.emplace_back([raw] { raw->decode_loop(); });
    raw->issuer = std::thread([raw] { raw->issue_loop(); });
    *out = raw;
    return TDM_OK;
}

int tdm_fusion_add_png(tdm_fusion *f, const char *depth_png, const float cam2base[16]) {
    tdm::clear_error();
    TDM_CHECK(f && depth_png && cam2base, TDM_ERR_INVALID_ARG, "null pointer argument");
    {
        std::lock_guard<std::mutex> lk(f->mu);
        TDM_CHECK(!f->stopping, TDM_ERR_INVALID_ARG, "fusion already finished");
        Job j;
        j.seq = f->submitted++;
        j.path = depth_png;
        std::memcpy(j.pose, cam2base, sizeof j.pose);
        f->jobs.push_back(std::move(j));
    }
    f->cv.notify_all();
    return TDM_OK;
}

int tdm_fusion_wait(tdm_fusion *f) {
    tdm::clear_error();
    TDM_CHECK(f, TDM_ERR_INVALID_ARG, "NULL fusion");
    {
        std::unique_lock<std::mutex> lk(f->mu);
        f->cv.wait(lk, [&] { return f->issued >= f->submitted; });
    }
    TDM_HIP(hipSetDevice(f->device));
    TDM_HIP(hipStreamSynchronize(f->stream));
    std::lock_guard<std::mutex> lk(f->mu);
    if (f->error != TDM_OK) return tdm::fail(f->error, "%s", f->error_msg.c_str());
    return TDM_OK;
}

int tdm_fusion_grids(tdm_fusion *f, float **tsdf, float **weight) {
    tdm::clear_error();
    TDM_CHECK(f && tsdf && weight, TDM_ERR_INVALID_ARG, "null pointer argument");
    *tsdf = f->tsdf;
    *weight = f->weight;
    return TDM_OK;
}

int tdm_fusion_copy_grids(tdm_fusion *f, float *tsdf, float *weight, void *stream) {
    tdm::clear_error();
    TDM_CHECK(f && tsdf && weight, TDM_ERR_INVALID_ARG, "null pointer argument");
    TDM_TRY(tdm_fusion_wait(f));
    const size_t bytes = sizeof(float) * f->dims[0] * f->dims[1] * f->dims[2];
    hipStream_t s = tdm::as_stream(stream);
    TDM_HIP(hipMemcpyAsync(tsdf, f->tsdf, bytes, hipMemcpyDeviceToDevice, s));
    TDM_HIP(hipMemcpyAsync(weight, f->weight, bytes, hipMemcpyDeviceToDevice, s));
    return TDM_OK;
}

int tdm_fusion_destroy(tdm_fusion *f) {
    tdm::clear_error();
    if (!f) return TDM_OK;
    {
        std::lock_guard<std::mutex> lk(f->mu);
        f->stopping = true;
    }
    f->cv.notify_all();
    for (auto &t : f->decoders) t.join();
    if (f->issuer.joinable()) f->issuer.join();
    (void)hipSetDevice(f->device);
    (void)hipStreamSynchronize(f->stream);
    for (Slot &s : f->slots) {
        if (s.h_mm) (void)hipHostFree(s.h_mm);
        if (s.d_mm) (void)hipFree(s.d_mm);
        if (s.d_depth) (void)hipFree(s.d_depth);
        if (s.d_pose) (void)hipFree(s.d_pose);
        if (s.done) (void)hipEventDestroy(s.done);
    }
    if (f->tsdf) (void)hipFree(f->tsdf);
    if (f->weight) (void)hipFree(f->weight);
    if (f->d_K) (void)hipFree(f->d_K);
    if (f->stream) (void)hipStreamDestroy(f->stream);
    delete f;
    return TDM_OK;
}

}  // extern "C"
