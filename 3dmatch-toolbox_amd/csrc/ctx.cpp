// Library plumbing: error strings, contexts and their scratch arenas.
#include <cstdio>
#include <cstdlib>
#include <map>

#include "common.hpp"

namespace tdm {

static thread_local std::string g_last_error;

int fail(int status, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return status;
}

void clear_error() { g_last_error.clear(); }

static std::mutex g_ctx_mu;
static std::map<int, tdm_ctx *> g_default_ctx;

int resolve_ctx(tdm_ctx *in, tdm_ctx **out) {
    // an error a caller's earlier HIP call left behind (and already reported
    // to that caller) must not be taken for a failure of this entry point's
    // launches
    (void)hipGetLastError();
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(TDM_ERR_NO_DEVICE, "no HIP device available");
    }
    if (in) {
        TDM_CHECK(in->device >= 0 && in->device < count, TDM_ERR_INVALID_ARG,
                  "context device %d out of range", in->device);
        TDM_HIP(hipSetDevice(in->device));
        *out = in;
        return TDM_OK;
    }
    int dev = 0;
    TDM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_ctx_mu);
    auto it = g_default_ctx.find(dev);
    if (it == g_default_ctx.end()) {
        tdm_ctx *c = new tdm_ctx();
        c->device = dev;
        c->owned_by_library = true;
        it = g_default_ctx.emplace(dev, c).first;
    }
    *out = it->second;
    return TDM_OK;
}

int ctx_scratch(tdm_ctx *ctx, int slot, size_t bytes, void **ptr, hipStream_t stream) {
    TDM_CHECK(slot >= 0 && slot < 4, TDM_ERR_INVALID_ARG, "bad scratch slot %d", slot);
    std::lock_guard<std::mutex> lock(ctx->mu);
    DevBuf &b = ctx->scratch[slot];
    if (b.bytes < bytes) {
        if (b.ptr) {
            // Work already queued may still use the old buffer.
            TDM_HIP(hipStreamSynchronize(stream));
            TDM_HIP(hipDeviceSynchronize());
            TDM_HIP(hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t want = bytes + bytes / 8 + 4096;
        hipError_t e = hipMalloc(&b.ptr, want);
        if (e != hipSuccess) {
            b.ptr = nullptr;
            return fail(TDM_ERR_NOMEM, "scratch allocation of %zu bytes failed: %s", want,
                        hipGetErrorString(e));
        }
        b.bytes = want;
    }
    *ptr = b.ptr;
    return TDM_OK;
}

}  // namespace tdm

extern "C" {

int tdm_version(void) { return TDM_VERSION; }

const char *tdm_status_string(int status) {
    switch (status) {
    case TDM_OK: return "ok";
    case TDM_ERR_INVALID_ARG: return "invalid argument";
    case TDM_ERR_IO: return "i/o error";
    case TDM_ERR_FORMAT: return "bad file format";
    case TDM_ERR_HIP: return "HIP runtime error";
    case TDM_ERR_NOMEM: return "out of memory";
    case TDM_ERR_UNSUPPORTED: return "unsupported";
    case TDM_ERR_NO_DEVICE: return "no GPU device";
    default: return "unknown status";
    }
}

const char *tdm_last_error(void) { return tdm::g_last_error.c_str(); }

void tdm_free(void *p) { std::free(p); }

int tdm_ctx_create(int device, tdm_ctx **out) {
    TDM_CHECK(out, TDM_ERR_INVALID_ARG, "out is NULL");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return tdm::fail(TDM_ERR_NO_DEVICE, "no HIP device available");
    }
    TDM_CHECK(device >= 0 && device < count, TDM_ERR_INVALID_ARG, "device %d out of range", device);
    tdm_ctx *c = new tdm_ctx();
    c->device = device;
    *out = c;
    return TDM_OK;
}

int tdm_ctx_destroy(tdm_ctx *ctx) {
    if (!ctx) return TDM_OK;
    TDM_CHECK(!ctx->owned_by_library, TDM_ERR_INVALID_ARG, "library context cannot be destroyed");
    int prev = 0;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto &b : ctx->scratch)
        if (b.ptr) (void)hipFree(b.ptr);
    (void)hipSetDevice(prev);
    delete ctx;
    return TDM_OK;
}

}  // extern "C"
