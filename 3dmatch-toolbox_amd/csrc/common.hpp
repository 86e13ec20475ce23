// Shared host-side plumbing of lib3dmatch_amd: status/error reporting,
// HIP error checks, and the per-device context with its scratch arena.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>

#include "tdm.h"

namespace tdm {

// Records a message for tdm_last_error() on this thread and returns `status`.
int fail(int status, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

// A failed HIP call also leaves its code in HIP's per-thread last error,
// which the next launch check would read; it is cleared here.
#define TDM_HIP(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (void)hipGetLastError();                                                       \
            return ::tdm::fail(TDM_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                \
                               hipGetErrorString(e_), __FILE__, __LINE__);                 \
        }                                                                                  \
    } while (0)

#define TDM_CHECK(cond, status, ...)                                                       \
    do {                                                                                   \
        if (!(cond))                                                                       \
            return ::tdm::fail((status), __VA_ARGS__);                                     \
    } while (0)

#define TDM_TRY(expr)                                                                      \
    do {                                                                                   \
        int s_ = (expr);                                                                   \
        if (s_ != TDM_OK)                                                                  \
            return s_;                                                                     \
    } while (0)

// After a kernel launch: report launch errors.
#define TDM_LAUNCHED() TDM_HIP(hipGetLastError())

// A growable device buffer owned by a context.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace tdm

// Per-device state: the device id and scratch buffers that kernels need
// beyond their inputs and outputs (bit grids, reduction partials, staging).
// Scratch is reused across calls on the same context, so calls that overlap
// on one context must be ordered on one stream.
struct tdm_ctx {
    int device = 0;
    std::mutex mu;
    tdm::DevBuf scratch[4];
    bool owned_by_library = false;
};

namespace tdm {

// Resolves ctx (NULL -> library context of the current device), makes its
// device current, and verifies a device is present.
int resolve_ctx(tdm_ctx *in, tdm_ctx **out);
// Scratch slot `slot` of at least `bytes` bytes (grown with hipMalloc; the
// stream is synchronized before an existing buffer is replaced).
int ctx_scratch(tdm_ctx *ctx, int slot, size_t bytes, void **ptr, hipStream_t stream);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned blocks_for(size_t n, unsigned threads) {
    return static_cast<unsigned>((n + threads - 1) / threads);
}

}  // namespace tdm
