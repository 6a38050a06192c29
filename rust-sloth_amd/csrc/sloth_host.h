// Host-side helpers shared by the C-ABI and the tiled planner (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <stddef.h>

#include "sloth_hip.h"

namespace sloth {

// Records a message for sloth_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);

// Looks up the current device's address of the no-allocation entry point's
// scratch array (once per device).
int noalloc_init(int dev);

// Device allocation owned by a handle; grow-only, contents not preserved.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int reserve(size_t n)
    {
        if (n <= bytes && p) return SLOTH_OK;
        release();
        if (n == 0) return SLOTH_OK;
        hipError_t e = hipMalloc(&p, n);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(SLOTH_ENOMEM, "hipMalloc(%zu bytes): %s", n, hipGetErrorString(e));
        }
        bytes = n;
        return SLOTH_OK;
    }
    template <class T> T* as() const { return (T*)p; }
};

}  // namespace sloth

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return sloth::fail(SLOTH_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

#define ARG_CHECK(cond, msg)                                      \
    do {                                                          \
        if (!(cond)) return sloth::fail(SLOTH_EINVAL, "%s", msg); \
    } while (0)
