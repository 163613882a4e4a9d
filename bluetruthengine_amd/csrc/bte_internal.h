// Host-side helpers shared by the renderer's C ABI (bte_api.cpp) and the
// denoiser's (bte_denoise_api.cpp): error reporting, owning device buffers and
// a view of the current render target.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/bte_api.h"

namespace bte {

// Records the message returned by bte_last_error() and returns code.
int fail(int code, const char* fmt, ...);

#define HIPCHK(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return ::bte::fail(BTE_ERR_DEVICE, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                               __FILE__, __LINE__);                                           \
    } while (0)
#define CHECK_RC(expr)               \
    do {                             \
        int rc_ = (expr);            \
        if (rc_) return rc_;         \
    } while (0)

// Owning device allocation.
struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    DBuf(DBuf&& o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DBuf& operator=(DBuf&& o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    int ensure(size_t n) {
        if (n <= bytes && p) return BTE_OK;
        release();
        if (n == 0) return BTE_OK;
        if (hipMalloc(&p, n) != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            return fail(BTE_ERR_MEMORY, "hipMalloc(%zu) failed", n);
        }
        bytes = n;
        return BTE_OK;
    }
    int upload(const void* src, size_t n) {
        int rc = ensure(n);
        if (rc) return rc;
        if (n) HIPCHK(hipMemcpy(p, src, n, hipMemcpyHostToDevice));
        return BTE_OK;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// The renderer's current accumulation target.
// A render target of the batch API or the single-scene one.
struct RenderView {
    hipStream_t stream = nullptr;
    int device = 0;
    int width = 0, height = 0;
    const float4* accum = nullptr;     // running mean, row 0 = bottom
    const float4* aovAlb = nullptr;    // AOV means (null when AOVs are off)
    const float4* aovNrm = nullptr;
    uint32_t* fb = nullptr;            // display buffer
    uint32_t frameIndex = 0;
};
int ensure_device();
hipStream_t main_stream();             // the library's stream (after ensure_device)
int render_view(RenderView& v);        // BTE_ERR_STATE when nothing was accumulated

// Denoiser hooks called by the render entry points.
bool denoiser_active();
int denoiser_present();                // denoise the accumulation into the display buffer
// Denoise one view into `out` (float4 per pixel, linear) and the view's display
// buffer; asynchronous on v.stream.
// captured: replay the launches as a HIP graph (cached by the buffers and
// sizes they use) when graphs are on; single frames only, a batch's many
// views would churn the graph cache.
int denoise_view(const RenderView& v, float4* out, bool captured = false);
// Runs emit's launches on the library's stream, captured once per key into a
// HIP graph and replayed (plain launches when graphs are off).
int run_captured(const std::vector<unsigned char>& key, const std::function<void()>& emit);
void denoiser_release();               // bte_cleanup

}  // namespace bte
