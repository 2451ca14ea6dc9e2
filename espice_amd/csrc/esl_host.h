// Host-side helpers shared by the lane engine and the warm-start network:
// device buffers, HIP error capture and generation-checked handles.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/espice_lanes.h"

namespace eslh {

extern thread_local std::string g_last_error;
esl_status fail_hip(hipError_t e, const char *what);

#define HIPCHK(x)                                             \
    do {                                                      \
        hipError_t e_ = (x);                                  \
        if (e_ != hipSuccess) return eslh::fail_hip(e_, #x);  \
    } while (0)

// One device allocation that grows on demand.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    hipError_t ensure(size_t b)
    {
        if (b <= bytes && p) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&p, b ? b : 8);
        if (e == hipSuccess) bytes = b;
        return e;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Handles: generation-checked tokens (never addresses); type 1 engine,
// 2 tape, 3 network.
void *make_handle(void *obj, int type);
void *lookup(void *h, int type);
void *release(void *h, int type);
bool device_present();

// The body of esl_tape_replay_noalloc: esl_tape_replay with its device
// buffers carved from `stage` (stage_bytes of device memory); no allocation.
esl_status tape_replay_staged(esl_tape *h, const double *a, const double *rhs, size_t lanes, double growth,
                              double *lx, double *ux, double *udiag, double *dx, uint32_t *fail, void *stage,
                              size_t stage_bytes);

}  // namespace eslh
