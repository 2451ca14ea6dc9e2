// esl_tape_replay_noalloc: the LU replay of esl_tape_replay for callers that
// cannot allocate device memory (tight serving loops, harnesses that forbid
// hipMalloc). Its staging buffers (the matrices and right-hand sides copied
// in, the factors, solutions and peel flags copied out) live in one fixed
// device array, `scratch`, that every GPU the library runs on holds until the
// process ends. Calls that overlap on one GPU share that array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "esl_host.h"

namespace {

constexpr size_t kScratchBytes = size_t(32) << 20;   // 32 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of `scratch` on each GPU, looked up once per device
// with that device current. Host code uses only these addresses: the array's
// own name is a host-side stand-in, not a device address.
void *g_scratch_ptr[kMaxDevices];
std::once_flag g_scratch_once[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

hipError_t scratch_on(int dev, void **out)
{
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::call_once(g_scratch_once[dev], [dev] {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_err[dev] = e;
        if (e == hipSuccess) g_scratch_ptr[dev] = p;
    });
    *out = g_scratch_ptr[dev];
    return g_scratch_err[dev];
}

}  // namespace

extern "C" esl_status esl_tape_replay_noalloc(esl_tape *t, const double *a, const double *rhs, size_t lanes,
                                              double growth, double *lx, double *ux, double *udiag, double *dx,
                                              uint32_t *fail)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    void *stage = nullptr;
    HIPCHK(scratch_on(dev, &stage));
    return eslh::tape_replay_staged(t, a, rhs, lanes, growth, lx, ux, udiag, dx, fail, stage, kScratchBytes);
}
