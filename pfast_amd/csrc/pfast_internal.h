// pfast_internal.h -- shared internals of libpfast (HIP/gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#define PFAST_INF 9999999.0

namespace pfast {

// Thread-local last error string, surfaced by pfast_last_error().
void set_error(const std::string &msg);

struct Error {
    int code;
    std::string msg;
};

}  // namespace pfast

// HIP call check inside functions returning int status.
#define HIPCHK(call)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            pfast::set_error(std::string(#call) + ": " + hipGetErrorString(e_));          \
            return PFAST_ERR_HIP;                                                          \
        }                                                                                  \
    } while (0)

#define ARGCHK(cond, msg)                                                                  \
    do {                                                                                   \
        if (!(cond)) {                                                                     \
            pfast::set_error(std::string("invalid argument: ") + (msg));                  \
            return PFAST_ERR_ARG;                                                          \
        }                                                                                  \
    } while (0)

// LDS-only workgroup barrier: waits for this wave's LDS traffic, then s_barrier.
// Outstanding global loads/stores stay in flight (no vmcnt wait), which is what
// lets the batch prefetch overlap the sweep steps.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}
