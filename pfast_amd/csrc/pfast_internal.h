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

// Raw buffer access: 32-bit byte offsets from a uniform base, and an offset at or
// past num_records (a position outside the array, e.g. where an unsigned offset
// wraps below the base) reads 0 without touching memory and never faults.
using rsrc_t = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ rsrc_t make_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes,
                                             0x00020000);
}
// A pointer every lane holds the same value of, as a scalar (a buffer descriptor must
// be; the compiler cannot always tell, e.g. for a pointer read from memory that the
// kernel also writes, and would otherwise loop over the lanes for every access).
template <class T>
__device__ __forceinline__ T *uniform_ptr(T *p)
{
    const unsigned long long v = reinterpret_cast<unsigned long long>(p);
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return reinterpret_cast<T *>(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ float bload(rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ float2 bload2(rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ double bload_d(rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ unsigned bload_u8(rsrc_t r, unsigned off)
{
    return __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}
__device__ __forceinline__ void bstore(rsrc_t r, unsigned off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, 0);
}

// Device-scope (sc1) buffer access for data that other work-groups of the same
// launch write or read: on gfx950 each XCD has its own L2, and an access at device
// scope is coherent across them (the same encoding the compiler gives a relaxed
// agent-scope atomic).  The dataflow sweep kernels access the fields this way; the
// per-diagonal ones, which read only what earlier launches wrote, keep the default
// policy (device scope there measured 23% slower in the forward, 3% in the adjoint).
constexpr int kCpolDevice = 16;  // CPol::SC1
__device__ __forceinline__ float bload_dev(rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, kCpolDevice));
}
__device__ __forceinline__ float2 bload2_dev(rsrc_t r, unsigned off)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, kCpolDevice));
}
__device__ __forceinline__ void bstore_dev(rsrc_t r, unsigned off, float v)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, off, 0, kCpolDevice);
}
template <class T>
__device__ __forceinline__ T ld_agent(const T *p)
{
    return __hip_atomic_load(const_cast<T *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ __forceinline__ void st_agent(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// vmcnt(0) (expcnt and lgkmcnt left at their maxima).  Issued through the builtin
// so that the compiler's own wait insertion knows nothing is outstanding after it.
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// LDS-only workgroup barrier: waits for this wave's LDS traffic, then s_barrier.
// Outstanding global loads/stores stay in flight (no vmcnt wait), which is what
// lets the batch prefetch overlap the sweep steps.
__device__ __forceinline__ void lds_barrier()
{
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}
