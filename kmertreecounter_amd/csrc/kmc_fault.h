// kmc_fault.h -- the engine's device and pinned-host allocations and its
// copies, through one place each, so that a test hook (kc_debug_fault, or the
// KMC_FAULT_ALLOC / KMC_FAULT_COPY environment variables) can make the n-th one
// fail and every error path can be exercised on the host side.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace kmc {

enum : uint32_t { kFaultAlloc = 1, kFaultCopy = 2 };

// Counts a call of `kind`; true when the armed hook fails this one.
bool fault_hit(uint32_t kind);

// A failed allocation is returned to the caller, which handles it; the
// runtime's last-error state is cleared so that the next launch check
// (hipGetLastError after a kernel launch) does not report it again.
inline hipError_t dmalloc(void** p, size_t bytes)
{
    if (fault_hit(kFaultAlloc)) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
}
template <class T>
inline hipError_t dmalloc(T** p, size_t bytes)
{
    return dmalloc(reinterpret_cast<void**>(p), bytes);
}

inline hipError_t dhost_malloc(void** p, size_t bytes, unsigned flags)
{
    if (fault_hit(kFaultAlloc)) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    const hipError_t e = hipHostMalloc(p, bytes, flags);
    if (e != hipSuccess) (void)hipGetLastError();
    return e;
}
template <class T>
inline hipError_t dhost_malloc(T** p, size_t bytes, unsigned flags)
{
    return dhost_malloc(reinterpret_cast<void**>(p), bytes, flags);
}

inline hipError_t dcopy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
{
    if (fault_hit(kFaultCopy)) return hipErrorInvalidValue;
    return hipMemcpy(dst, src, bytes, kind);
}

inline hipError_t dcopy_async(void* dst, const void* src, size_t bytes, hipMemcpyKind kind, hipStream_t st)
{
    if (fault_hit(kFaultCopy)) return hipErrorInvalidValue;
    return hipMemcpyAsync(dst, src, bytes, kind, st);
}

// frees a device pointer and clears it
template <class T>
inline void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace kmc
