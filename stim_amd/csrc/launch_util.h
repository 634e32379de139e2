// Host-side launch helper shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <mutex>

namespace sa {

// Raises a kernel's dynamic-LDS limit to at least lds bytes and never lowers it: the check and the
// update happen under one lock, so launches of different sizes from several threads cannot lower
// the limit below another thread's pending launch. `reserved` is the kernel's current limit (one
// variable per kernel).
inline hipError_t reserve_lds(const void *kernel, size_t lds, size_t &reserved) {
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    if (lds <= reserved) return hipSuccess;
    const hipError_t e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) reserved = lds;
    return e;
}

}  // namespace sa
