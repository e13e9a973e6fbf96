// Device helpers shared by the kernel files.
#pragma once
#include <hip/hip_runtime.h>

namespace pth {

// Workgroups are dispatched to the 8 XCDs of the MI355X round-robin by linear id, and each XCD
// has its own L2. Remapping the linear id w of W workgroups this way gives every XCD one
// contiguous range of logical workgroups, so neighbouring tiles that touch the same cache lines
// share one L2 instead of fetching them into several.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned XcdRemap(unsigned w, unsigned W) {
  const unsigned x = w % kXcds, i = w / kXcds, per = W / kXcds, rem = W % kXcds;
  return x * per + (x < rem ? x : rem) + i;
}

// Global-address-space view of a device pointer. Loads and stores through it are global_load
// / global_store, counted on vmcnt alone; through a generic pointer read from memory (a table
// entry, a struct of arrays) the compiler emits flat accesses, which also count on lgkmcnt.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *Glob(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

}  // namespace pth
