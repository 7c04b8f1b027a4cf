// lus_launch.hpp -- the one launch path of the library's kernels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "lus_internal.hpp"

namespace lus {

// hipLaunchKernelGGL; while kernel_times().on, hipExtLaunchKernelGGL with an
// event pair that the runtime stamps at this dispatch's own start and end (the
// kernel's execution time, as a kernel trace reports it, without the gaps
// between dispatches that events recorded between launches include).
template <class F, class... Args>
inline void launch_kernel(F kernel, dim3 grid, dim3 block, unsigned shmem, hipStream_t s, Args... args) {
    KernelTimes& kt = kernel_times();
    if (kt.on) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) == hipSuccess) {
            if (hipEventCreate(&e1) == hipSuccess) {
                kt.start.push_back(e0);
                kt.stop.push_back(e1);
                hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, e0, e1, 0u, args...);
                return;
            }
            (void)hipEventDestroy(e0);
        }
        kt.failed = true;
    }
    hipLaunchKernelGGL(kernel, grid, block, shmem, s, args...);
}

}  // namespace lus
