// kernels.hpp — host-side launch interface of kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "engine_types.hpp"

namespace alien {

// Launch shape of one world set: number of blocks of the per-world block tables and their
// threads per block.  Changing it never changes results (no launch-shape-dependent arithmetic).
struct LaunchShape {
    int cellBlocks = 0, cellThreads = 256;
    int partBlocks = 0, partThreads = 256;
    int binBlocks = 0, pbinBlocks = 0, binThreads = 256;
};

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
void launch_reset_density(const Ctx& c, const LaunchShape& sh, hipStream_t s);
void launch_sort(const Ctx& c, const LaunchShape& sh, hipStream_t s);
void launch_step(const Ctx& c, const LaunchShape& sh, hipStream_t s);
void launch_debug_neighbors(const Ctx& c, const LaunchShape& sh, const uint64_t* offsets, int* out, int* counts,
                            hipStream_t s);

}  // namespace alien
