// kernels.hpp — host-side launch interface of the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/alien_hip.h"
#include "engine_types.hpp"

namespace alien {

// Launch shape of one world set: sizes of the per-world block tables and grid sizes of the
// candidate-driven kernels.  None of it changes results: no arithmetic depends on the launch shape.
struct LaunchShape {
    int cellBlocks = 0, cellThreads = 256;
    int partBlocks = 0, partThreads = 256;
    int binBlocks = 0, pbinBlocks = 0, binThreads = 256;
    int tileBlocks = 0;    // neighbour-pass workgroups (including XCD padding)
    int partGrid = 64;     // grid-stride particle kernels
    int bondBlocks = 0;    // bond kernels: cell blocks of the worlds that can have bonded cells
    int candGrid = 16;     // grid-stride kernels over fusion candidates
    bool hasFluid = false, hasCollision = false;  // motion modes present in the set
    bool hasBondWork = true;  // some active world has (or can get by fusion) a bonded or barrier cell
    // some world has live energy particles or radiates (only cell radiation creates particles, so
    // once all are absorbed or merged in a set without it the particle kernels, which would find
    // no work, are not launched)
    bool hasParticles = true;
    bool hasRadiation = false;  // some active world radiates (DevParams::radiates)
    bool hasSplitting = false;  // some active world's particles can split (DevParams::splits)
    int livingBlocks = 0;           // cell blocks of the worlds whose living states live on the device
    bool hasInnerFriction = false;  // some active world with bond work has innerFriction > 0
};

// Optional per-launch hook: mark(name) is called right before each kernel launch and
// mark(nullptr) after the last one, so a recorder can bracket every kernel with events.
struct Recorder {
    virtual ~Recorder() = default;
    virtual void mark(const char* name) = 0;
};

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
// bin tiles of the one-pass map scan (one status word each)
int map_scan_tiles(int numBins);
// bins of the current state (A) for the next sort; the step keeps them up to date itself
void launch_seed_bins(const Ctx& c, const LaunchShape& sh, hipStream_t s);
// workgroups of the row-streaming (collision) neighbour pass one CU keeps resident
int neighbor_blocks_per_cu();
// A second stream of the same priority and two events: the particle kernels run on it beside the
// neighbour pass (sort, movement, targets and fusion claims always; fusions and absorptions too
// when no world has bond work, so that nothing in the pass reads the energies they change),
// joined before the kernels after the pass.
struct SideStream {
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, mapped = nullptr, join = nullptr;  // after the scan, the scatter, the particles
};
// rec != nullptr (per-kernel timing) launches everything on s, one kernel after another
void launch_step(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec = nullptr,
                 const SideStream* side = nullptr);
// cell map of the current state in logical order (sorted/newidx), for introspection; consumes
// the seeded bins and seeds them again afterwards
void launch_debug_map(const Ctx& c, const LaunchShape& sh, hipStream_t s);
// per-colour statistics of one world: block count for its cells and particle slots, and the
// launch writing per-block counts and energies (kColors per block)
int statistics_blocks(int numCells, int partCap, int& cellBlocks);
void launch_statistics(const Ctx& c, int wi, int cellBlocks, int blocks, int* cnt, double* en, hipStream_t s);
// the same without allocating: partials and result in the library's per-device `scratch` array
// (at most kStatScratchBlocks blocks); statistics_scratch gives the current device's address of
// it (looked up once per device), launch_statistics_noalloc writes the result into *hostOut and
// returns when it is there (timestep left 0)
constexpr int kStatScratchBlocks = 65536;
hipError_t statistics_scratch(void** ptr);
hipError_t launch_statistics_noalloc(const Ctx& c, int wi, int cellBlocks, int blocks, void* scratchPtr,
                                     alien_statistics* hostOut, hipStream_t s);
void launch_debug_neighbors(const Ctx& c, const LaunchShape& sh, const uint64_t* offsets, int* out, int* counts,
                            hipStream_t s);

}  // namespace alien
