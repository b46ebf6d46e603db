// launch.hip — the kernel sequence of one timestep and the introspection passes.
//
// Step order follows _SimulationKernelsLauncher::calcTimestep
// (source/EngineGpuKernels/SimulationKernelsLauncher.cu:25-91): maps, neighbour forces, particle
// movement/collision, connection forces around the Verlet updates, friction, structural operations.
#include <mutex>

#include "../../include/alien_hip.h"
#include "kernels_common.hpp"
#include "kernels.hpp"

namespace alien {

void launch_map_scan(const Ctx& c, hipStream_t s, StepCounters* open, bool advanceTimestep);
void launch_cell_scatter(const Ctx& c, const LaunchShape& sh, hipStream_t s);
void launch_part_sort(const Ctx& c, const LaunchShape& sh, hipStream_t s);
void launch_part_scatter(const Ctx& c, const LaunchShape& sh, hipStream_t s);
void launch_cell_binsort(const Ctx& c, const LaunchShape& sh, hipStream_t s);
void launch_neighbor(const Ctx& c, const LaunchShape& sh, hipStream_t s, bool fluid, bool collision);
void launch_particles_pre(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec);
void launch_particles_post(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec);
void launch_bonds(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec);
void launch_radiation(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec);

void launch_step(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec, const SideStream* side) {
    // cell radiation comes first: the particles it emits enter this step's particle map
    launch_radiation(c, sh, s, rec);
    // without bond work the end-of-step kernel would only advance the timesteps: the map scan does
    // it
    if (rec) rec->mark("map_scan");
    launch_map_scan(c, s, c.sc, !sh.hasBondWork);
    // The particle kernels run beside the cell scatter and the neighbour pass: their sort and
    // movement need only the scan (the particle bins' offsets), their targets and fusion claims
    // also the cell map.
    const bool fork = side && !rec && sh.hasParticles && sh.partBlocks;
    // Absorption adds to cell energies, which the neighbour pass reads for fusion and the bond
    // kernels for bond decay: without bond work (no bonded, barrier or bondable cell) neither
    // reads them, and the fusions and absorptions run beside the neighbour pass as well.
    const bool allBeside = fork && !sh.hasBondWork;
    if (fork) {
        (void)hipEventRecord(side->fork, s);
        if (rec) rec->mark("cell_scatter");
        launch_cell_scatter(c, sh, s);
        (void)hipEventRecord(side->mapped, s);
        launch_neighbor(c, sh, s, sh.hasFluid, sh.hasCollision);
        (void)hipStreamWaitEvent(side->stream, side->fork, 0);
        launch_part_sort(c, sh, side->stream);
        (void)hipStreamWaitEvent(side->stream, side->mapped, 0);
        launch_particles_pre(c, sh, side->stream, rec);
        if (allBeside) launch_particles_post(c, sh, side->stream, rec);
        (void)hipEventRecord(side->join, side->stream);
        (void)hipStreamWaitEvent(s, side->join, 0);
    } else {
        if (rec) rec->mark("cell_scatter");
        launch_cell_scatter(c, sh, s);
        if (rec && sh.hasParticles) rec->mark("part_sort");
        launch_part_sort(c, sh, s);
        launch_particles_pre(c, sh, s, rec);
        if (rec) rec->mark("neighbor");
        launch_neighbor(c, sh, s, sh.hasFluid, sh.hasCollision);
    }
    if (!allBeside) launch_particles_post(c, sh, s, rec);
    launch_bonds(c, sh, s, rec);
    if (rec) rec->mark(nullptr);
}

// The map of the current state: consumes the seeded bins and mover lists, which seeding restores
// for the next step.
void launch_debug_map(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    launch_map_scan(c, s, nullptr, false);
    launch_cell_scatter(c, sh, s);
    launch_part_scatter(c, sh, s);
    launch_cell_binsort(c, sh, s);
    launch_seed_bins(c, sh, s);
}

__global__ void k_debug_neighbors(Ctx c, const uint64_t* offsets, int* out, int* counts) {
    int2 blk = c.cellBlocks[blockIdx.x];
    const int wi = blk.x;
    const DevWorld w = c.worlds[wi];
    const int k = blk.y + int(threadIdx.x);
    if (k >= w.cellOff + w.numCells) return;
    const DevParams& P = c.params[wi];
    const int i = c.sorted[k];
    const f2 pi = to_f2(c.A.pos[i]);
    const int deti = flags_detached(c.A.flags[i]);
    int sx, sy;
    slot_of(pi, w.W, w.H, sx, sy);
    const int bx = min(sx / w.binSize, w.nbx - 1), by = min(sy / w.binSize, w.nby - 1);
    int cnt = 0;
    uint64_t o = offsets ? offsets[i] : 0;
    for (int dy = -1; dy <= 1; ++dy) {
        int y = by + dy;
        y = y < 0 ? y + w.nby : (y >= w.nby ? y - w.nby : y);
        for (int dx = -1; dx <= 1; ++dx) {
            int x = bx + dx;
            x = x < 0 ? x + w.nbx : (x >= w.nbx ? x - w.nbx : x);
            int s, e;
            cell_bin_range(c, w, y * w.nbx + x, s, e);
            for (int m = s; m < e; ++m) {
                const int j = c.sorted[m];
                if (j == i) continue;
                const f2 pj = to_f2(c.A.pos[j]);
                if (deti + flags_detached(c.A.flags[j]) == 1) continue;
                bool nb;
                if (P.motion == 0)
                    nb = len2(torus_delta(pi, pj, float(w.W), float(w.H))) <= P.smoothing * 2.0f;
                else
                    nb = len2(sub(pj, pi)) <= P.collisionDistance;
                if (!nb) continue;
                if (out) out[o + cnt] = j;
                ++cnt;
            }
        }
    }
    if (counts) counts[i] = cnt;
}

void launch_debug_neighbors(const Ctx& c, const LaunchShape& sh, const uint64_t* offsets, int* out, int* counts,
                            hipStream_t s) {
    if (sh.cellBlocks)
        hipLaunchKernelGGL(k_debug_neighbors, dim3(sh.cellBlocks), dim3(sh.cellThreads), 0, s, c, offsets, out, counts);
}

// Per-colour statistics of world wi (StatisticsKernels.cu:11-45): blocks [0, cellBlocks) take 256
// cells each, the rest 256 particle slots each (living particles only); every block writes its
// colours' counts and energies, summed in a fixed order (a tree per wave, then the waves in turn),
// and the host adds the blocks in order.
constexpr int kStatThreads = 256;
__global__ void __launch_bounds__(kStatThreads) k_statistics(Ctx c, int wi, int cellBlocks, int* cnt, double* en) {
    __shared__ int wc[kStatThreads / 64][kColors];
    __shared__ double we[kStatThreads / 64][kColors];
    const DevWorld w = c.worlds[wi];
    const bool cells = int(blockIdx.x) < cellBlocks;
    const int idx = (cells ? int(blockIdx.x) : int(blockIdx.x) - cellBlocks) * kStatThreads + int(threadIdx.x);
    int col = -1, col2 = -1;
    double e = 0.0, e2 = 0.0;
    if (cells) {
        if (idx < w.numCells) {
            const int i = w.cellOff + idx;
            col = flags_color(c.A.flags[i]) % kColors;
            e = double(c.energy[w.cellOff + c.A.slot[i]]);
        }
    } else if (idx < c.ctr[wi].numParticles) {
        // (particles absorbed in the last step stay in the live range, marked, until the next sort)
        const int j = w.partOff + idx;
        const uint8_t pf = c.A.pflags[j];
        if (!(pf & kParticleDead)) {
            col = int(pf & kParticleColorMask) % kColors;
            e = double(c.A.pen[j]);
        }
        // a half split off it at the end of the last step, waiting for the next map
        if (c.splitStage > 0) {
            const StagedParticle r = c.splitStaged[j];
            if (r.color & kStagedPending) {
                col2 = int(r.color & kParticleColorMask) % kColors;
                e2 = double(r.energy);
            }
        }
    }
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
#pragma unroll
    for (int k = 0; k < kColors; ++k) {
        int n = (col == k ? 1 : 0) + (col2 == k ? 1 : 0);
        double s = (col == k ? e : 0.0) + (col2 == k ? e2 : 0.0);
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            n += __shfl_down(n, o, 64);
            s += __shfl_down(s, o, 64);
        }
        if (lane == 0) {
            wc[wave][k] = n;
            we[wave][k] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x < kColors) {
        const int k = int(threadIdx.x);
        int n = 0;
        double s = 0.0;
        for (int v = 0; v < kStatThreads / 64; ++v) {
            n += wc[v][k];
            s += we[v][k];
        }
        cnt[blockIdx.x * kColors + k] = n;
        en[blockIdx.x * kColors + k] = s;
    }
}

int statistics_blocks(int numCells, int partCap, int& cellBlocks) {
    cellBlocks = (numCells + kStatThreads - 1) / kStatThreads;
    return cellBlocks + (partCap + kStatThreads - 1) / kStatThreads;
}

void launch_statistics(const Ctx& c, int wi, int cellBlocks, int blocks, int* cnt, double* en, hipStream_t s) {
    if (blocks > 0) hipLaunchKernelGGL(k_statistics, dim3(blocks), dim3(kStatThreads), 0, s, c, wi, cellBlocks, cnt, en);
}

// ---------------------------------------------------------------------------------------------
// Statistics without allocation.  The per-block partials of k_statistics and the final record
// live in one file-scope device array, `scratch`, that every GPU the library runs on holds until
// the process ends.  Host code never passes the array's name to a kernel (on the host it is only
// a stand-in, not a device address): the address of each device's copy is looked up once per
// device with hipGetSymbolAddress, that device current, and only that address is used.
// Every call on one device shares the array: calls that overlap on one device race.  Nothing in
// it needs to start at zero: k_statistics writes every partial it covers and k_statistics_total
// the whole final record.
// ---------------------------------------------------------------------------------------------
constexpr size_t kStatEnBytes = sizeof(double) * size_t(kStatScratchBlocks) * kColors;
constexpr size_t kStatCntBytes = sizeof(int) * size_t(kStatScratchBlocks) * kColors;
constexpr size_t kStatTotalOff = (kStatEnBytes + kStatCntBytes + 255) / 256 * 256;
constexpr size_t kStatScratchBytes = kStatTotalOff + (sizeof(alien_statistics) + 255) / 256 * 256;
static __device__ __align__(256) unsigned char scratch[kStatScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratchMutex;
void* g_scratchPtr[kMaxDevices] = {};
}  // namespace

hipError_t statistics_scratch(void** ptr) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratchPtr[dev] = p;
    }
    *ptr = g_scratchPtr[dev];
    return hipSuccess;
}

// The final record from the partials: per colour, counts summed (order-free in integers) and
// energies added block after block from +0, the same sequence of double additions as the host
// loop of the allocating path, so both paths return the same bits.  One wave per colour; a wave
// reads 64 partials at once and its first lane adds them in order.
__global__ void __launch_bounds__(64 * kColors) k_statistics_total(const int* cnt, const double* en, int cellBlocks,
                                                                   int blocks, alien_statistics* out) {
    const int k = int(threadIdx.x) >> 6, lane = int(threadIdx.x) & 63;
    int cells = 0, parts = 0;
    double e = 0.0;
    for (int b0 = 0; b0 < blocks; b0 += 64) {
        const int b = b0 + lane;
        const int n = b < blocks ? cnt[size_t(b) * kColors + k] : 0;
        const double v = b < blocks ? en[size_t(b) * kColors + k] : 0.0;
        cells += b < cellBlocks ? n : 0;
        parts += b < cellBlocks ? 0 : n;
        const int m = min(64, blocks - b0);
        for (int l = 0; l < m; ++l) e += __shfl(v, l, 64);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        cells += __shfl_down(cells, o, 64);
        parts += __shfl_down(parts, o, 64);
    }
    if (lane == 0) {
        out->numCells[k] = cells;
        out->numParticles[k] = parts;
        out->totalEnergy[k] = e;
        if (k == 0) out->timestep = 0;
    }
}

hipError_t launch_statistics_noalloc(const Ctx& c, int wi, int cellBlocks, int blocks, void* scratchPtr,
                                     alien_statistics* hostOut, hipStream_t s) {
    if (blocks <= 0 || blocks > kStatScratchBlocks) return hipErrorInvalidValue;
    unsigned char* base = static_cast<unsigned char*>(scratchPtr);
    double* en = reinterpret_cast<double*>(base);
    int* cnt = reinterpret_cast<int*>(base + kStatEnBytes);
    alien_statistics* total = reinterpret_cast<alien_statistics*>(base + kStatTotalOff);
    hipLaunchKernelGGL(k_statistics, dim3(blocks), dim3(kStatThreads), 0, s, c, wi, cellBlocks, cnt, en);
    hipLaunchKernelGGL(k_statistics_total, dim3(1), dim3(64 * kColors), 0, s, cnt, en, cellBlocks, blocks, total);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(hostOut, total, sizeof(alien_statistics), hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

}  // namespace alien
