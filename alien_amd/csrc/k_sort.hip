// k_sort.hip — spatial hash: counting sort of cells and energy particles by bin.
//
// Replaces ALIEN's CellMap / ParticleMap (source/EngineGpuKernels/Map.cuh:71-280), which insert
// pointers into a dense world-size grid with atomicCAS chains (cell map) or last-writer-wins
// stores (particle map).  Here every bin of every world is one counter; the bins of the next step
// are counted by the kernels that produce the end-of-step positions (bin_for_next_step), then one
// exclusive scan, one scatter and, for particles, an in-bin sort give each bin its entries.  Cell
// bins are ranked inside the neighbour-pass tiles instead (see k_neighbor.hip).
#include "kernels_common.hpp"
#include "kernels.hpp"

namespace alien {

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// bins of the current state, used after data is set and after introspection
__global__ void k_cell_bin(Ctx c) {
    int2 blk = c.cellBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int i = blk.y + int(threadIdx.x);
    if (i >= w.cellOff + w.numCells) return;
    bin_for_next_step(c, w, i, to_f2(c.A.pos[i]), c.count);
}

__global__ void k_part_bin(Ctx c) {
    for_live_particles(c, [&](int, const DevWorld& w, int i) {
        if (c.A.pflags[i] & kParticleDead)
            c.pbinOf[i] = -1;
        else
            pbin_for_next_step(c, w, i, to_f2(c.A.ppos[i]), c.count);
    });
}

// exclusive scan of the bin sizes into start[0..n]: tile totals, then tile scans that each add
// the totals of the tiles before them
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// bin sizes base..base+7: shared count plus, for cell bins, the owner-counted part
__device__ __forceinline__ void load8(const uint32_t* count, const uint32_t* priv, int nCell, size_t base, int n,
                                      uint32_t v[8]) {
    if (base + 8 <= size_t(n)) {
        const uint4 a = *reinterpret_cast<const uint4*>(count + base);
        const uint4 b = *reinterpret_cast<const uint4*>(count + base + 4);
        v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = base + k < size_t(n) ? count[base + k] : 0u;
    }
    if (base + 8 <= size_t(nCell)) {
        const uint4 a = *reinterpret_cast<const uint4*>(priv + base);
        const uint4 b = *reinterpret_cast<const uint4*>(priv + base + 4);
        v[0] += a.x, v[1] += a.y, v[2] += a.z, v[3] += a.w, v[4] += b.x, v[5] += b.y, v[6] += b.z, v[7] += b.w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (base + k < size_t(nCell)) v[k] += priv[base + k];
    }
}

// tile totals; also opens the step: clears the per-step list counters, and advances the worlds'
// timesteps if `advance` is set (a step without bond work has no later reader of them)
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* count, const uint32_t* priv, int nCell,
                                                              uint32_t* blockSums, int n, StepCounters* sc,
                                                              DevCounters* advance, int numWorlds) {
    const size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanItems;
    uint32_t v[8];
    load8(count, priv, nCell, base, n, v);
    uint32_t s = v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[7];
    uint32_t total;
    block_exclusive_scan(s, total);
    if (threadIdx.x == 0) {
        blockSums[blockIdx.x] = total;
        if (sc && blockIdx.x == 0) {
            sc->candidates = 0;
            sc->liveParticles = 0;
        }
    }
    if (advance && blockIdx.x == 0)
        for (int wi = threadIdx.x; wi < numWorlds; wi += kScanThreads) advance[wi].timestep += 1ull;
}

// writes start[]: every tile adds the totals of the tiles before it (a few thousand words, read
// from L2) to its own exclusive scan.  `count` stays: the scatters clear it (cells) or count it down (particles) to zero,
// which leaves it cleared for the step after next (the two histograms alternate)
__global__ void __launch_bounds__(kScanThreads) k_scan_down(const uint32_t* count, const uint32_t* priv, int nCell,
                                                            const uint32_t* blockSums, uint32_t* start, int n) {
    uint32_t before = 0;
    for (int t = int(threadIdx.x); t < int(blockIdx.x); t += kScanThreads) before += blockSums[t];
    uint32_t prefix;
    block_exclusive_scan(before, prefix);
    const size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanItems;
    uint32_t v[8];
    load8(count, priv, nCell, base, n, v);
    uint32_t s = v[0] + v[1] + v[2] + v[3] + v[4] + v[5] + v[6] + v[7];
    uint32_t total;
    uint32_t off = block_exclusive_scan(s, total) + prefix;
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) start[n] = prefix + total;
    uint32_t o[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        o[k] = off;
        off += v[k];
    }
    if (base + 8 <= size_t(n)) {
        *reinterpret_cast<uint4*>(start + base) = make_uint4(o[0], o[1], o[2], o[3]);
        *reinterpret_cast<uint4*>(start + base + 4) = make_uint4(o[4], o[5], o[6], o[7]);
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (base + k < size_t(n)) start[base + k] = o[k];
    }
}

// entries land in their bin (the in-bin order is fixed later by ranking), with a copy of the
// state the neighbour pass reads: owner-counted cells at the rank they got when counted, the
// shared ones after them at the place their count returned.  A shared cell also clears its bin's
// shared counter (plain stores of zero), which leaves the histogram empty for the step after next.
__global__ void k_cell_scatter(Ctx c) {
    int2 blk = c.cellBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int i = blk.y + int(threadIdx.x);
    if (i >= w.cellOff + w.numCells) return;
    const float2 p = c.A.pos[i], v = c.A.vel[i];
    // the neighbour pass of a fluid world reads 1 / density
    const float d = c.params[blk.x].motion == 0 ? 1.0f / c.A.dens[i] : c.A.dens[i];
    const uint32_t f = c.A.flags[i];
    const int gb = c.binOf[i];
    const int arr = c.arrival[i];
    int rank = arr;
    if (arr & kArrivalShared) {
        rank = int(c.priv[gb]) + (arr & ~kArrivalShared);
        c.count[gb] = 0u;
    }
    const int dest = w.cellOff + int(c.start[gb] - c.start[w.binOff]) + rank;
    c.sorted[dest] = i;
    c.mpos[dest] = p;
    c.mvel[dest] = v;
    c.mdens[dest] = d;
    c.mflags[dest] = f;
}

__global__ void k_part_scatter(Ctx c) {
    for_live_particles(c, [&](int, const DevWorld& w, int i) {
        int gb = c.pbinOf[i];
        if (gb < 0) return;
        int dest = w.partOff + int(c.start[gb] - c.start[w.pbinOff]) + int(atomicSub(&c.count[gb], 1u)) - 1;
        c.psorted[dest] = i;
    });
}

// sort the entries of one bin ascending (R1 stable order), small bins in registers
template <bool kWriteInverse>
__device__ __forceinline__ void sort_bin(int* arr, int s, int e, int* inverse) {
    const int n = e - s;
    if (n <= 0) return;
    if (n == 1) {
        if (kWriteInverse) inverse[arr[s]] = s;
        return;
    }
    if (n <= 16) {
        int v[16];
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (k < n) v[k] = arr[s + k];
        for (int k = 1; k < n; ++k) {
            int x = v[k];
            int j = k - 1;
            while (j >= 0 && v[j] > x) {
                v[j + 1] = v[j];
                --j;
            }
            v[j + 1] = x;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (k < n) {
                arr[s + k] = v[k];
                if (kWriteInverse) inverse[v[k]] = s + k;
            }
        return;
    }
    for (int k = s + 1; k < e; ++k) {
        int x = arr[k];
        int j = k - 1;
        while (j >= s && arr[j] > x) {
            arr[j + 1] = arr[j];
            --j;
        }
        arr[j + 1] = x;
    }
    if (kWriteInverse)
        for (int k = s; k < e; ++k) inverse[arr[k]] = k;
}

// introspection only: cell map in logical order
__global__ void k_cell_binsort(Ctx c) {
    int2 blk = c.binBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int gb = blk.y + int(threadIdx.x);
    if (gb >= w.binOff + w.nbx * w.nby) return;
    int s, e;
    cell_bin_range(c, w, gb - w.binOff, s, e);
    sort_bin<true>(c.sorted, s, e, c.newidx);
}

// particle map in logical order; also publishes the new number of live particles per world
__global__ void k_part_binsort(Ctx c) {
    if (blockIdx.x == 0) {
        for (int wi = threadIdx.x; wi < c.numWorlds; wi += blockDim.x) {
            const DevWorld& w = c.worlds[wi];
            const int live = int(c.start[w.pbinOff + w.pnbx * w.pnby] - c.start[w.pbinOff]);
            c.ctr[wi].numParticles = live;
            if (live > 0) atomicAdd(&c.sc->liveParticles, live);
        }
    }
    int2 blk = c.pbinBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int gb = blk.y + int(threadIdx.x);
    if (gb >= w.pbinOff + w.pnbx * w.pnby) return;
    int s, e;
    part_bin_range(c, w, gb - w.pbinOff, s, e);
    sort_bin<false>(c.psorted, s, e, nullptr);
}

// ---------------------------------------------------------------------------------------------
void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return;
    size_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(k_fill_u32, dim3(unsigned(g)), dim3(256), 0, s, p, v, n);
}

void launch_seed_bins(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    // every cell is counted in the shared histogram; the owner-counted part starts empty
    launch_fill_u32(c.priv, 0u, size_t(c.numCellBins), s);
    if (sh.cellBlocks) hipLaunchKernelGGL(k_cell_bin, dim3(sh.cellBlocks), dim3(sh.cellThreads), 0, s, c);
    if (sh.partBlocks) hipLaunchKernelGGL(k_part_bin, dim3(sh.partGrid), dim3(sh.partThreads), 0, s, c);
}

void launch_scan(const Ctx& c, hipStream_t s, StepCounters* sc, bool advanceTimestep) {
    const int tiles = (c.numBins + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(kScanThreads), 0, s, c.count, c.priv, c.numCellBins,
                       c.blockSums, c.numBins, sc, advanceTimestep ? c.ctr : nullptr, c.numWorlds);
    hipLaunchKernelGGL(k_scan_down, dim3(tiles), dim3(kScanThreads), 0, s, c.count, c.priv, c.numCellBins,
                       c.blockSums, c.start, c.numBins);
}

void launch_cell_scatter(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    if (sh.cellBlocks) hipLaunchKernelGGL(k_cell_scatter, dim3(sh.cellBlocks), dim3(sh.cellThreads), 0, s, c);
}

void launch_part_scatter(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    if (sh.partBlocks) hipLaunchKernelGGL(k_part_scatter, dim3(sh.partGrid), dim3(sh.partThreads), 0, s, c);
}

void launch_part_sort(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    if (!sh.hasParticles) return;
    if (sh.partBlocks) hipLaunchKernelGGL(k_part_scatter, dim3(sh.partGrid), dim3(sh.partThreads), 0, s, c);
    if (sh.pbinBlocks) hipLaunchKernelGGL(k_part_binsort, dim3(sh.pbinBlocks), dim3(sh.binThreads), 0, s, c);
}

void launch_cell_binsort(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    if (sh.binBlocks) hipLaunchKernelGGL(k_cell_binsort, dim3(sh.binBlocks), dim3(sh.binThreads), 0, s, c);
}

}  // namespace alien
