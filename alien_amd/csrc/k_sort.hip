// k_sort.hip — spatial hash: counting sort of cells and energy particles by bin.
//
// Replaces ALIEN's CellMap / ParticleMap (source/EngineGpuKernels/Map.cuh:71-280), which insert
// pointers into a dense world-size grid with atomicCAS chains (cell map) or last-writer-wins
// stores (particle map).  Here every bin of every world is one counter and one offset.  Cells:
// the kernels that produce the end-of-step positions record each cell's next bin; at the start of
// the next step the map count gives every bin its size from the same bin of the last map (the
// cells that stay) plus the list of cells that moved in, one exclusive scan gives the offsets, and
// the scatter places every cell at its rank among the bin's cells by previous index (R1), so the
// map order is the logical order of the step.  Particles: counted with atomics, scattered, then sorted inside their bins.
#include "kernels_common.hpp"
#include "kernels.hpp"

namespace alien {

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// bins of the current state, used after data is set and after introspection: the current state
// is in no map order yet, so every cell is a mover of the coming step
__global__ void k_cell_bin(Ctx c) {
    int2 blk = c.cellBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int i = blk.y + int(threadIdx.x);
    if (i >= w.cellOff + w.numCells) return;
    const int gb = cell_bin_of(w, to_f2(c.A.pos[i]));
    c.binOf[i] = int(uint32_t(gb) | kMoverBit);
    push_mover(c.moverHead, c.moverNext, gb, i);
    atomicAdd(&c.delta[gb], 1);
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

// bin sizes base..base+7: a cell bin's size is its size in the last map plus its delta, a particle
// bin's its count
__device__ __forceinline__ void load8(const uint32_t* count, const uint32_t* startPrev, const int* delta, int nCell,
                                      size_t base, int n, uint32_t v[8]) {
    if (base + 8 <= size_t(n)) {
        const uint4 a = *reinterpret_cast<const uint4*>(count + base);
        const uint4 b = *reinterpret_cast<const uint4*>(count + base + 4);
        v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = base + k < size_t(n) ? count[base + k] : 0u;
    }
    if (base >= size_t(nCell)) return;
    uint32_t sp[9];
    int dl[8];
    if (base + 8 <= size_t(nCell)) {
        const uint4 a = *reinterpret_cast<const uint4*>(startPrev + base);
        const uint4 b = *reinterpret_cast<const uint4*>(startPrev + base + 4);
        sp[0] = a.x, sp[1] = a.y, sp[2] = a.z, sp[3] = a.w, sp[4] = b.x, sp[5] = b.y, sp[6] = b.z, sp[7] = b.w;
        sp[8] = startPrev[base + 8];
        const int4 c0 = *reinterpret_cast<const int4*>(delta + base);
        const int4 c1 = *reinterpret_cast<const int4*>(delta + base + 4);
        dl[0] = c0.x, dl[1] = c0.y, dl[2] = c0.z, dl[3] = c0.w, dl[4] = c1.x, dl[5] = c1.y, dl[6] = c1.z, dl[7] = c1.w;
    } else {
        // (the end of the last cell bin is the start of the particle bins)
#pragma unroll
        for (int k = 0; k < 9; ++k) sp[k] = base + k <= size_t(nCell) ? startPrev[base + k] : 0u;
#pragma unroll
        for (int k = 0; k < 8; ++k) dl[k] = base + k < size_t(nCell) ? delta[base + k] : 0;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k)
        if (base + k < size_t(nCell)) v[k] = uint32_t(int(sp[k + 1] - sp[k]) + dl[k]);
}

// tile totals; also opens the step: clears the per-step list counters, and advances the worlds'
// timesteps if `advance` is set (a step without bond work has no later reader of them)
__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* count, const uint32_t* startPrev,
                                                              const int* delta, int nCell, uint32_t* blockSums, int n,
                                                              StepCounters* sc,
                                                              DevCounters* advance, int numWorlds) {
    const size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanItems;
    uint32_t v[8];
    load8(count, startPrev, delta, nCell, base, n, v);
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
// from L2) to its own exclusive scan.  `count` stays: the particle scatter counts its particle
// bins down to zero, which leaves them cleared for the step after next (the two histograms
// alternate); the map count writes the cell bins every step
__global__ void __launch_bounds__(kScanThreads) k_scan_down(const uint32_t* count, const uint32_t* startPrev,
                                                            const int* delta, int nCell, const uint32_t* blockSums,
                                                            uint32_t* start, int n, int* deltaNext, int* moverHeadNext) {
    uint32_t before = 0;
    for (int t = int(threadIdx.x); t < int(blockIdx.x); t += kScanThreads) before += blockSums[t];
    uint32_t prefix;
    block_exclusive_scan(before, prefix);
    const size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanItems;
    uint32_t v[8];
    load8(count, startPrev, delta, nCell, base, n, v);
    // the next step's deltas and mover lists start empty (their last readers ran a step ago)
    if (deltaNext) {
        if (base + 8 <= size_t(nCell)) {
            *reinterpret_cast<int4*>(deltaNext + base) = make_int4(0, 0, 0, 0);
            *reinterpret_cast<int4*>(deltaNext + base + 4) = make_int4(0, 0, 0, 0);
            *reinterpret_cast<int4*>(moverHeadNext + base) = make_int4(kNone, kNone, kNone, kNone);
            *reinterpret_cast<int4*>(moverHeadNext + base + 4) = make_int4(kNone, kNone, kNone, kNone);
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (base + k < size_t(nCell)) {
                    deltaNext[base + k] = 0;
                    moverHeadNext[base + k] = kNone;
                }
        }
    }
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

// Map scatter: cell i (A index, i.e. its place in the last map) goes to its bin at its rank by A
// index among the bin's cells (R1).  The bin's stayers hold their places of the last map's bin in
// order, so a stayer's rank among them is its place in that bin minus the movers that left the bin
// before it (marked lanes of the wave, and cells of the bin in earlier waves), and a mover from a
// lower place precedes them all, one from a higher place follows them all; every cell then adds
// the bin's movers with a smaller A index.
__global__ void k_cell_scatter(Ctx c) {
    int2 blk = c.cellBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    const int i = blk.y + int(threadIdx.x);
    const int lane = int(threadIdx.x) & 63;
    const int i0 = i - lane;  // the wave's first cell
    const bool active = i < w.cellOff + w.numCells;
    const int bo = active ? c.binOf[i] : 0;
    const bool mover = (uint32_t(bo) & kMoverBit) != 0u;
    const unsigned long long movers = __ballot(active && mover);
    if (!active) return;
    const int gb = int(uint32_t(bo) & ~kMoverBit);
    const uint32_t base = c.startPrev[w.binOff];
    const int s = w.cellOff + int(c.startPrev[gb] - base), e = w.cellOff + int(c.startPrev[gb + 1] - base);
    const int head = c.moverHead[gb];
    const int dest0 = w.cellOff + int(c.start[gb] - c.start[w.binOff]);
    int rank;
    if (!mover) {
        // a stayer: s <= i < e
        const int from = max(s, i0) - i0;
        const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
        const unsigned long long fromMask = from >= 64 ? 0ull : ~((1ull << from) - 1ull);
        int left = __builtin_popcountll(movers & below & fromMask);
        for (int j = s; j < i0; ++j) left += (uint32_t(c.binOf[j]) & kMoverBit) ? 1 : 0;
        rank = (i - s) - left;
    } else {
        rank = 0;
        if (i >= e)
            for (int j = s; j < e; ++j) rank += c.binOf[j] == gb ? 1 : 0;
    }
    for (int j = head; j != kNone; j = c.moverNext[j]) rank += j < i ? 1 : 0;
    const int dest = dest0 + rank;
    c.sorted[dest] = i;
    // the row-staged collision pass reads a copy of the state in map order; the fluid pass reads
    // the state through `sorted`
    if (c.params[blk.x].motion != 0) {
        c.mpos[dest] = c.A.pos[i];
        c.mvel[dest] = c.A.vel[i];
        c.mdens[dest] = c.A.dens[i];
        c.mflags[dest] = c.A.flags[i];
    }
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
    // no last map (every bin of it empty), every cell a mover of the coming step
    launch_fill_u32(const_cast<uint32_t*>(c.startPrev), 0u, size_t(c.numBins) + 1, s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(c.moverHead), uint32_t(kNone), size_t(c.numCellBins), s);
    launch_fill_u32(reinterpret_cast<uint32_t*>(c.delta), 0u, size_t(c.numCellBins), s);
    if (sh.cellBlocks) hipLaunchKernelGGL(k_cell_bin, dim3(sh.cellBlocks), dim3(sh.cellThreads), 0, s, c);
    if (sh.partBlocks) hipLaunchKernelGGL(k_part_bin, dim3(sh.partGrid), dim3(sh.partThreads), 0, s, c);
}

// exclusive scan of the bin sizes into start (c.start); with `resetNext` it also empties the
// next step's deltas and mover lists
void launch_scan(const Ctx& c, hipStream_t s, StepCounters* sc, bool advanceTimestep, bool resetNext) {
    const int tiles = (c.numBins + kScanTile - 1) / kScanTile;
    hipLaunchKernelGGL(k_scan_reduce, dim3(tiles), dim3(kScanThreads), 0, s, c.count, c.startPrev, c.delta,
                       c.numCellBins, c.blockSums, c.numBins, sc, advanceTimestep ? c.ctr : nullptr, c.numWorlds);
    hipLaunchKernelGGL(k_scan_down, dim3(tiles), dim3(kScanThreads), 0, s, c.count, c.startPrev, c.delta,
                       c.numCellBins, c.blockSums, c.start, c.numBins, resetNext ? c.deltaNext : nullptr,
                       c.moverHeadNext);
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
