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
        // a half split off particle i at the end of the last step
        if (c.splitStage > 0) {
            const StagedParticle& r = c.splitStaged[i];
            if (r.color & kStagedPending)
                pbin_for_next_step(c, w, c.splitStage + i, to_f2(r.pos), c.count);
            else
                c.pbinOf[c.splitStage + i] = -1;
        }
    });
}

constexpr int kScanItems = 8;  // consecutive bins per group of the map scan

// bin sizes base..base+7: a cell bin's size is its size in the last map plus its delta, a particle
// bin's its count (the cell bins come first; `count` is read for particle bins only)
__device__ __forceinline__ void load8(const uint32_t* count, const uint32_t* startPrev, const int* delta, int nCell,
                                      size_t base, int n, uint32_t v[8]) {
    if (base >= size_t(nCell) && base + 8 <= size_t(n)) {
        const uint4 a = *reinterpret_cast<const uint4*>(count + base);
        const uint4 b = *reinterpret_cast<const uint4*>(count + base + 4);
        v[0] = a.x, v[1] = a.y, v[2] = a.z, v[3] = a.w, v[4] = b.x, v[5] = b.y, v[6] = b.z, v[7] = b.w;
    } else if (base + 8 > size_t(nCell)) {
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = base + k >= size_t(nCell) && base + k < size_t(n) ? count[base + k] : 0u;
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

// ---------------------------------------------------------------------------------------------
// Offsets of a step's map in one pass over the bins.  Every workgroup takes the next tile of 8192
// bins (tickets in the order the workgroups start), sums its bin sizes and publishes the sum, then
// finds the sum of all bins before it by looking back over what the tiles before it published
// (decoupled look-back: a tile publishes its own sum before it looks back, and every tile before
// it holds an earlier ticket and has started, so each wait ends).  With that it writes the offsets
// of its bins.  It also empties the next step's deltas and mover lists of its cell bins (their
// last readers ran a step ago), and the first tile opens the step: clears the per-step list
// counters and advances the worlds' timesteps if asked (a step without bond work has no later
// reader of them).  `count` stays: the particle scatter counts its particle bins down to zero,
// which leaves them cleared for the step after next (the two histograms alternate).
// ---------------------------------------------------------------------------------------------
constexpr int kMapScanThreads = 1024;
constexpr int kScanGroups = 1;  // groups of kScanItems consecutive bins per thread (2 measured slower)
constexpr int kMapScanTile = kMapScanThreads * kScanGroups * kScanItems;
// tile status word: flag (bits 62-63: 1 the tile's sum, 2 the sum up to and including the tile),
// launch epoch (bits 32-61), value (bits 0-31)
constexpr unsigned long long kTileSum = 1ull << 62, kTileInclusive = 2ull << 62;
constexpr uint32_t kEpochMask = 0x3fffffffu;
constexpr int kLookBackSpins = 1 << 22;  // far beyond any wait for a running tile

__device__ __forceinline__ bool tile_status_ready(unsigned long long w, uint32_t epoch) {
    return (w >> 62) != 0ull && uint32_t((w >> 32) & kEpochMask) == epoch;
}

__global__ void __launch_bounds__(kMapScanThreads) k_map_scan(Ctx c, int tiles, StepCounters* open, int advance) {
    __shared__ int sTile;
    __shared__ uint32_t sEpoch, sPrefix;
    const int tid = int(threadIdx.x), lane = tid & 63;
    // All atomics here are relaxed (agent scope): a status word carries its whole message (flag,
    // epoch, value), and the epoch and ticket a launch reads were written by the launches before
    // it.  Ordered ones would write back or invalidate the XCD's L2 at every store and load.
    if (tid == 0) {
        const uint32_t e = __hip_atomic_load(c.mapEpoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // the epoch has been read when the ticket is taken; the last ticket's holder advances it
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        sEpoch = e & kEpochMask;
        sTile = int(__hip_atomic_fetch_add(c.mapTicket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    }
    __syncthreads();
    const int tile = sTile;
    const uint32_t epoch = sEpoch;
    const unsigned long long ep = (unsigned long long)epoch << 32;
    const int n = c.numBins, nCell = c.numCellBins;
    const size_t base0 = size_t(tile) * kMapScanTile + size_t(tid) * (kScanGroups * kScanItems);
    uint32_t v[kScanGroups][kScanItems];
#pragma unroll
    for (int g = 0; g < kScanGroups; ++g) load8(c.count, c.startPrev, c.delta, nCell, base0 + g * kScanItems, n, v[g]);
    uint32_t s = 0;
#pragma unroll
    for (int g = 0; g < kScanGroups; ++g)
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) s += v[g][k];
    uint32_t total;
    const uint32_t off = block_exclusive_scan(s, total);
    if (tid == 0)
        __hip_atomic_store(&c.mapStatus[tile], (tile == 0 ? kTileInclusive : kTileSum) | ep | total, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    if (tid < 64) {
        // the first wave looks back, 64 tiles at a time, until a tile with its inclusive sum
        uint32_t prefix = 0;
        if (tile > 0) {
            for (int j = tile - 1;; j -= 64) {
                const int idx = j - lane;
                unsigned long long w = kTileInclusive | ep;  // before the first tile: sum 0
                if (idx >= 0) {
                    // (bounded: a tile that never published would be a bug; the step then reports
                    // an error instead of hanging)
                    for (int spin = 0;; ++spin) {
                        w = __hip_atomic_load(&c.mapStatus[idx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (tile_status_ready(w, epoch)) break;
                        if (spin >= kLookBackSpins) {
                            atomicOr(&c.sc->errors, 2);
                            w = kTileInclusive | ep;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                const unsigned long long incl = __ballot((w >> 62) == 2ull);
                const int stop = incl ? __builtin_ctzll(incl) : 64;
                uint32_t val = lane <= stop ? uint32_t(w) : 0u;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o, 64);
                prefix += val;
                if (incl) break;
            }
            if (lane == 0)
                __hip_atomic_store(&c.mapStatus[tile], kTileInclusive | ep | (prefix + total), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) sPrefix = prefix;
    }
    // the next step's deltas and mover lists start empty
#pragma unroll
    for (int g = 0; g < kScanGroups; ++g) {
        const size_t base = base0 + g * kScanItems;
        if (base + 8 <= size_t(nCell)) {
            *reinterpret_cast<int4*>(c.deltaNext + base) = make_int4(0, 0, 0, 0);
            *reinterpret_cast<int4*>(c.deltaNext + base + 4) = make_int4(0, 0, 0, 0);
            *reinterpret_cast<int4*>(c.moverHeadNext + base) = make_int4(kNone, kNone, kNone, kNone);
            *reinterpret_cast<int4*>(c.moverHeadNext + base + 4) = make_int4(kNone, kNone, kNone, kNone);
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (base + k < size_t(nCell)) {
                    c.deltaNext[base + k] = 0;
                    c.moverHeadNext[base + k] = kNone;
                }
        }
    }
    if (tile == 0) {
        if (open && tid == 0) {
            open->candidates = 0;
            open->liveParticles = 0;
        }
        if (advance)
            for (int wi = tid; wi < c.numWorlds; wi += kMapScanThreads) c.ctr[wi].timestep += 1ull;
    }
    __syncthreads();
    uint32_t o = sPrefix + off;
#pragma unroll
    for (int g = 0; g < kScanGroups; ++g) {
        const size_t base = base0 + g * kScanItems;
        uint32_t st8[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            st8[k] = o;
            o += v[g][k];
        }
        if (base + 8 <= size_t(n)) {
            *reinterpret_cast<uint4*>(c.start + base) = make_uint4(st8[0], st8[1], st8[2], st8[3]);
            *reinterpret_cast<uint4*>(c.start + base + 4) = make_uint4(st8[4], st8[5], st8[6], st8[7]);
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if (base + k < size_t(n)) c.start[base + k] = st8[k];
        }
        if (base <= size_t(n - 1) && size_t(n - 1) < base + 8) c.start[n] = o;
    }
    // the last ticket: every ticket is gone; reset the ticket and advance the epoch for the next
    // launch
    if (tile == tiles - 1 && tid == 0) {
        __hip_atomic_store(c.mapTicket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(c.mapEpoch, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Map scatter: cell i (A index, i.e. its place in the last map) goes to its bin at its rank by A
// index among the bin's cells (R1).  The bin's stayers hold their places of the last map's bin in
// order, so a stayer's rank among them is its place in that bin minus the movers that left the bin
// before it (marked lanes of the wave, and cells of the bin in earlier waves), and a mover from a
// lower place precedes them all, one from a higher place follows them all; every cell then adds
// the bin's movers with a smaller A index.  A workgroup places the cells of kScatterRuns entries of
// the cell block table, one cell of each per thread, with the loads of each phase issued for all
// of them together (the pass is a chain of dependent loads: bin, then the bin's offsets and mover
// list, then the store).
constexpr int kScatterRuns = 4;

__global__ void k_cell_scatter(Ctx c, int tableBlocks) {
    const int lane = int(threadIdx.x) & 63;
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    int i[kScatterRuns], bo[kScatterRuns], wi[kScatterRuns], cellOff[kScatterRuns], binOff[kScatterRuns];
    bool active[kScatterRuns];
#pragma unroll
    for (int q = 0; q < kScatterRuns; ++q) {
        const int e = int(blockIdx.x) * kScatterRuns + q;
        const int2 blk = e < tableBlocks ? c.cellBlocks[e] : make_int2(0, 0);
        const DevWorld& w = c.worlds[blk.x];
        wi[q] = blk.x;
        cellOff[q] = w.cellOff;
        binOff[q] = w.binOff;
        i[q] = blk.y + int(threadIdx.x);
        active[q] = e < tableBlocks && i[q] < w.cellOff + w.numCells;
        bo[q] = active[q] ? c.binOf[i[q]] : 0;
    }
    int gb[kScatterRuns], s[kScatterRuns], e[kScatterRuns], head[kScatterRuns], dest0[kScatterRuns];
    unsigned long long movers[kScatterRuns];
#pragma unroll
    for (int q = 0; q < kScatterRuns; ++q) {
        const bool mover = (uint32_t(bo[q]) & kMoverBit) != 0u;
        movers[q] = __ballot(active[q] && mover);
        gb[q] = int(uint32_t(bo[q]) & ~kMoverBit);
        const uint32_t spBase = c.startPrev[binOff[q]], stBase = c.start[binOff[q]];
        const int g = active[q] ? gb[q] : binOff[q];
        s[q] = cellOff[q] + int(c.startPrev[g] - spBase);
        e[q] = cellOff[q] + int(c.startPrev[g + 1] - spBase);
        head[q] = active[q] ? c.moverHead[g] : kNone;
        dest0[q] = cellOff[q] + int(c.start[g] - stBase);
    }
    int rank[kScatterRuns];
#pragma unroll
    for (int q = 0; q < kScatterRuns; ++q) {
        rank[q] = 0;
        if (!active[q]) continue;
        const int iq = i[q], i0 = iq - lane;  // the wave's first cell
        if ((uint32_t(bo[q]) & kMoverBit) == 0u) {
            // a stayer: s <= i < e
            const int from = max(s[q], i0) - i0;
            const unsigned long long fromMask = from >= 64 ? 0ull : ~((1ull << from) - 1ull);
            int left = __builtin_popcountll(movers[q] & below & fromMask);
            for (int j = s[q]; j < i0; ++j) left += (uint32_t(c.binOf[j]) & kMoverBit) ? 1 : 0;
            rank[q] = (iq - s[q]) - left;
        } else if (iq >= e[q]) {
            for (int j = s[q]; j < e[q]; ++j) rank[q] += c.binOf[j] == gb[q] ? 1 : 0;
        }
    }
    // the bins' mover lists, walked side by side (one load per list in flight per round)
    int jw[kScatterRuns];
#pragma unroll
    for (int q = 0; q < kScatterRuns; ++q) jw[q] = active[q] ? head[q] : kNone;
    for (;;) {
        bool more = false;
#pragma unroll
        for (int q = 0; q < kScatterRuns; ++q)
            if (jw[q] != kNone) {
                rank[q] += jw[q] < i[q] ? 1 : 0;
                more = true;
            }
        if (!more) break;
#pragma unroll
        for (int q = 0; q < kScatterRuns; ++q) jw[q] = jw[q] != kNone ? c.moverNext[jw[q]] : kNone;
    }
#pragma unroll
    for (int q = 0; q < kScatterRuns; ++q) {
        if (!active[q]) continue;
        const int iq = i[q];
        const int dest = dest0[q] + rank[q];
        c.sorted[dest] = iq;
        // the row-staged collision pass reads a copy of the state in map order; the fluid pass
        // reads the state through `sorted` (a copy for it costs the scatter more than it saves)
        if (c.params[wi[q]].motion != 0) {
            c.mpos[dest] = c.A.pos[iq];
            c.mvel[dest] = c.A.vel[iq];
            c.mdens[dest] = c.A.dens[iq];
            c.mflags[dest] = c.A.flags[iq];
        }
    }
}

// particle i (A index or staging slot) of bin gb to its place in the map (the order inside the bin
// is fixed by the bin sort)
__device__ __forceinline__ void scatter_particle(const Ctx& c, const DevWorld& w, int gb, int i) {
    const int rel = int(c.start[gb] - c.start[w.pbinOff]) + int(atomicSub(&c.count[gb], 1u)) - 1;
    if (rel < w.partCap)
        c.psorted[w.partOff + rel] = i;
    else
        atomicOr(&c.sc->errors, 4);  // (the host keeps room for a step's emissions: not reached)
}

__global__ void k_part_scatter(Ctx c) {
    for_live_particles(c, [&](int, const DevWorld& w, int i) {
        const int gb = c.pbinOf[i];
        if (gb >= 0) scatter_particle(c, w, gb, i);
        // the particle split off particle i at the end of the last step (cleared once read)
        if (c.splitStage > 0) {
            const int st = c.splitStage + i;
            const int g2 = c.pbinOf[st];
            if (g2 >= 0) {
                c.pbinOf[st] = -1;
                scatter_particle(c, w, g2, st);
            }
        }
    });
}

// the particles cell radiation emitted this step, from their staging slots (cleared once read); one
// workgroup per cell block
__global__ void k_part_scatter_staged(Ctx c) {
    const int2 blk = c.cellBlocks[blockIdx.x];
    const DevWorld& w = c.worlds[blk.x];
    const int i = blk.y + int(threadIdx.x);
    if (i >= w.cellOff + w.numCells) return;
    const int st = c.partStage + i;
    const int gb = c.pbinOf[st];
    if (gb < 0) return;
    c.pbinOf[st] = -1;
    scatter_particle(c, w, gb, st);
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

// Particle map in logical order (the particles cell radiation emitted this step, from their staging
// slots, sort behind every live particle's index); also publishes the new number of live
// particles per world and the table of live particle blocks.
__global__ void k_part_binsort(Ctx c) {
    if (blockIdx.x == 0) {
        for (int wi = threadIdx.x; wi < c.numWorlds; wi += blockDim.x) {
            const DevWorld& w = c.worlds[wi];
            const int live = min(int(c.start[w.pbinOff + w.pnbx * w.pnby] - c.start[w.pbinOff]), w.partCap);
            c.ctr[wi].numParticles = live;
            if (live > 0) atomicAdd(&c.sc->liveParticles, live);
        }
        __syncthreads();
        // the table of live particle blocks, world by world
        const int ppb = c.partBlockParticles;
        int base = 0;
        for (int wi = 0; wi < c.numWorlds; ++wi) {
            const DevWorld& w = c.worlds[wi];
            if (!w.active) continue;
            const int nb = (c.ctr[wi].numParticles + ppb - 1) / ppb;
            for (int j = int(threadIdx.x); j < nb; j += int(blockDim.x)) c.partBlocks[base + j] = make_int2(wi, w.partOff + j * ppb);
            base += nb;
        }
        if (threadIdx.x == 0) *c.numPartBlocks = base;
    }
    int2 blk = c.pbinBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int gb = blk.y + int(threadIdx.x);
    if (gb >= w.pbinOff + w.pnbx * w.pnby) return;
    int s, e;
    part_bin_range(c, w, gb - w.pbinOff, s, e);
    sort_bin<false>(c.psorted, s, e, nullptr);
}

// The movement of the particles in map order (RadiationProcessor.cuh:61-70: position += velocity *
// dt on the torus) into the end-of-step arrays, one thread per particle of the new live ranges; a
// particle emitted this step comes from its staging record.
__global__ void k_part_move(Ctx c) {
    // (the start-of-step and end-of-step arrays are distinct)
    const float2* __restrict__ appos = c.A.ppos;
    const float2* __restrict__ apvel = c.A.pvel;
    const float* __restrict__ apen = c.A.pen;
    const int* __restrict__ aplast = c.A.plast;
    const int* __restrict__ apslot = c.A.pslot;
    const uint8_t* __restrict__ apflags = c.A.pflags;
    const unsigned long long* __restrict__ apid = c.A.pid;
    float2* __restrict__ bppos = c.B.ppos;
    float2* __restrict__ bpvel = c.B.pvel;
    float* __restrict__ bpen = c.B.pen;
    int* __restrict__ bplast = c.B.plast;
    int* __restrict__ bpslot = c.B.pslot;
    uint8_t* __restrict__ bpflags = c.B.pflags;
    unsigned long long* __restrict__ bpid = c.B.pid;
    for_live_particles(c, [&](int wi, const DevWorld& w, int k) {
        const float dt = c.params[wi].dt;
        const int src = c.psorted[k];
        const bool rad = c.partStage > 0 && src >= c.partStage;
        if (rad || (c.splitStage > 0 && src >= c.splitStage)) {
            const StagedParticle r = rad ? c.staged[src - c.partStage] : c.splitStaged[src - c.splitStage];
            if (!rad) c.splitStaged[src - c.splitStage].color = r.color & ~kStagedPending;  // (in the arrays now)
            const f2 v = to_f2(r.vel);
            bppos[k] = to_float2(torus_pos(add(to_f2(r.pos), mul(v, dt)), w.W, w.H));
            bpvel[k] = r.vel;
            bpen[k] = r.energy;
            bplast[k] = kNone;
            bpslot[k] = kNone;
            bpflags[k] = uint8_t(r.color & kParticleColorMask);
            bpid[k] = r.id;
            return;
        }
        const f2 v = to_f2(apvel[src]);
        bppos[k] = to_float2(torus_pos(add(to_f2(appos[src]), mul(v, dt)), w.W, w.H));
        bpvel[k] = to_float2(v);
        bpen[k] = apen[src];
        bplast[k] = aplast[src];
        bpslot[k] = apslot[src];
        bpflags[k] = apflags[src];
        bpid[k] = apid[src];
    });
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

// exclusive scan of the bin sizes into start (c.start); with `open` it also opens the step (its
// list counters cleared) and with `advanceTimestep` advances the worlds' timesteps
void launch_map_scan(const Ctx& c, hipStream_t s, StepCounters* open, bool advanceTimestep) {
    const int tiles = map_scan_tiles(c.numBins);
    if (tiles == 0) return;
    hipLaunchKernelGGL(k_map_scan, dim3(tiles), dim3(kMapScanThreads), 0, s, c, tiles, open, advanceTimestep ? 1 : 0);
}

int map_scan_tiles(int numBins) { return (numBins + kMapScanTile - 1) / kMapScanTile; }

void launch_cell_scatter(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    if (sh.cellBlocks)
        hipLaunchKernelGGL(k_cell_scatter, dim3((sh.cellBlocks + kScatterRuns - 1) / kScatterRuns), dim3(sh.cellThreads),
                           0, s, c, sh.cellBlocks);
}

void launch_part_scatter(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    if (sh.partBlocks) hipLaunchKernelGGL(k_part_scatter, dim3(sh.partGrid), dim3(sh.partThreads), 0, s, c);
}

void launch_part_sort(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    if (!sh.hasParticles) return;
    if (sh.partBlocks) hipLaunchKernelGGL(k_part_scatter, dim3(sh.partGrid), dim3(sh.partThreads), 0, s, c);
    if (c.partStage > 0 && sh.cellBlocks)
        hipLaunchKernelGGL(k_part_scatter_staged, dim3(sh.cellBlocks), dim3(sh.cellThreads), 0, s, c);
    if (sh.pbinBlocks) hipLaunchKernelGGL(k_part_binsort, dim3(sh.pbinBlocks), dim3(sh.binThreads), 0, s, c);
    if (sh.partBlocks) hipLaunchKernelGGL(k_part_move, dim3(sh.partGrid), dim3(sh.partThreads), 0, s, c);
}

void launch_cell_binsort(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    if (sh.binBlocks) hipLaunchKernelGGL(k_cell_binsort, dim3(sh.binBlocks), dim3(sh.binThreads), 0, s, c);
}

}  // namespace alien
