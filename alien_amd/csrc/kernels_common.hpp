// kernels_common.hpp — device helpers shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "engine_types.hpp"

namespace alien {

__device__ __forceinline__ void slot_of(f2 p, int W, int H, int& sx, int& sy) {
    sx = wrap_int(floor_int(p.x), W);
    sy = wrap_int(floor_int(p.y), H);
}
// floor(a / b) for 0 <= a < 2^24, 0 < b < 2^24: the product with the approximate reciprocal is
// within a few units of the quotient, the corrections make it exact
__device__ __forceinline__ int div_small(int a, int b) {
    int q = int(float(a) * __builtin_amdgcn_rcpf(float(b)));
    while ((q + 1) * b <= a) ++q;
    while (q > 0 && q * b > a) --q;
    return q;
}
// floor(s / binSize) for a slot coordinate (a shift for power-of-two bins)
__device__ __forceinline__ int bin_coord(const DevWorld& w, int s) {
    return w.binShift >= 0 ? s >> w.binShift : div_small(s, w.binSize);
}
__device__ __forceinline__ int cell_bin_local(const DevWorld& w, int sx, int sy) {
    int bx = min(bin_coord(w, sx), w.nbx - 1), by = min(bin_coord(w, sy), w.nby - 1);
    return by * w.nbx + bx;
}
__device__ __forceinline__ int part_bin_local(const DevWorld& w, int sx, int sy) {
    int bx = min(sx / kParticleBin, w.pnbx - 1), by = min(sy / kParticleBin, w.pnby - 1);
    return by * w.pnbx + bx;
}
__device__ __forceinline__ int wrap_small(int x, int n) { return x < 0 ? x + n : (x >= n ? x - n : x); }

// global new-cell index range of local cell bin b
__device__ __forceinline__ void cell_bin_range(const Ctx& c, const DevWorld& w, int b, int& s, int& e) {
    uint32_t base = c.start[w.binOff];
    s = w.cellOff + int(c.start[w.binOff + b] - base);
    e = w.cellOff + int(c.start[w.binOff + b + 1] - base);
}
// (clamped to the world's particle capacity: a map beyond it is reported by the particle scatter)
__device__ __forceinline__ void part_bin_range(const Ctx& c, const DevWorld& w, int b, int& s, int& e) {
    uint32_t base = c.start[w.pbinOff];
    s = w.partOff + min(int(c.start[w.pbinOff + b] - base), w.partCap);
    e = w.partOff + min(int(c.start[w.pbinOff + b + 1] - base), w.partCap);
}

// cell bin of a position (global bin index), and its bin coordinates
__device__ __forceinline__ void cell_bin_xy(const DevWorld& w, f2 p, int& bx, int& by) {
    int sx, sy;
    slot_of(p, w.W, w.H, sx, sy);
    bx = min(bin_coord(w, sx), w.nbx - 1);
    by = min(bin_coord(w, sy), w.nby - 1);
}
__device__ __forceinline__ int cell_bin_of(const DevWorld& w, f2 p) {
    int bx, by;
    cell_bin_xy(w, p, bx, by);
    return w.binOff + by * w.nbx + bx;
}

// The map of a step orders cells by bin, and inside a bin by their index in the last step's order
// (R1).  The kernels that produce a cell's end-of-step position record its next bin (binOf).  A
// cell that stays in its bin keeps its place among the bin's cells of the last map; a cell that
// changes bins (a "mover"; after seeding every cell) is marked (kMoverBit), counted out of its old
// bin and into its new one (delta), and entered in the list of its new bin: head[bin] is the last
// one entered, next[i] the one before (kNone ends a list; ranks count the list's smaller indices,
// so its order does not matter).
__device__ __forceinline__ void push_mover(int* head, int* next, int gb, int k) { next[k] = atomicExch(&head[gb], k); }

// next-step bin of the cell at end-of-step index k: end-of-step position pNew; gbOld, its bin in
// this step's map (the bin of its start-of-step position)
__device__ __forceinline__ void bin_next(const Ctx& c, const DevWorld& w, int k, f2 pNew, int gbOld) {
    const int gb = cell_bin_of(w, pNew);
    if (gb == gbOld) {
        c.binOf[k] = gb;
        return;
    }
    c.binOf[k] = int(uint32_t(gb) | kMoverBit);
    push_mover(c.moverHeadNext, c.moverNext, gb, k);
    atomicAdd(&c.deltaNext[gb], 1);
    atomicSub(&c.deltaNext[gbOld], 1);
}

__device__ __forceinline__ void pbin_for_next_step(const Ctx& c, const DevWorld& w, int k, f2 p, uint32_t* hist) {
    int sx, sy;
    slot_of(p, w.W, w.H, sx, sy);
    int gb = w.pbinOff + part_bin_local(w, sx, sy);
    c.pbinOf[k] = gb;
    atomicAdd(&hist[gb], 1u);
}

// Inclusive sum over the 64 lanes of a wave (all lanes active), with DPP row shifts and row
// broadcasts instead of LDS permutes: sums of up to four lanes, of eight, of the 16-lane row, then
// row 0's total into row 1 and row 2's into row 3, then lane 31's prefix into rows 2 and 3.
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t x) {
    uint32_t v = x;
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x111, 0xf, 0xf, true));   // row_shr:1
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x112, 0xf, 0xf, true));   // row_shr:2
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(x), 0x113, 0xf, 0xf, true));   // row_shr:3
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x114, 0xf, 0xe, false));  // row_shr:4, lanes 4-15
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x118, 0xf, 0xc, false));  // row_shr:8, lanes 8-15
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x142, 0xa, 0xf, false));  // row_bcast:15, rows 1, 3
    v += uint32_t(__builtin_amdgcn_update_dpp(0, int(v), 0x143, 0xc, 0xf, false));  // row_bcast:31, rows 2, 3
    return v;
}

// the value of lane 63 (a wave total after an inclusive scan), read as a scalar
__device__ __forceinline__ int wave_last(int v) { return __builtin_amdgcn_readlane(v, 63); }

// exclusive scan across the block (blockDim multiple of 64, at most 1024); block total in `total`
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t waveSums[17];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int numWaves = blockDim.x >> 6;
    uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) waveSums[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        for (int k = 0; k < numWaves; ++k) {
            uint32_t t = waveSums[k];
            waveSums[k] = s;
            s += t;
        }
        waveSums[16] = s;
    }
    __syncthreads();
    uint32_t res = inc - v + waveSums[wave];
    total = waveSums[16];
    __syncthreads();
    return res;
}

// grid-stride iteration over the blocks of live particles (the table the particle bin sort writes
// for the live ranges it publishes)
template <class F> __device__ __forceinline__ void for_live_particles(const Ctx& c, F&& f) {
    const int n = *c.numPartBlocks;
    for (int e = blockIdx.x; e < n; e += gridDim.x) {
        const int2 blk = c.partBlocks[e];
        const int live = c.ctr[blk.x].numParticles;
        const DevWorld& w = c.worlds[blk.x];
        const int k = blk.y + int(threadIdx.x);
        if (k < w.partOff + live) f(blk.x, w, k);
    }
}

}  // namespace alien
