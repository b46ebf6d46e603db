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
__device__ __forceinline__ int cell_bin_local(const DevWorld& w, int sx, int sy) {
    int bx = min(div_small(sx, w.binSize), w.nbx - 1), by = min(div_small(sy, w.binSize), w.nby - 1);
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
__device__ __forceinline__ void part_bin_range(const Ctx& c, const DevWorld& w, int b, int& s, int& e) {
    uint32_t base = c.start[w.pbinOff];
    s = w.partOff + int(c.start[w.pbinOff + b] - base);
    e = w.partOff + int(c.start[w.pbinOff + b + 1] - base);
}

// rank of entry value v among the entries of sorted[s, e): the cell's position in its bin under
// the ascending-previous-index order (R1)
__device__ __forceinline__ int rank_in_bin(const int* sorted, int s, int e, int v) {
    int r = 0;
    for (int t = s; t < e; ++t) r += sorted[t] < v ? 1 : 0;
    return r;
}

// counting-sort bookkeeping: a cell counted in the shared (atomic) histogram records its place
// among the bin's shared arrivals (the count before its own) with this mark
constexpr int kArrivalShared = 0x40000000;

// cell bin of a position (global bin index)
__device__ __forceinline__ int cell_bin_of(const DevWorld& w, f2 p) {
    int sx, sy;
    slot_of(p, w.W, w.H, sx, sy);
    return w.binOff + cell_bin_local(w, sx, sy);
}

// counting-sort bookkeeping for the next step: bin of the end-of-step position, counted into the
// shared histogram `hist`, and the cell's place among the bin's shared arrivals
__device__ __forceinline__ void bin_for_next_step(const Ctx& c, const DevWorld& w, int k, f2 p, uint32_t* hist) {
    const int gb = cell_bin_of(w, p);
    c.binOf[k] = gb;
    c.arrival[k] = kArrivalShared | int(atomicAdd(&hist[gb], 1u));
}
// bin_for_next_step for a whole wave (every lane calls it; lanes with !valid count nothing): one
// atomic per run of consecutive lanes whose cells land in the same bin (cells in map order mostly
// stay together), each cell ranked inside its run's share of the bin's shared arrivals
__device__ __forceinline__ void bin_for_next_step_runs(const Ctx& c, const DevWorld& w, int k, f2 p, uint32_t* hist,
                                                       bool valid) {
    const int lane = int(threadIdx.x) & 63;
    const int gb = valid ? cell_bin_of(w, p) : -1;
    const int prev = __shfl_up(gb, 1, 64);
    const bool start = valid && (lane == 0 || prev != gb);
    const unsigned long long starts = __ballot(start);
    const unsigned long long bounds = starts | __ballot(!valid);  // a run ends at a start or an invalid lane
    const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;  // lanes 0..lane
    const unsigned long long sm = starts & upto;
    const int rs = sm ? 63 - __clzll(sm) : lane;  // the run's first lane
    const unsigned long long after = bounds & ~upto;
    const int re = after ? __ffsll(after) - 1 : 64;  // one past the run's last lane
    int base = 0;
    if (start) base = int(atomicAdd(&hist[gb], uint32_t(re - lane)));
    base = __shfl(base, rs, 64);
    if (valid) {
        c.binOf[k] = gb;
        c.arrival[k] = kArrivalShared | (base + lane - rs);
    }
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

// grid-stride iteration over the particle block table, live particles only
template <class F> __device__ __forceinline__ void for_live_particles(const Ctx& c, F&& f) {
    for (int e = blockIdx.x; e < c.numPartBlocks; e += gridDim.x) {
        const int2 blk = c.partBlocks[e];
        const int live = c.ctr[blk.x].numParticles;
        const DevWorld& w = c.worlds[blk.x];
        const int k = blk.y + int(threadIdx.x);
        if (k < w.partOff + live) f(blk.x, w, k);
    }
}

}  // namespace alien
