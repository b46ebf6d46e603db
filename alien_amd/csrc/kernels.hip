// kernels.hip — gfx950 kernels of the physics timestep.
//
// One timestep of a world set is the fixed kernel sequence below; every kernel is data parallel
// over the cells, particles or bins of all worlds at once (a block never spans two worlds, so
// per-world parameters are uniform within a block):
//
//   bin build (cells, particles)  -> scan -> scatter -> in-bin sort   spatial hash, counting sort
//   neighbour pass                                                    collision or SPH forces,
//                                                                     overlap correction, Verlet
//                                                                     of unbonded cells, fusion
//                                                                     candidates
//   particle move / target / claim / apply / absorb                   energy particles
//   connection forces x2 with gathers                                 springs + angle forces,
//                                                                     Verlet of bonded cells
//   mark / fusion claim / apply / reset / delete / clear              structural operations
//
// Reference behaviour: source/EngineGpuKernels/SimulationKernelsLauncher.cu:25-91 and the
// processors it calls (CellProcessor.cuh, RadiationProcessor.cuh, CellConnectionProcessor.cuh).
// All floating-point work is plain IEEE single precision without contraction, and every reduction
// runs in a fixed order, so a world's results are bit-reproducible and independent of launch
// shape and of the other worlds in the set.
#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "engine_types.hpp"
#include "kernels.hpp"

namespace alien {

// ---------------------------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void slot_of(f2 p, int W, int H, int& sx, int& sy) {
    sx = wrap_int(floor_int(p.x), W);
    sy = wrap_int(floor_int(p.y), H);
}
__device__ __forceinline__ int cell_bin_local(const DevWorld& w, int sx, int sy) {
    int bx = min(sx / w.binSize, w.nbx - 1), by = min(sy / w.binSize, w.nby - 1);
    return by * w.nbx + bx;
}
__device__ __forceinline__ int part_bin_local(const DevWorld& w, int sx, int sy) {
    int bx = min(sx / kParticleBin, w.pnbx - 1), by = min(sy / kParticleBin, w.pnby - 1);
    return by * w.pnbx + bx;
}
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

// ---------------------------------------------------------------------------------------------
// step bookkeeping
// ---------------------------------------------------------------------------------------------
__global__ void k_step_begin(Ctx c) {
    if (threadIdx.x == 0) {
        c.sc->staged = 0;
        c.sc->candidates = 0;
    }
}

__global__ void k_step_end(Ctx c) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < c.numWorlds) c.ctr[w].timestep += 1ull;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    size_t stride = size_t(gridDim.x) * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_reset_density(Ctx c) {
    int2 blk = c.cellBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int i = blk.y + int(threadIdx.x);
    if (i >= w.cellOff + w.numCells) return;
    c.A.dens[i] = 1.0f;
}

// ---------------------------------------------------------------------------------------------
// spatial hash: counting sort of cells and particles by bin (cell map + particle map)
// ---------------------------------------------------------------------------------------------
__global__ void k_cell_bin(Ctx c) {
    int2 blk = c.cellBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int i = blk.y + int(threadIdx.x);
    if (i >= w.cellOff + w.numCells) return;
    f2 p = to_f2(c.A.pos[i]);
    int sx, sy;
    slot_of(p, w.W, w.H, sx, sy);
    int gb = w.binOff + cell_bin_local(w, sx, sy);
    c.binOf[i] = gb;
    c.rank[i] = int(atomicAdd(&c.count[gb], 1u));
}

__global__ void k_part_bin(Ctx c) {
    int2 blk = c.partBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int i = blk.y + int(threadIdx.x);
    if (i >= w.partOff + c.ctr[blk.x].numParticles) return;
    if (c.A.pflags[i] & kParticleDead) {
        c.pbinOf[i] = -1;
        return;
    }
    f2 p = to_f2(c.A.ppos[i]);
    int sx, sy;
    slot_of(p, w.W, w.H, sx, sy);
    int gb = w.pbinOff + part_bin_local(w, sx, sy);
    c.pbinOf[i] = gb;
    c.prank[i] = int(atomicAdd(&c.count[gb], 1u));
}

// exclusive scan of count[0..n) into start[0..n], clearing count
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

// exclusive scan across the block (blockDim multiple of 64, at most 1024); returns block total
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

__global__ void __launch_bounds__(kScanThreads) k_scan_reduce(const uint32_t* count, uint32_t* blockSums, int n) {
    size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanItems;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + k;
        if (i < size_t(n)) s += count[i];
    }
    uint32_t total;
    block_exclusive_scan(s, total);
    if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) k_scan_top(uint32_t* blockSums, int numTiles, uint32_t* start, int n) {
    const int per = (numTiles + int(blockDim.x) - 1) / int(blockDim.x);
    const int b0 = int(threadIdx.x) * per;
    uint32_t s = 0;
    for (int k = 0; k < per; ++k)
        if (b0 + k < numTiles) s += blockSums[b0 + k];
    uint32_t total;
    uint32_t off = block_exclusive_scan(s, total);
    for (int k = 0; k < per; ++k) {
        if (b0 + k < numTiles) {
            uint32_t t = blockSums[b0 + k];
            blockSums[b0 + k] = off;
            off += t;
        }
    }
    if (threadIdx.x == 0) start[n] = total;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_down(uint32_t* count, const uint32_t* blockSums,
                                                            uint32_t* start, int n) {
    size_t base = size_t(blockIdx.x) * kScanTile + size_t(threadIdx.x) * kScanItems;
    uint32_t v[kScanItems];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + k;
        v[k] = i < size_t(n) ? count[i] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan(s, total) + blockSums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        size_t i = base + k;
        if (i < size_t(n)) {
            start[i] = off;
            count[i] = 0u;
        }
        off += v[k];
    }
}

__global__ void k_cell_scatter(Ctx c) {
    int2 blk = c.cellBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int i = blk.y + int(threadIdx.x);
    if (i >= w.cellOff + w.numCells) return;
    int gb = c.binOf[i];
    int dest = w.cellOff + int(c.start[gb] - c.start[w.binOff]) + c.rank[i];
    c.sorted[dest] = i;
}

__global__ void k_part_scatter(Ctx c) {
    int2 blk = c.partBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int i = blk.y + int(threadIdx.x);
    if (i >= w.partOff + c.ctr[blk.x].numParticles) return;
    int gb = c.pbinOf[i];
    if (gb < 0) return;
    int dest = w.partOff + int(c.start[gb] - c.start[w.pbinOff]) + c.prank[i];
    c.psorted[dest] = i;
}

__global__ void k_part_count(Ctx c) {
    int wi = blockIdx.x * blockDim.x + threadIdx.x;
    if (wi >= c.numWorlds) return;
    const DevWorld w = c.worlds[wi];
    c.ctr[wi].numParticles = int(c.start[w.pbinOff + w.pnbx * w.pnby] - c.start[w.pbinOff]);
}

// sort the entries of one bin ascending (R1: stable order), small bins in registers
template <bool kWriteInverse>
__device__ __forceinline__ void sort_bin(int* arr, int s, int e, int* inverse) {
    const int n = e - s;
    if (n <= 0) return;
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

__global__ void k_cell_binsort(Ctx c) {
    int2 blk = c.binBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int gb = blk.y + int(threadIdx.x);
    if (gb >= w.binOff + w.nbx * w.nby) return;
    int s, e;
    cell_bin_range(c, w, gb - w.binOff, s, e);
    sort_bin<true>(c.sorted, s, e, c.newidx);
}

__global__ void k_part_binsort(Ctx c) {
    int2 blk = c.pbinBlocks[blockIdx.x];
    const DevWorld w = c.worlds[blk.x];
    int gb = blk.y + int(threadIdx.x);
    if (gb >= w.pbinOff + w.pnbx * w.pnby) return;
    int s, e;
    part_bin_range(c, w, gb - w.pbinOff, s, e);
    sort_bin<false>(c.psorted, s, e, nullptr);
}

// ---------------------------------------------------------------------------------------------
// neighbour pass (CellProcessor.cuh:122-365) + applyForces (:388-406) + Verlet of unbonded cells
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void push_candidate(const Ctx& c, int a, int b, int w) {
    int q = atomicAdd(&c.sc->candidates, 1);
    if (q < c.candCap)
        c.cand[q] = make_int4(a, b, w, 1);
    else
        atomicOr(&c.sc->errors, 1);
}

__global__ void __launch_bounds__(256) k_neighbor(Ctx c) {
    int2 blk = c.cellBlocks[blockIdx.x];
    const int wi = blk.x;
    const DevWorld w = c.worlds[wi];
    const int k = blk.y + int(threadIdx.x);  // new index
    if (k >= w.cellOff + w.numCells) return;
    const DevParams& P = c.params[wi];
    const int cap = c.cellCap;
    const float Wf = float(w.W), Hf = float(w.H);
    const int i = c.sorted[k];  // A index

    const f2 pi = to_f2(c.A.pos[i]);
    const f2 vi = to_f2(c.A.vel[i]);
    const uint32_t fi = c.A.flags[i];
    const int nci = flags_num_conn(fi);
    const bool bari = flags_barrier(fi);
    const int deti = flags_detached(fi);
    const float densi = c.A.dens[i];
    int bt[kMaxBonds];
#pragma unroll
    for (int b = 0; b < kMaxBonds; ++b) bt[b] = b < nci ? c.A.btgt[b * cap + i] : -1;
    const int iLocal = i - w.cellOff;
    (void)iLocal;

    int sx, sy;
    slot_of(pi, w.W, w.H, sx, sy);
    const int bx = min(sx / w.binSize, w.nbx - 1), by = min(sy / w.binSize, w.nby - 1);
    const bool fluid = P.motion == 0;
    const float cmd = P.minDistance;

    f2 F = {0.0f, 0.0f};
    f2 pc = pi;       // collision: corrected position (sequential adds)
    f2 D = {0.0f, 0.0f};  // fluid: accumulated overlap correction
    f2 Fp = {0.0f, 0.0f}, Fv = {0.0f, 0.0f};
    float rho = 0.0f;
    int barrierList[10];
    int numBarrier = 0;
    float energyI = -1.0f;  // loaded lazily for fusion checks

    const float h = P.smoothing;
    const float r = P.collisionDistance;

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
                const f2 pj = to_f2(c.A.pos[j]);
                const uint32_t fj = c.A.flags[j];
                const int jLocal = j - w.cellOff;
                if (!fluid) {
                    // collision mode (CellProcessor.cuh:285-365), seam-blind predicate (Map.cuh:189)
                    if (j == i) continue;
                    if (deti + flags_detached(fj) == 1) continue;
                    if (!(len2(sub(pj, pi)) <= r)) continue;
                    const f2 vj = to_f2(c.A.vel[j]);
                    const f2 pd = torus_dir(sub(pi, pj), Wf, Hf);
                    const float d = len2(pd);
                    if (!bari && d < cmd) pc = add(pc, dvd(mul(pd, cmd), 5.0f));
                    bool bonded = false;
#pragma unroll
                    for (int b = 0; b < kMaxBonds; ++b) bonded = bonded || (bt[b] == jLocal);
                    if (bonded) continue;
                    const f2 vd = sub(vi, vj);
                    const bool appr = dot2(pd, vd) < 0.0f;
                    const bool barj = flags_barrier(fj);
                    f2 fij, fji;
                    {
                        float bf = bari ? 2.0f : 1.0f;
                        if (len2(vi) > 0.5f && appr) {
                            float dsq = d * d + 0.25f;
                            fij = mul(dvd(mul(pd, dot2(vd, pd)), -2.0f * dsq), bf);
                        } else {
                            fij = mul(mul(mul(unit(pd), r - len2(pd)), P.repulsion), bf);
                        }
                    }
                    {
                        const f2 pdn = neg(pd), vdn = neg(vd);
                        float bf = barj ? 2.0f : 1.0f;
                        if (len2(vj) > 0.5f && appr) {
                            float dsq = d * d + 0.25f;
                            fji = mul(dvd(mul(pdn, dot2(vdn, pdn)), -2.0f * dsq), bf);
                        } else {
                            fji = mul(mul(mul(unit(pdn), r - len2(pdn)), P.repulsion), bf);
                        }
                    }
                    F = add(F, fij);
                    F = add(F, neg(fji));
                    if (k < m && nci < flags_max_conn(fi) && flags_num_conn(fj) < flags_max_conn(fj) &&
                        len2(vd) >= P.fusionVelocity && appr && !bari && !barj) {
                        if (energyI < 0.0f) energyI = c.energy[w.cellOff + c.A.slot[i]];
                        float energyJ = c.energy[w.cellOff + c.A.slot[j]];
                        if (energyI <= P.maxBindingEnergy && energyJ <= P.maxBindingEnergy) push_candidate(c, k, m, wi);
                    }
                } else {
                    // fluid mode (CellProcessor.cuh:122-283)
                    const f2 pd = torus_dir(sub(pi, pj), Wf, Hf);
                    const float d = len2(pd);
                    if (!(d <= h * 2.0f) || deti + flags_detached(fj) == 1) continue;
                    if (flags_barrier(fj)) {
                        if (numBarrier < 10) barrierList[numBarrier] = j;
                        ++numBarrier;
                        continue;
                    }
                    rho = rho + sph_w(d / h) / (h * h);
                    if (j == i) continue;
                    if (!bari && d < cmd) D = add(D, dvd(mul(pd, cmd), 5.0f));
                    bool bonded = false;
#pragma unroll
                    for (int b = 0; b < kMaxBonds; ++b) bonded = bonded || (bt[b] == jLocal);
                    if (bonded) continue;
                    const f2 vj = to_f2(c.A.vel[j]);
                    const float dj = c.A.dens[j];
                    const f2 vd = sub(vi, vj);
                    const float factor = densi / (densi * densi) + dj / (dj * dj);
                    if (fabsf(d) > kNearZero) {
                        const float kd = sph_dw(d / h) / (h * h * h);
                        Fp = add(Fp, mul(mul(dvd(pd, -d), factor), kd));
                        Fv = add(Fv, dvd(mul(mul(dvd(vd, dj), d), kd), d * d + 0.25f));
                    }
                    if (k < m && len2(vd) >= P.fusionVelocity && nci < flags_max_conn(fi) &&
                        flags_num_conn(fj) < flags_max_conn(fj) && !bari) {
                        if (energyI < 0.0f) energyI = c.energy[w.cellOff + c.A.slot[i]];
                        float energyJ = c.energy[w.cellOff + c.A.slot[j]];
                        if (energyI <= P.maxBindingEnergy && energyJ <= P.maxBindingEnergy) push_candidate(c, k, m, wi);
                    }
                }
            }
        }
    }

    if (fluid) {
        numBarrier = min(numBarrier, 10);
        if (numBarrier > 0) {  // barrier reflection, CellProcessor.cuh:234-274
            int cb = -1;
            float cbDist = 0.0f;
            for (int q = 0; q < numBarrier; ++q) {
                float d = len2(torus_dir(sub(pi, to_f2(c.A.pos[barrierList[q]])), Wf, Hf));
                if (cb < 0 || d < cbDist) {
                    cb = barrierList[q];
                    cbDist = d;
                }
            }
            const f2 pb = to_f2(c.A.pos[cb]);
            const f2 vb = to_f2(c.A.vel[cb]);
            const int nb = flags_num_conn(c.A.flags[cb]);
            f2 rr = {0.0f, 0.0f};
            if (nb <= 1) {
                rr = torus_dir(sub(pi, pb), Wf, Hf);
            } else {
                float angleToCell = angle_of(torus_dir(sub(pi, pb), Wf, Hf));
                for (int q = 0; q < nb; ++q) {
                    const f2 o1 = to_f2(c.A.pos[w.cellOff + c.A.btgt[q * cap + cb]]);
                    const f2 o2 = to_f2(c.A.pos[w.cellOff + c.A.btgt[((q + 1) % nb) * cap + cb]]);
                    float a1 = angle_of(torus_dir(sub(o1, pb), Wf, Hf));
                    float a2 = angle_of(torus_dir(sub(o2, pb), Wf, Hf));
                    if (angle_between(a1, a2, angleToCell)) {
                        rr = sub(o2, o1);
                        rot_ccw(rr);
                        break;
                    }
                }
            }
            const f2 vr = sub(vi, vb);
            const float dd = dot2(vr, rr);
            if (dd < 0.0f) {
                float trs = fmaxf(0.05f, lensq(rr));
                float td = fmaxf(0.05f, cbDist);
                F = add(F, dvd(sub(add(sub(vr, dvd(mul(mul(rr, 2.0f), dd), trs)), vb), vi), td));
            }
        }
        F = add(F, add(mul(Fp, P.pressure), mul(Fv, P.viscosity)));
        pc = add(pi, D);
    }

    // hot copy into the new order, bond targets renumbered
    c.B.flags[k] = fi;
    c.B.slot[k] = c.A.slot[i];
    c.B.dens[k] = fluid ? rho : densi;
    for (int b = 0; b < nci; ++b) {
        c.B.btgt[b * cap + k] = c.newidx[w.cellOff + bt[b]] - w.cellOff;
        c.B.brest[b * cap + k] = c.A.brest[b * cap + i];
        c.B.bang[b * cap + k] = c.A.bang[b * cap + i];
    }
    f2 v = vi;
    if (!bari) {  // applyForces
        v = add(v, F);
        if (len2(v) > P.maxVelocity) v = mul(unit(v), P.maxVelocity);
    }
    if (!bari && nci == 0) {
        // unbonded cell: its connection forces are zero, so the Verlet updates and friction of the
        // step (CellProcessor.cuh:525-559, :707-723) complete here with the same operations
        const f2 zero = {0.0f, 0.0f};
        const float dt = P.dt;
        f2 p2 = torus_pos(add(pc, add(mul(v, dt), dvd(mul(mul(zero, dt), dt), 2.0f))), w.W, w.H);
        f2 acc = dvd(add(zero, zero), 2.0f);
        v = add(v, mul(acc, dt));
        v = mul(v, 1.0f - P.friction);
        c.B.pos[k] = to_float2(p2);
        c.B.vel[k] = to_float2(v);
    } else {
        c.B.pos[k] = to_float2(pc);
        c.B.vel[k] = to_float2(v);
        int q = atomicAdd(&c.sc->staged, 1);
        c.staged[q] = make_int2(k, wi);
    }
}

// ---------------------------------------------------------------------------------------------
// energy particles (RadiationProcessor.cuh:61-170)
// ---------------------------------------------------------------------------------------------
__global__ void k_part_move(Ctx c) {
    int2 blk = c.partBlocks[blockIdx.x];
    const int wi = blk.x;
    const DevWorld w = c.worlds[wi];
    int k = blk.y + int(threadIdx.x);
    if (k >= w.partOff + c.ctr[wi].numParticles) return;
    const DevParams& P = c.params[wi];
    int src = c.psorted[k];
    f2 p = to_f2(c.A.ppos[src]);
    f2 v = to_f2(c.A.pvel[src]);
    p = torus_pos(add(p, mul(v, P.dt)), w.W, w.H);
    c.B.ppos[k] = to_float2(p);
    c.B.pvel[k] = to_float2(v);
    c.B.pen[k] = c.A.pen[src];
    c.B.plast[k] = c.A.plast[src];
    c.B.pslot[k] = c.A.pslot[src];
    c.B.pflags[k] = c.A.pflags[src];
}

__global__ void k_part_target(Ctx c) {
    int2 blk = c.partBlocks[blockIdx.x];
    const int wi = blk.x;
    const DevWorld w = c.worlds[wi];
    int k = blk.y + int(threadIdx.x);
    if (k >= w.partOff + c.ctr[wi].numParticles) return;
    const f2 p = to_f2(c.B.ppos[k]);
    int sx, sy;
    slot_of(p, w.W, w.H, sx, sy);
    // particle map winner of the slot (R2): first particle of the bin whose pre-move slot matches
    int s, e;
    part_bin_range(c, w, part_bin_local(w, sx, sy), s, e);
    int t = -1;
    for (int m = s; m < e; ++m) {
        int qx, qy;
        slot_of(to_f2(c.A.ppos[c.psorted[m]]), w.W, w.H, qx, qy);
        if (qx == sx && qy == sy) {
            t = m;
            break;
        }
    }
    int want = -1;
    if (t >= 0 && t != k && lensq(sub(p, to_f2(c.B.ppos[t]))) < 0.5f) want = t;
    c.want[k] = want;
    // cell hit by the particle this step: first cell of the slot at map-build time
    int cell = -1;
    if (want < 0) {
        const f2 q = add(p, to_f2(c.B.pvel[k]));
        int qx, qy;
        slot_of(q, w.W, w.H, qx, qy);
        int cs, ce;
        cell_bin_range(c, w, cell_bin_local(w, qx, qy), cs, ce);
        for (int m = cs; m < ce; ++m) {
            int ax, ay;
            slot_of(to_f2(c.A.pos[c.sorted[m]]), w.W, w.H, ax, ay);
            if (ax == qx && ay == qy) {
                cell = m;
                break;
            }
        }
    }
    c.absorbCell[k] = cell;
}

__global__ void k_part_claim(Ctx c) {
    int2 blk = c.partBlocks[blockIdx.x];
    const int wi = blk.x;
    const DevWorld w = c.worlds[wi];
    int k = blk.y + int(threadIdx.x);
    if (k >= w.partOff + c.ctr[wi].numParticles) return;
    int t = c.want[k];
    if (t >= 0 && c.want[t] < 0) atomicMin(&c.fclaim[t], k);
}

__global__ void k_part_apply(Ctx c) {
    int2 blk = c.partBlocks[blockIdx.x];
    const int wi = blk.x;
    const DevWorld w = c.worlds[wi];
    int k = blk.y + int(threadIdx.x);
    if (k >= w.partOff + c.ctr[wi].numParticles) return;
    const DevParams& P = c.params[wi];
    int q = c.fclaim[k];
    if (q != kClaimFree) {
        // sink particle: merge the lowest-index particle that fuses into it (R8)
        c.fclaim[k] = kClaimFree;
        float eq = c.B.pen[q], ek = c.B.pen[k];
        if (eq > kNearZero && ek > kNearZero) {
            float f1 = eq / (eq + ek);
            f2 vq = to_f2(c.B.pvel[q]), vk = to_f2(c.B.pvel[k]);
            c.B.pvel[k] = to_float2(add(mul(vq, f1), mul(vk, 1.0f - f1)));
            c.B.pen[k] = ek + eq;
            c.B.plast[k] = kNone;
            c.B.pen[q] = 0.0f;
            c.B.pflags[q] = uint8_t(c.B.pflags[q] | kParticleDead);
        }
        c.absorbCell[k] = -1;
        return;
    }
    if (c.want[k] >= 0) return;
    int cm = c.absorbCell[k];
    if (cm < 0) return;
    const uint32_t fc = c.B.flags[cm];
    if (flags_barrier(fc)) {
        const f2 vr = sub(to_f2(c.B.pvel[k]), to_f2(c.B.vel[cm]));
        const f2 rr = torus_dir(sub(to_f2(c.B.ppos[k]), to_f2(c.B.pos[cm])), float(w.W), float(w.H));
        float d = dot2(vr, rr);
        if (d < 0.0f) {
            float trs = fmaxf(0.1f, lensq(rr));
            c.B.pvel[k] = to_float2(add(sub(vr, dvd(mul(mul(rr, 2.0f), d), trs)), to_f2(c.B.vel[cm])));
        }
        c.absorbCell[k] = -1;
        return;
    }
    if (c.B.plast[k] == c.B.slot[cm] || P.absorption[flags_color(fc) % kColors] < kNearZero) {
        c.absorbCell[k] = -1;
        return;
    }
    atomicMin(&c.cclaim[cm], k);
}

__global__ void k_part_absorb(Ctx c) {
    int2 blk = c.partBlocks[blockIdx.x];
    const int wi = blk.x;
    const DevWorld w = c.worlds[wi];
    int k = blk.y + int(threadIdx.x);
    if (k >= w.partOff + c.ctr[wi].numParticles) return;
    int cm = c.absorbCell[k];
    if (cm < 0) return;
    if (c.cclaim[cm] != k) return;
    c.cclaim[cm] = kClaimFree;
    const DevParams& P = c.params[wi];
    const uint32_t fc = c.B.flags[cm];
    const float ra = P.absorption[flags_color(fc) % kColors];
    float e = c.B.pen[k];
    float transfer = e * ra;
    if (e < 0.01f) transfer = e;
    const int cslot = c.B.slot[cm];
    float* ce = &c.energy[w.cellOff + cslot];
    *ce = *ce + transfer;
    e = e - transfer;
    c.B.pen[k] = e;
    if (e < kNearZero)
        c.B.pflags[k] = uint8_t(c.B.pflags[k] | kParticleDead);
    else
        c.B.plast[k] = cslot;
}

// ---------------------------------------------------------------------------------------------
// connection forces (CellProcessor.cuh:408-491), gathered (R7), and the Verlet updates of bonded
// and barrier cells (CellProcessor.cuh:493-559, :707-723, :775-791)
// ---------------------------------------------------------------------------------------------
__global__ void k_conn_own(Ctx c, int pass) {
    const int n = c.sc->staged;
    const int cap = c.cellCap;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int2 ent = c.staged[q];
        const int x = ent.x, wi = ent.y;
        const DevWorld& w = c.worlds[wi];
        const DevParams& P = c.params[wi];
        const uint32_t fx = c.B.flags[x];
        const int nc = flags_num_conn(fx);
        if (nc == 0 || flags_barrier(fx)) {
            c.valid[x] = 0u;
            continue;
        }
        const float Wf = float(w.W), Hf = float(w.H);
        const bool considerAngles = (c.ctr[wi].timestep % 3ull) == 0ull;
        const f2 px = to_f2(c.B.pos[x]);
        const float sx = c.stiff[w.cellOff + c.B.slot[x]];
        const float kx2 = sx * sx;
        int tg[kMaxBonds];
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b) tg[b] = b < nc ? w.cellOff + c.B.btgt[b * cap + x] : -1;
        f2 prevDisp = torus_dir(sub(to_f2(c.B.pos[tg[nc - 1]]), px), Wf, Hf);
        f2 force = {0.0f, 0.0f};
        uint32_t v = 0u;
        bool tooLong = false;
        const float maxBind = P.maxBindingDistance[flags_color(fx) % kColors];
        for (int b = 0; b < nc; ++b) {
            const int o = tg[b];
            const float so = c.stiff[w.cellOff + c.B.slot[o]];
            const float kc2 = so * so;
            const f2 disp = torus_dir(sub(to_f2(c.B.pos[o]), px), Wf, Hf);
            const float ad = len2(disp);
            if (pass == 1 && ad > maxBind) tooLong = true;  // checkConnections on new positions
            const float dev = ad - c.B.brest[b * cap + x];
            force = add(force, dvd(mul(mul(unit(disp), dev), kx2 + kc2), 6.0f));
            if (considerAngles && (nc > 2 || (nc == 2 && b == 0))) {
                const int last = (b + nc - 1) % nc;
                const int lc = tg[last];
                const uint32_t fo = c.B.flags[o];
                const int no = flags_num_conn(fo);
                bool tri = false;
                for (int j = 0; j < no; ++j)
                    if (w.cellOff + c.B.btgt[j * cap + o] == lc) {
                        tri = true;
                        break;
                    }
                if (!tri) {
                    const float angle = angle_of(disp);
                    const float prevAngle = angle_of(prevDisp);
                    const float act = sub_angle(angle, prevAngle);
                    if (act < 0.0f) continue;
                    const float ref = c.B.bang[b * cap + x];
                    const float strength = fabsf(ref - act) / 2000.0f * kx2;
                    f2 f1 = mul(dvd(unit(disp), fmaxf(len2(disp), P.minDistance)), strength);
                    rot_cw(f1);
                    f2 fb = mul(dvd(unit(prevDisp), fmaxf(len2(prevDisp), P.minDistance)), strength);
                    rot_ccw(fb);
                    if (ref < act) {
                        f1 = mul(f1, -1.0f);
                        fb = mul(fb, -1.0f);
                    }
                    if (!flags_barrier(fo)) {
                        c.out1[b * cap + x] = to_float2(f1);
                        v |= 1u << b;
                    }
                    if (!flags_barrier(c.B.flags[lc])) {
                        c.out2[b * cap + x] = to_float2(fb);
                        v |= 1u << (6 + b);
                    }
                    force = sub(force, add(f1, fb));
                }
            }
            prevDisp = disp;
        }
        c.own[x] = to_float2(force);
        c.valid[x] = v;
        if (tooLong) c.delFlag[x] = 1;
    }
}

__device__ __forceinline__ f2 gather_conn(const Ctx& c, const DevWorld& w, int x, int nc) {
    const int cap = c.cellCap;
    f2 acc = {0.0f, 0.0f};
    acc = add(acc, to_f2(c.own[x]));
    const int xl = x - w.cellOff;
    for (int q = 0; q < nc; ++q) {
        const int p = w.cellOff + c.B.btgt[q * cap + x];
        const int np = flags_num_conn(c.B.flags[p]);
        int b = -1;
        for (int kk = 0; kk < np; ++kk)
            if (c.B.btgt[kk * cap + p] == xl) {
                b = kk;
                break;
            }
        if (b < 0) continue;
        const uint32_t vp = c.valid[p];
        if (vp & (1u << b)) acc = add(acc, to_f2(c.out1[b * cap + p]));
        const int b2 = (b + 1) % np;
        if (vp & (1u << (6 + b2))) acc = add(acc, to_f2(c.out2[b2 * cap + p]));
    }
    return acc;
}

__global__ void k_conn_verlet_pos(Ctx c) {
    const int n = c.sc->staged;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int2 ent = c.staged[q];
        const int x = ent.x, wi = ent.y;
        const DevWorld& w = c.worlds[wi];
        const DevParams& P = c.params[wi];
        const uint32_t fx = c.B.flags[x];
        const float dt = P.dt;
        const f2 p = to_f2(c.B.pos[x]);
        const f2 v = to_f2(c.B.vel[x]);
        if (flags_barrier(fx)) {
            c.B.pos[x] = to_float2(torus_pos(add(p, mul(v, dt)), w.W, w.H));
            continue;
        }
        const f2 s1 = gather_conn(c, w, x, flags_num_conn(fx));
        c.B.pos[x] = to_float2(torus_pos(add(p, add(mul(v, dt), dvd(mul(mul(s1, dt), dt), 2.0f))), w.W, w.H));
        c.s2[x] = to_float2(s1);
    }
}

__global__ void k_conn_verlet_vel(Ctx c) {
    const int n = c.sc->staged;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int2 ent = c.staged[q];
        const int x = ent.x, wi = ent.y;
        const DevWorld& w = c.worlds[wi];
        const DevParams& P = c.params[wi];
        const uint32_t fx = c.B.flags[x];
        if (flags_barrier(fx)) continue;
        const f2 s1 = gather_conn(c, w, x, flags_num_conn(fx));
        const f2 acc = dvd(add(s1, to_f2(c.s2[x])), 2.0f);
        f2 v = to_f2(c.B.vel[x]);
        v = add(v, mul(acc, P.dt));
        v = mul(v, 1.0f - P.friction);
        c.B.vel[x] = to_float2(v);
        if (c.energy[w.cellOff + c.B.slot[x]] > P.maxBindingEnergy) c.delFlag[x] = 1;  // decay, :786-790
    }
}

// ---------------------------------------------------------------------------------------------
// structural operations (CellConnectionProcessor.cuh), R5/R6
// ---------------------------------------------------------------------------------------------
__global__ void k_mark(Ctx c) {
    const int n = c.sc->staged;
    const int cap = c.cellCap;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int2 ent = c.staged[q];
        const int x = ent.x;
        const DevWorld& w = c.worlds[ent.y];
        uint8_t m = c.delFlag[x];
        const int nc = flags_num_conn(c.B.flags[x]);
        for (int b = 0; b < nc && !m; ++b)
            if (c.delFlag[w.cellOff + c.B.btgt[b * cap + x]]) m = 1;
        c.mark[x] = m;
    }
}

template <class F> __device__ __forceinline__ void for_claim_set(const Ctx& c, const DevWorld& w, int a, int b, F&& f) {
    const int cap = c.cellCap;
    f(a);
    int na = flags_num_conn(c.B.flags[a]);
    for (int k = 0; k < na; ++k) f(w.cellOff + c.B.btgt[k * cap + a]);
    f(b);
    int nb = flags_num_conn(c.B.flags[b]);
    for (int k = 0; k < nb; ++k) f(w.cellOff + c.B.btgt[k * cap + b]);
}

__device__ __forceinline__ unsigned long long claim_key(const DevWorld& w, int a, int b) {
    return (static_cast<unsigned long long>(uint32_t(a - w.cellOff)) << 32) | uint32_t(b - w.cellOff);
}

__global__ void k_fuse_claim(Ctx c) {
    const int n = min(c.sc->candidates, c.candCap);
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int4 cd = c.cand[q];
        const DevWorld& w = c.worlds[cd.z];
        if (c.mark[cd.x] || c.mark[cd.y]) {
            c.cand[q].w = 0;
            continue;
        }
        unsigned long long key = claim_key(w, cd.x, cd.y);
        for_claim_set(c, w, cd.x, cd.y, [&](int x) { atomicMin(&c.claim[x], key); });
    }
}

// tryAddConnectionOneWay (CellConnectionProcessor.cuh:274-427) without angle alignment
__device__ bool would_overlap(const Ctx& c, const DevWorld& w, int c1, f2 otherPos) {
    const int cap = c.cellCap;
    const int n = flags_num_conn(c.B.flags[c1]);
    if (n < 2) return false;
    const f2 p1 = to_f2(c.B.pos[c1]);
    for (int i = 0; i < n; ++i) {
        const int cc = w.cellOff + c.B.btgt[i * cap + c1];
        const int nc = w.cellOff + c.B.btgt[((i + 1) % n) * cap + c1];
        const int ncl = nc - w.cellOff;
        const int ncc = flags_num_conn(c.B.flags[cc]);
        bool both = false;
        for (int j = 0; j < ncc; ++j)
            if (c.B.btgt[j * cap + cc] == ncl) {
                both = true;
                break;
            }
        if (!both) continue;
        if (segments_cross(p1, otherPos, to_f2(c.B.pos[cc]), to_f2(c.B.pos[nc]))) return true;
    }
    return false;
}

__device__ bool add_one_way(const Ctx& c, const DevWorld& w, int c1, int c2, f2 posDelta) {
    const int cap = c.cellCap;
    const float Wf = float(w.W), Hf = float(w.H);
    uint32_t f = c.B.flags[c1];
    const int n = flags_num_conn(f);
    if (n == kMaxBonds) return false;
    if (would_overlap(c, w, c1, to_f2(c.B.pos[c2]))) return false;
    const float newAngle = angle_of(posDelta);
    const float dist = len2(posDelta);
    const int c2l = c2 - w.cellOff;
    const f2 p1 = to_f2(c.B.pos[c1]);
    if (n == 0) {
        c.B.btgt[c1] = c2l;
        c.B.brest[c1] = dist;
        c.B.bang[c1] = 360.0f;
        c.B.flags[c1] = flags_with_num_conn(f, 1);
        return true;
    }
    if (n == 1) {
        const f2 cd = torus_dir(sub(to_f2(c.B.pos[w.cellOff + c.B.btgt[c1]]), p1), Wf, Hf);
        const float prevAngle = angle_of(cd);
        float angleDiff = newAngle - prevAngle;
        if (angleDiff < 0.0f) angleDiff += 360.0f;
        if (fabsf(angleDiff) < kNearZero || fabsf(angleDiff - 360.0f) < kNearZero ||
            fabsf(angleDiff + 360.0f) < kNearZero)
            return false;
        c.B.bang[cap + c1] = angleDiff;
        c.B.bang[c1] = 360.0f - angleDiff;
        c.B.btgt[cap + c1] = c2l;
        c.B.brest[cap + c1] = dist;
        c.B.flags[c1] = flags_with_num_conn(f, 2);
        return true;
    }
    int index = 0;
    float prevAngle = 0.0f, nextAngle = 0.0f;
    bool found = false;
    for (; index < n; ++index) {
        const int prevIndex = (index + n - 1) % n;
        prevAngle = angle_of(torus_dir(sub(to_f2(c.B.pos[w.cellOff + c.B.btgt[prevIndex * cap + c1]]), p1), Wf, Hf));
        nextAngle = angle_of(torus_dir(sub(to_f2(c.B.pos[w.cellOff + c.B.btgt[index * cap + c1]]), p1), Wf, Hf));
        if (angle_between(prevAngle, nextAngle, newAngle) || prevIndex == index) {
            found = true;
            break;
        }
    }
    if (!found) return false;
    const float refAngle = c.B.bang[index * cap + c1];
    const float angleDiff1 = sub_angle(newAngle, prevAngle);
    const float angleDiff2 = sub_angle(nextAngle, prevAngle);
    const float frac = angleDiff2 != 0.0f ? angleDiff1 / angleDiff2 : 0.5f;
    float afp = refAngle * frac;
    afp = fminf(afp, refAngle);
    if (afp < kNearZero) return false;
    if (index == 0) index = n;
    for (int j = n; j > index; --j) {
        c.B.btgt[j * cap + c1] = c.B.btgt[(j - 1) * cap + c1];
        c.B.brest[j * cap + c1] = c.B.brest[(j - 1) * cap + c1];
        c.B.bang[j * cap + c1] = c.B.bang[(j - 1) * cap + c1];
    }
    c.B.btgt[index * cap + c1] = c2l;
    c.B.brest[index * cap + c1] = dist;
    c.B.bang[index * cap + c1] = afp;
    c.B.bang[((index + 1) % (n + 1)) * cap + c1] = refAngle - afp;
    c.B.flags[c1] = flags_with_num_conn(f, n + 1);
    return true;
}

__global__ void k_fuse_apply(Ctx c) {
    const int n = min(c.sc->candidates, c.candCap);
    const int cap = c.cellCap;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int4 cd = c.cand[q];
        if (!cd.w) continue;
        const DevWorld& w = c.worlds[cd.z];
        const int a = cd.x, b = cd.y;
        unsigned long long key = claim_key(w, a, b);
        bool ok = true;
        for_claim_set(c, w, a, b, [&](int x) { ok = ok && (c.claim[x] == key); });
        if (!ok) continue;
        // lockAndtryAddConnections (CellConnectionProcessor.cuh:249-272)
        const uint32_t fa = c.B.flags[a], fb = c.B.flags[b];
        const int na = flags_num_conn(fa), nb = flags_num_conn(fb);
        bool already = false;
        for (int k = 0; k < na; ++k) already = already || (c.B.btgt[k * cap + a] == b - w.cellOff);
        if (already) continue;
        if (!(na < flags_max_conn(fa) && nb < flags_max_conn(fb))) continue;
        // tryAddConnections (:211-240)
        const f2 posDelta = torus_dir(sub(to_f2(c.B.pos[b]), to_f2(c.B.pos[a])), float(w.W), float(w.H));
        int st[kMaxBonds];
        float sr[kMaxBonds], sa[kMaxBonds];
        for (int k = 0; k < kMaxBonds; ++k) {
            st[k] = c.B.btgt[k * cap + a];
            sr[k] = c.B.brest[k * cap + a];
            sa[k] = c.B.bang[k * cap + a];
        }
        if (!add_one_way(c, w, a, b, posDelta)) continue;
        if (!add_one_way(c, w, b, a, mul(posDelta, -1.0f))) {
            c.B.flags[a] = fa;
            for (int k = 0; k < kMaxBonds; ++k) {
                c.B.btgt[k * cap + a] = st[k];
                c.B.brest[k * cap + a] = sr[k];
                c.B.bang[k * cap + a] = sa[k];
            }
        }
    }
}

__global__ void k_fuse_reset(Ctx c) {
    const int n = min(c.sc->candidates, c.candCap);
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        int4 cd = c.cand[q];
        if (!cd.w) continue;
        const DevWorld& w = c.worlds[cd.z];
        for_claim_set(c, w, cd.x, cd.y, [&](int x) { c.claim[x] = ~0ull; });
    }
}

// deleteConnectionOneWay (CellConnectionProcessor.cuh:429-448)
__device__ void delete_one_way(const Ctx& c, int c1, int targetLocal) {
    const int cap = c.cellCap;
    uint32_t f = c.B.flags[c1];
    const int n = flags_num_conn(f);
    for (int i = 0; i < n; ++i) {
        if (c.B.btgt[i * cap + c1] == targetLocal) {
            const float angleToAdd = c.B.bang[i * cap + c1];
            for (int j = i; j < n - 1; ++j) {
                c.B.btgt[j * cap + c1] = c.B.btgt[(j + 1) * cap + c1];
                c.B.brest[j * cap + c1] = c.B.brest[(j + 1) * cap + c1];
                c.B.bang[j * cap + c1] = c.B.bang[(j + 1) * cap + c1];
            }
            if (i < n - 1)
                c.B.bang[i * cap + c1] += angleToAdd;
            else
                c.B.bang[c1] += angleToAdd;
            c.B.flags[c1] = flags_with_num_conn(f, n - 1);
            return;
        }
    }
}

__global__ void k_delete(Ctx c) {
    const int n = c.sc->staged;
    const int cap = c.cellCap;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int2 ent = c.staged[q];
        const int x = ent.x;
        if (!c.mark[x]) continue;
        const DevWorld& w = c.worlds[ent.y];
        const int nc = flags_num_conn(c.B.flags[x]);
        int snap[kMaxBonds];
        for (int b = 0; b < nc; ++b) snap[b] = c.B.btgt[b * cap + x];
        const bool self = c.delFlag[x] != 0;
        for (int b = 0; b < nc; ++b)
            if (self || c.delFlag[w.cellOff + snap[b]]) delete_one_way(c, x, snap[b]);
    }
}

__global__ void k_clear_flags(Ctx c) {
    const int n = c.sc->staged;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int x = c.staged[q].x;
        c.mark[x] = 0;
        c.delFlag[x] = 0;
    }
}

// ---------------------------------------------------------------------------------------------
// introspection: neighbour lists under the step predicate (same visiting order as k_neighbor)
// ---------------------------------------------------------------------------------------------
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
                    nb = len2(torus_dir(sub(pi, pj), float(w.W), float(w.H))) <= P.smoothing * 2.0f;
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

// ---------------------------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------------------------
#define ALIEN_LAUNCH(kernel, grid, block, stream, ...) \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, stream, __VA_ARGS__)

static int persistent_grid(int n, int block) {
    int g = (n + block - 1) / block;
    if (g > 2048) g = 2048;
    if (g < 1) g = 1;
    return g;
}

void launch_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return;
    size_t g = (n + 255) / 256;
    if (g > 4096) g = 4096;
    ALIEN_LAUNCH(k_fill_u32, int(g), 256, s, p, v, n);
}

void launch_reset_density(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    if (sh.cellBlocks) ALIEN_LAUNCH(k_reset_density, sh.cellBlocks, sh.cellThreads, s, c);
}

void launch_sort(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    if (sh.cellBlocks) ALIEN_LAUNCH(k_cell_bin, sh.cellBlocks, sh.cellThreads, s, c);
    if (sh.partBlocks) ALIEN_LAUNCH(k_part_bin, sh.partBlocks, sh.partThreads, s, c);
    const int tiles = (c.numBins + kScanTile - 1) / kScanTile;
    ALIEN_LAUNCH(k_scan_reduce, tiles, kScanThreads, s, c.count, c.blockSums, c.numBins);
    ALIEN_LAUNCH(k_scan_top, 1, 1024, s, c.blockSums, tiles, c.start, c.numBins);
    ALIEN_LAUNCH(k_scan_down, tiles, kScanThreads, s, c.count, c.blockSums, c.start, c.numBins);
    if (sh.cellBlocks) ALIEN_LAUNCH(k_cell_scatter, sh.cellBlocks, sh.cellThreads, s, c);
    if (sh.partBlocks) ALIEN_LAUNCH(k_part_scatter, sh.partBlocks, sh.partThreads, s, c);
    ALIEN_LAUNCH(k_part_count, (c.numWorlds + 63) / 64, 64, s, c);
    if (sh.binBlocks) ALIEN_LAUNCH(k_cell_binsort, sh.binBlocks, sh.binThreads, s, c);
    if (sh.pbinBlocks) ALIEN_LAUNCH(k_part_binsort, sh.pbinBlocks, sh.binThreads, s, c);
}

void launch_step(const Ctx& c, const LaunchShape& sh, hipStream_t s) {
    ALIEN_LAUNCH(k_step_begin, 1, 64, s, c);
    launch_sort(c, sh, s);
    if (sh.cellBlocks) ALIEN_LAUNCH(k_neighbor, sh.cellBlocks, sh.cellThreads, s, c);
    if (sh.partBlocks) {
        ALIEN_LAUNCH(k_part_move, sh.partBlocks, sh.partThreads, s, c);
        ALIEN_LAUNCH(k_part_target, sh.partBlocks, sh.partThreads, s, c);
        ALIEN_LAUNCH(k_part_claim, sh.partBlocks, sh.partThreads, s, c);
        ALIEN_LAUNCH(k_part_apply, sh.partBlocks, sh.partThreads, s, c);
        ALIEN_LAUNCH(k_part_absorb, sh.partBlocks, sh.partThreads, s, c);
    }
    const int sg = persistent_grid(c.cellCap, 256);
    ALIEN_LAUNCH(k_conn_own, sg, 256, s, c, 0);
    ALIEN_LAUNCH(k_conn_verlet_pos, sg, 256, s, c);
    ALIEN_LAUNCH(k_conn_own, sg, 256, s, c, 1);
    ALIEN_LAUNCH(k_conn_verlet_vel, sg, 256, s, c);
    ALIEN_LAUNCH(k_mark, sg, 256, s, c);
    const int cg = persistent_grid(c.candCap, 256);
    ALIEN_LAUNCH(k_fuse_claim, cg, 256, s, c);
    ALIEN_LAUNCH(k_fuse_apply, cg, 256, s, c);
    ALIEN_LAUNCH(k_fuse_reset, cg, 256, s, c);
    ALIEN_LAUNCH(k_delete, sg, 256, s, c);
    ALIEN_LAUNCH(k_clear_flags, sg, 256, s, c);
    ALIEN_LAUNCH(k_step_end, (c.numWorlds + 63) / 64, 64, s, c);
}

void launch_debug_neighbors(const Ctx& c, const LaunchShape& sh, const uint64_t* offsets, int* out, int* counts,
                            hipStream_t s) {
    if (sh.cellBlocks) ALIEN_LAUNCH(k_debug_neighbors, sh.cellBlocks, sh.cellThreads, s, c, offsets, out, counts);
}

}  // namespace alien
