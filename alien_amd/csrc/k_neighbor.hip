// k_neighbor.hip — the neighbour pass of the physics timestep.
//
// ALIEN: calcCollisions_reconnectCells_correctOverlap / calcFluidForces_reconnectCells_correctOverlap
// (source/EngineGpuKernels/CellProcessor.cuh:122-365), then checkForces/applyForces (:367-406); for
// cells without bonds also the Verlet updates and friction of the same step (:525-559, :707-723),
// which reduce to local operations when the connection forces are zero.
//
// One workgroup owns a tile of TX x TY bins.  It stages the cells of the tile and of its one-bin
// halo into LDS, one thread per bin, placing every entry at its rank by previous index inside the
// bin (the stable in-bin order of the cell map, R1, without a separate sort pass).  Then every
// thread walks the 3x3 bin stencil of one interior cell in LDS.  A tile whose halo does not fit
// the LDS budget walks the same stencil in the same order straight from global memory.  The kernel
// is instantiated per motion mode; a workgroup of a world in the other mode exits at once.
#include "kernels_common.hpp"
#include "kernels.hpp"

namespace alien {

struct Nb {
    f2 pos, vel;
    float dens;
    uint32_t flags;
    int aidx;  // index in the start-of-step arrays (A)
    int lidx;  // index in the end-of-step arrays (B), the logical index of this step
};

__device__ __forceinline__ void push_candidate(const Ctx& c, int a, int b, int w) {
    int q = atomicAdd(&c.sc->candidates, 1);
    if (q < c.candCap)
        c.cand[q] = make_int4(a, b, w, 1);
    else
        atomicOr(&c.sc->errors, 1);
}

// Physics of one cell (self = its own start-of-step data) against its neighbours, visited in
// stencil order by forEachNb.
template <bool kFluid, class ForEach>
__device__ __forceinline__ void cell_update(const Ctx& c, const DevWorld& w, const DevParams& P, int wi, const Nb& self,
                                            ForEach&& forEachNb) {
    const int cap = c.cellCap;
    const float Wf = float(w.W), Hf = float(w.H);
    const int i = self.aidx, k = self.lidx;
    const f2 pi = self.pos;
    const f2 vi = self.vel;
    const uint32_t fi = self.flags;
    const int nci = flags_num_conn(fi);
    const bool bari = flags_barrier(fi);
    const int deti = flags_detached(fi);
    const float densi = self.dens;
    int bt[kMaxBonds];
#pragma unroll
    for (int b = 0; b < kMaxBonds; ++b) bt[b] = b < nci ? c.A.btgt[b * cap + i] : -1;

    const float cmd = P.minDistance;
    const float h = P.smoothing;
    const float r = P.collisionDistance;

    f2 F = {0.0f, 0.0f};
    f2 pc = pi;           // collision: corrected position (sequential adds, CellProcessor.cuh:304-308)
    f2 D = {0.0f, 0.0f};  // fluid: summed overlap correction (:187-191, :276)
    f2 Fp = {0.0f, 0.0f}, Fv = {0.0f, 0.0f};
    float rho = 0.0f;
    int barrierList[10];
    int numBarrier = 0;
    float energyI = -1.0f;

    forEachNb([&](const Nb& o) {
        const int j = o.aidx;
        const int jLocal = j - w.cellOff;
        if constexpr (!kFluid) {
            // collision mode (CellProcessor.cuh:285-365), seam-blind neighbour predicate (Map.cuh:189)
            if (j == i) return;
            if (deti + flags_detached(o.flags) == 1) return;
            if (!(len2(sub(o.pos, pi)) <= r)) return;
            const f2 pd = torus_dir(sub(pi, o.pos), Wf, Hf);
            const float d = len2(pd);
            if (!bari && d < cmd) pc = add(pc, dvd(mul(pd, cmd), 5.0f));
            bool bonded = false;
#pragma unroll
            for (int b = 0; b < kMaxBonds; ++b) bonded = bonded || (bt[b] == jLocal);
            if (bonded) return;
            const f2 vd = sub(vi, o.vel);
            const bool appr = dot2(pd, vd) < 0.0f;
            const bool barj = flags_barrier(o.flags);
            f2 fij, fji;
            {
                const float bf = bari ? 2.0f : 1.0f;
                if (len2(vi) > 0.5f && appr) {
                    const float dsq = d * d + 0.25f;
                    fij = mul(dvd(mul(pd, dot2(vd, pd)), -2.0f * dsq), bf);
                } else {
                    fij = mul(mul(mul(unit(pd), r - len2(pd)), P.repulsion), bf);
                }
            }
            {
                const f2 pdn = neg(pd), vdn = neg(vd);
                const float bf = barj ? 2.0f : 1.0f;
                if (len2(o.vel) > 0.5f && appr) {
                    const float dsq = d * d + 0.25f;
                    fji = mul(dvd(mul(pdn, dot2(vdn, pdn)), -2.0f * dsq), bf);
                } else {
                    fji = mul(mul(mul(unit(pdn), r - len2(pdn)), P.repulsion), bf);
                }
            }
            F = add(F, fij);
            F = add(F, neg(fji));
            if (k < o.lidx && nci < flags_max_conn(fi) && flags_num_conn(o.flags) < flags_max_conn(o.flags) &&
                len2(vd) >= P.fusionVelocity && appr && !bari && !barj) {
                if (energyI < 0.0f) energyI = c.energy[w.cellOff + c.A.slot[i]];
                const float energyJ = c.energy[w.cellOff + c.A.slot[j]];
                if (energyI <= P.maxBindingEnergy && energyJ <= P.maxBindingEnergy) push_candidate(c, k, o.lidx, wi);
            }
        } else {
            // fluid mode (CellProcessor.cuh:122-283)
            const f2 pd = torus_dir(sub(pi, o.pos), Wf, Hf);
            const float d = len2(pd);
            if (!(d <= h * 2.0f) || deti + flags_detached(o.flags) == 1) return;
            if (flags_barrier(o.flags)) {
                if (numBarrier < 10) barrierList[numBarrier] = j;
                ++numBarrier;
                return;
            }
            rho = rho + sph_w(d / h) / (h * h);
            if (j == i) return;
            if (!bari && d < cmd) D = add(D, dvd(mul(pd, cmd), 5.0f));
            bool bonded = false;
#pragma unroll
            for (int b = 0; b < kMaxBonds; ++b) bonded = bonded || (bt[b] == jLocal);
            if (bonded) return;
            const float dj = o.dens;
            const f2 vd = sub(vi, o.vel);
            const float factor = densi / (densi * densi) + dj / (dj * dj);
            if (fabsf(d) > kNearZero) {
                const float kd = sph_dw(d / h) / (h * h * h);
                Fp = add(Fp, mul(mul(dvd(pd, -d), factor), kd));
                Fv = add(Fv, dvd(mul(mul(dvd(vd, dj), d), kd), d * d + 0.25f));
            }
            if (k < o.lidx && len2(vd) >= P.fusionVelocity && nci < flags_max_conn(fi) &&
                flags_num_conn(o.flags) < flags_max_conn(o.flags) && !bari) {
                if (energyI < 0.0f) energyI = c.energy[w.cellOff + c.A.slot[i]];
                const float energyJ = c.energy[w.cellOff + c.A.slot[j]];
                if (energyI <= P.maxBindingEnergy && energyJ <= P.maxBindingEnergy) push_candidate(c, k, o.lidx, wi);
            }
        }
    });

    if constexpr (kFluid) {
        numBarrier = min(numBarrier, 10);
        if (numBarrier > 0) {  // barrier reflection, CellProcessor.cuh:234-274
            int cb = -1;
            float cbDist = 0.0f;
            for (int q = 0; q < numBarrier; ++q) {
                const float d = len2(torus_dir(sub(pi, to_f2(c.A.pos[barrierList[q]])), Wf, Hf));
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
                const float angleToCell = angle_of(torus_dir(sub(pi, pb), Wf, Hf));
                for (int q = 0; q < nb; ++q) {
                    const f2 o1 = to_f2(c.A.pos[w.cellOff + c.A.btgt[q * cap + cb]]);
                    const f2 o2 = to_f2(c.A.pos[w.cellOff + c.A.btgt[((q + 1) % nb) * cap + cb]]);
                    const float a1 = angle_of(torus_dir(sub(o1, pb), Wf, Hf));
                    const float a2 = angle_of(torus_dir(sub(o2, pb), Wf, Hf));
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
                const float trs = fmaxf(0.05f, lensq(rr));
                const float td = fmaxf(0.05f, cbDist);
                F = add(F, dvd(sub(add(sub(vr, dvd(mul(mul(rr, 2.0f), dd), trs)), vb), vi), td));
            }
        }
        F = add(F, add(mul(Fp, P.pressure), mul(Fv, P.viscosity)));
        pc = add(pi, D);
    }

    // the cell in the end-of-step order; bond targets stay in start-of-step indices until the
    // staged cells are renumbered (k_remap)
    c.B.flags[k] = fi;
    c.B.slot[k] = c.A.slot[i];
    c.B.dens[k] = kFluid ? rho : densi;
    for (int b = 0; b < nci; ++b) {
        c.B.btgt[b * cap + k] = bt[b];
        c.B.brest[b * cap + k] = c.A.brest[b * cap + i];
        c.B.bang[b * cap + k] = c.A.bang[b * cap + i];
    }
    if (nci > 0) c.newidx[i] = k;
    f2 v = vi;
    if (!bari) {  // applyForces
        v = add(v, F);
        if (len2(v) > P.maxVelocity) v = mul(unit(v), P.maxVelocity);
    }
    if (!bari && nci == 0) {
        // no bonds: both connection-force passes yield zero, so the Verlet position and velocity
        // updates and friction of this step are applied here with the same operations
        const f2 zero = {0.0f, 0.0f};
        const float dt = P.dt;
        const f2 p2 = torus_pos(add(pc, add(mul(v, dt), dvd(mul(mul(zero, dt), dt), 2.0f))), w.W, w.H);
        const f2 acc = dvd(add(zero, zero), 2.0f);
        v = add(v, mul(acc, dt));
        v = mul(v, 1.0f - P.friction);
        c.B.pos[k] = to_float2(p2);
        c.B.vel[k] = to_float2(v);
        bin_for_next_step(c, w, k, p2, c.countNext);
    } else {
        c.B.pos[k] = to_float2(pc);
        c.B.vel[k] = to_float2(v);
        const int q = atomicAdd(&c.sc->staged, 1);
        c.staged[q] = make_int2(k, wi);
    }
}

struct TileShared {
    float2 pos[kTileCap];
    float2 vel[kTileCap];
    float dens[kTileCap];
    uint32_t flags[kTileCap];
    int aidx[kTileCap];
    int lidx[kTileCap];
    int hbStart[kMaxHaloBins + 1];  // LDS offset of every halo bin (exclusive scan)
    int hbGlob[kMaxHaloBins];       // global position of the bin's first entry in `sorted`
};

__device__ __forceinline__ Nb lds_nb(const TileShared& S, int s) {
    Nb o;
    o.pos = to_f2(S.pos[s]);
    o.vel = to_f2(S.vel[s]);
    o.dens = S.dens[s];
    o.flags = S.flags[s];
    o.aidx = S.aidx[s];
    o.lidx = S.lidx[s];
    return o;
}

__device__ __forceinline__ Nb global_nb(const Ctx& c, int a, int l) {
    Nb o;
    o.pos = to_f2(c.A.pos[a]);
    o.vel = to_f2(c.A.vel[a]);
    o.dens = c.A.dens[a];
    o.flags = c.A.flags[a];
    o.aidx = a;
    o.lidx = l;
    return o;
}

template <bool kFluid>
__global__ void __launch_bounds__(kTileThreads) k_neighbor_tile(Ctx c) {
    __shared__ TileShared S;
    const int2 tb = c.tileBlocks[blockIdx.x];
    if (tb.x < 0) return;
    const int wi = tb.x;
    const DevWorld& w = c.worlds[wi];
    const DevParams& P = c.params[wi];
    if ((P.motion == 0) != kFluid) return;
    const int tx = tb.y % w.ntx, ty = tb.y / w.ntx;
    const int x0 = tx * w.tileX, y0 = ty * w.tileY;
    const int IX = min(w.tileX, w.nbx - x0), IY = min(w.tileY, w.nby - y0);
    const int HX = IX + 2, HY = IY + 2, nh = HX * HY;
    const uint32_t base = c.start[w.binOff];
    const int tid = threadIdx.x;
    constexpr int kPer = kMaxHaloBins / kTileThreads;

    // halo bins: entry counts and LDS offsets
    int cnt[kPer];
    uint32_t local = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int hb = tid * kPer + q;
        int n = 0;
        if (hb < nh) {
            const int rr = hb / HX, cc = hb - rr * HX;
            const int gx = wrap_small(x0 - 1 + cc, w.nbx), gy = wrap_small(y0 - 1 + rr, w.nby);
            const int gb = w.binOff + gy * w.nbx + gx;
            const uint32_t s = c.start[gb];
            n = int(c.start[gb + 1] - s);
            S.hbGlob[hb] = w.cellOff + int(s - base);
        }
        cnt[q] = n;
        local += uint32_t(n);
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan(local, total);
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int hb = tid * kPer + q;
        if (hb < nh) S.hbStart[hb] = int(off);
        off += uint32_t(cnt[q]);
    }
    if (tid == 0) S.hbStart[nh] = int(total);
    const bool staged = total <= uint32_t(kTileCap);
    __syncthreads();
    if (staged) {
        // stage, one thread per halo bin: each entry goes to its rank by previous index
        for (int hb = tid; hb < nh; hb += kTileThreads) {
            const int s0 = S.hbStart[hb];
            const int n = S.hbStart[hb + 1] - s0;
            const int m0 = S.hbGlob[hb];
            for (int t = 0; t < n; ++t) {
                const int i = c.sorted[m0 + t];
                const int rank = n == 1 ? 0 : rank_in_bin(c.sorted, m0, m0 + n, i);
                const int slot = s0 + rank;
                S.pos[slot] = c.A.pos[i];
                S.vel[slot] = c.A.vel[i];
                S.dens[slot] = c.A.dens[i];
                S.flags[slot] = c.A.flags[i];
                S.aidx[slot] = i;
                S.lidx[slot] = m0 + rank;
            }
        }
        __syncthreads();
    }

    // interior cells: halo rows 1..IY, halo columns 1..IX (contiguous per row)
    int interior = 0;
    for (int rr = 1; rr <= IY; ++rr) interior += S.hbStart[rr * HX + 1 + IX] - S.hbStart[rr * HX + 1];
    for (int q = tid; q < interior; q += kTileThreads) {
        // locate the interior entry: its halo row, then its bin
        int acc = 0, lq = -1, row = 1;
        for (int rr = 1; rr <= IY; ++rr) {
            const int n = S.hbStart[rr * HX + 1 + IX] - S.hbStart[rr * HX + 1];
            if (q < acc + n) {
                lq = S.hbStart[rr * HX + 1] + (q - acc);
                row = rr;
                break;
            }
            acc += n;
        }
        int lo = row * HX + 1, hi = row * HX + 1 + IX;  // hbStart[lo] <= lq < hbStart[hi]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (S.hbStart[mid] <= lq)
                lo = mid;
            else
                hi = mid;
        }
        const int hb = lo;
        if (staged) {
            const Nb self = lds_nb(S, lq);
            cell_update<kFluid>(c, w, P, wi, self, [&](auto&& visit) {
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int nb = hb + dy * HX + dx;
                        const int e = S.hbStart[nb + 1];
                        for (int s2 = S.hbStart[nb]; s2 < e; ++s2) visit(lds_nb(S, s2));
                    }
            });
        } else {
            // dense tile: same stencil, same order, from global memory; entries of a bin are
            // visited by increasing previous index (selection over the unsorted bin)
            const int m0 = S.hbGlob[hb];
            const int n0 = S.hbStart[hb + 1] - S.hbStart[hb];
            const int i = c.sorted[m0 + (lq - S.hbStart[hb])];
            const Nb self = global_nb(c, i, m0 + rank_in_bin(c.sorted, m0, m0 + n0, i));
            cell_update<kFluid>(c, w, P, wi, self, [&](auto&& visit) {
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int nb = hb + dy * HX + dx;
                        const int s0 = S.hbGlob[nb];
                        const int n = S.hbStart[nb + 1] - S.hbStart[nb];
                        int last = -1;
                        for (int t = 0; t < n; ++t) {
                            int best = 0x7fffffff;
                            for (int u = 0; u < n; ++u) {
                                const int a = c.sorted[s0 + u];
                                if (a > last && a < best) best = a;
                            }
                            last = best;
                            visit(global_nb(c, best, s0 + t));
                        }
                    }
            });
        }
    }
}

void launch_neighbor(const Ctx& c, const LaunchShape& sh, hipStream_t s, bool fluid, bool collision) {
    if (!sh.tileBlocks) return;
    if (fluid) hipLaunchKernelGGL(k_neighbor_tile<true>, dim3(sh.tileBlocks), dim3(kTileThreads), 0, s, c);
    if (collision) hipLaunchKernelGGL(k_neighbor_tile<false>, dim3(sh.tileBlocks), dim3(kTileThreads), 0, s, c);
}

}  // namespace alien
