// k_neighbor.hip — the neighbour pass of the physics timestep.
//
// ALIEN: calcCollisions_reconnectCells_correctOverlap / calcFluidForces_reconnectCells_correctOverlap
// (source/EngineGpuKernels/CellProcessor.cuh:122-365), then checkForces/applyForces (:367-406); for
// cells without bonds also the Verlet updates and friction of the same step (:525-559, :707-723),
// which reduce to local operations when the connection forces are zero.
//
// One workgroup owns a tile of TX x TY bins.
//  1. Staging: the bin counters of the tile and its one-bin halo give every halo bin its range of
//     LDS entries; one thread per entry loads the cell, then every entry is ranked inside its bin by
//     previous index (the stable in-bin order of the cell map, R1) into the permutation `perm`.
//  2. Candidates: one thread per interior cell walks the 3x3 bin stencil (three contiguous slot
//     ranges) with the neighbour predicate evaluated on squared lengths, and records the
//     neighbours as a bit mask in visiting order.
//  3. Pair terms: the neighbour pairs of all threads are laid out back to back (block scan of the
//     mask popcounts) and evaluated 128 at a time with every lane busy; the terms go to LDS.
//  4. Each thread adds its own pairs' terms in visiting order (the float sums are the same as a
//     sequential walk of the stencil), then finishes the cell.
// A cell with more than 64 stencil candidates, or a tile whose halo does not fit LDS, walks its
// stencil sequentially instead (same order, same results).  The kernel is instantiated per motion
// mode; a workgroup of a world in the other mode exits at once.
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

// what a neighbour pair contributes
constexpr uint32_t kPairBarrier = 1u;  // fluid: a barrier neighbour (A index, distance)
constexpr uint32_t kPairSelf = 2u;     // fluid: the cell itself (density term only)
constexpr uint32_t kPairOverlap = 4u;  // overlap correction
constexpr uint32_t kPairForce = 8u;    // force terms
constexpr int kPairTerms = 7;

// neighbour predicate (fluid: CellProcessor.cuh:137-143; collision: Map.cuh:189 with
// CellProcessor.cuh:290-296): |d| <= R tested as |d|^2 <= reachSq, which is exactly equivalent
template <bool kFluid>
__device__ __forceinline__ bool is_neighbor(const Nb& s, const Nb& o, float Wf, float Hf, float reachSq) {
    if (flags_detached(s.flags) + flags_detached(o.flags) == 1) return false;
    if constexpr (kFluid)
        return lensq(torus_delta(s.pos, o.pos, Wf, Hf)) <= reachSq;
    else
        return o.aidx != s.aidx && lensq(sub(o.pos, s.pos)) <= reachSq;
}

// The contribution of neighbour o to cell s, for a pair that passed is_neighbor.
template <bool kFluid>
__device__ __forceinline__ uint32_t pair_term(const Ctx& c, const DevWorld& w, const DevParams& P, int wi, const Nb& s,
                                              const Nb& o, bool bonded, float* t) {
    const float Wf = float(w.W), Hf = float(w.H);
    const float cmd = P.minDistance;
    const bool bari = flags_barrier(s.flags);
    const f2 pd = torus_delta(s.pos, o.pos, Wf, Hf);
    const float d = len2(pd);
    uint32_t f = 0u;
    if constexpr (kFluid) {
        // CellProcessor.cuh:144-231
        if (flags_barrier(o.flags)) {
            t[0] = __int_as_float(o.aidx);
            t[1] = d;
            return kPairBarrier;
        }
        const float h = P.smoothing;
        t[0] = sph_w(d / h) / (h * h);
        if (o.aidx == s.aidx) return kPairSelf;
        if (!bari && d < cmd) {
            const f2 D = dvd(mul(pd, cmd), 5.0f);
            t[1] = D.x;
            t[2] = D.y;
            f |= kPairOverlap;
        }
        if (bonded) return f;
        const float di = s.dens, dj = o.dens;
        const f2 vd = sub(s.vel, o.vel);
        const float factor = di / (di * di) + dj / (dj * dj);
        if (fabsf(d) > kNearZero) {
            const float kd = sph_dw(d / h) / (h * h * h);
            const f2 Fp = mul(mul(dvd(pd, -d), factor), kd);
            const f2 Fv = dvd(mul(mul(dvd(vd, dj), d), kd), d * d + 0.25f);
            t[3] = Fp.x;
            t[4] = Fp.y;
            t[5] = Fv.x;
            t[6] = Fv.y;
            f |= kPairForce;
        }
        if (len2(vd) >= P.fusionVelocity && s.lidx < o.lidx && flags_num_conn(s.flags) < flags_max_conn(s.flags) &&
            flags_num_conn(o.flags) < flags_max_conn(o.flags) && !bari) {
            const float ei = c.energy[w.cellOff + c.A.slot[s.aidx]];
            const float ej = c.energy[w.cellOff + c.A.slot[o.aidx]];
            if (ei <= P.maxBindingEnergy && ej <= P.maxBindingEnergy) push_candidate(c, s.lidx, o.lidx, wi);
        }
    } else {
        // CellProcessor.cuh:297-361
        if (!bari && d < cmd) {
            const f2 D = dvd(mul(pd, cmd), 5.0f);
            t[0] = D.x;
            t[1] = D.y;
            f |= kPairOverlap;
        }
        if (bonded) return f;
        const float r = P.collisionDistance;
        const f2 vd = sub(s.vel, o.vel);
        const bool appr = dot2(pd, vd) < 0.0f;
        const bool barj = flags_barrier(o.flags);
        f2 fij, fji;
        {
            const float bf = bari ? 2.0f : 1.0f;
            if (len2(s.vel) > 0.5f && appr) {
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
        const f2 nfji = neg(fji);
        t[2] = fij.x;
        t[3] = fij.y;
        t[4] = nfji.x;
        t[5] = nfji.y;
        f |= kPairForce;
        if (s.lidx < o.lidx && flags_num_conn(s.flags) < flags_max_conn(s.flags) &&
            flags_num_conn(o.flags) < flags_max_conn(o.flags) && len2(vd) >= P.fusionVelocity && appr && !bari &&
            !barj) {
            const float ei = c.energy[w.cellOff + c.A.slot[s.aidx]];
            const float ej = c.energy[w.cellOff + c.A.slot[o.aidx]];
            if (ei <= P.maxBindingEnergy && ej <= P.maxBindingEnergy) push_candidate(c, s.lidx, o.lidx, wi);
        }
    }
    return f;
}

// per-cell sums of the pass, added in visiting order.  Of the barrier neighbours only the first
// ten count, and of those the first nearest one (CellProcessor.cuh:234-246); it is tracked as the
// barriers arrive.
struct Acc {
    f2 F, pc, D, Fp, Fv;
    float rho;
    int numBarrier;
    int cb;  // nearest barrier (A index), -1 if none
    float cbDist;
};

__device__ __forceinline__ void acc_init(Acc& a, const Nb& s) {
    a.F = {0.0f, 0.0f};
    a.pc = s.pos;
    a.D = {0.0f, 0.0f};
    a.Fp = {0.0f, 0.0f};
    a.Fv = {0.0f, 0.0f};
    a.rho = 0.0f;
    a.numBarrier = 0;
    a.cb = -1;
    a.cbDist = 0.0f;
}

template <bool kFluid>
__device__ __forceinline__ void acc_add(Acc& a, uint32_t f, const float* t) {
    if constexpr (kFluid) {
        if (f & kPairBarrier) {
            if (a.numBarrier < 10 && (a.cb < 0 || t[1] < a.cbDist)) {
                a.cb = __float_as_int(t[0]);
                a.cbDist = t[1];
            }
            ++a.numBarrier;
            return;
        }
        a.rho = a.rho + t[0];
        if (f & kPairSelf) return;
        if (f & kPairOverlap) a.D = add(a.D, f2{t[1], t[2]});
        if (f & kPairForce) {
            a.Fp = add(a.Fp, f2{t[3], t[4]});
            a.Fv = add(a.Fv, f2{t[5], t[6]});
        }
    } else {
        if (f & kPairOverlap) a.pc = add(a.pc, f2{t[0], t[1]});
        if (f & kPairForce) {
            a.F = add(a.F, f2{t[2], t[3]});
            a.F = add(a.F, f2{t[4], t[5]});
        }
    }
}

__device__ __forceinline__ void load_bonds(const Ctx& c, const Nb& s, int bt[kMaxBonds]) {
    const int nci = flags_num_conn(s.flags);
#pragma unroll
    for (int b = 0; b < kMaxBonds; ++b) bt[b] = b < nci ? c.A.btgt[b * c.cellCap + s.aidx] : -1;
}

__device__ __forceinline__ bool is_bonded(const int bt[kMaxBonds], int jLocal) {
    bool bonded = false;
#pragma unroll
    for (int b = 0; b < kMaxBonds; ++b) bonded = bonded || (bt[b] == jLocal);
    return bonded;
}

// Barrier reflection, applyForces, and the end-of-step writes of one cell.
template <bool kFluid>
__device__ __forceinline__ void cell_finish(const Ctx& c, const DevWorld& w, const DevParams& P, int wi, const Nb& s,
                                            const int bt[kMaxBonds], Acc& a) {
    const int cap = c.cellCap;
    const float Wf = float(w.W), Hf = float(w.H);
    const int i = s.aidx, k = s.lidx;
    const f2 pi = s.pos, vi = s.vel;
    const uint32_t fi = s.flags;
    const int nci = flags_num_conn(fi);
    const bool bari = flags_barrier(fi);
    f2 F = a.F, pc = a.pc;
    if constexpr (kFluid) {
        if (a.cb >= 0) {  // barrier reflection, CellProcessor.cuh:234-274
            const int cb = a.cb;
            const float cbDist = a.cbDist;
            const f2 pb = to_f2(c.A.pos[cb]);
            const f2 vb = to_f2(c.A.vel[cb]);
            const int nb = flags_num_conn(c.A.flags[cb]);
            f2 rr = {0.0f, 0.0f};
            if (nb <= 1) {
                rr = torus_delta(pi, pb, Wf, Hf);
            } else {
                const float angleToCell = angle_of(torus_delta(pi, pb, Wf, Hf));
                for (int q = 0; q < nb; ++q) {
                    const f2 o1 = to_f2(c.A.pos[w.cellOff + c.A.btgt[q * cap + cb]]);
                    const f2 o2 = to_f2(c.A.pos[w.cellOff + c.A.btgt[((q + 1) % nb) * cap + cb]]);
                    const float a1 = angle_of(torus_delta(o1, pb, Wf, Hf));
                    const float a2 = angle_of(torus_delta(o2, pb, Wf, Hf));
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
        F = add(F, add(mul(a.Fp, P.pressure), mul(a.Fv, P.viscosity)));
        pc = add(pi, a.D);
    }

    // the cell in the end-of-step order; bond targets stay in start-of-step indices until the
    // staged cells are renumbered (k_remap)
    c.B.flags[k] = fi;
    c.B.slot[k] = c.A.slot[i];
    c.B.dens[k] = kFluid ? a.rho : s.dens;
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

// One cell, neighbours visited one after another (forEachNb calls visit(o) in stencil order).
template <bool kFluid, class ForEach>
__device__ __forceinline__ void cell_direct(const Ctx& c, const DevWorld& w, const DevParams& P, int wi, const Nb& s,
                                            ForEach&& forEachNb) {
    const float Wf = float(w.W), Hf = float(w.H);
    int bt[kMaxBonds];
    load_bonds(c, s, bt);
    Acc a;
    acc_init(a, s);
    forEachNb([&](const Nb& o) {
        if (!is_neighbor<kFluid>(s, o, Wf, Hf, P.reachSq)) return;
        float t[kPairTerms];
        const uint32_t f = pair_term<kFluid>(c, w, P, wi, s, o, is_bonded(bt, o.aidx - w.cellOff), t);
        acc_add<kFluid>(a, f, t);
    });
    cell_finish<kFluid>(c, w, P, wi, s, bt, a);
}

struct TileShared {
    float2 pos[kTileCap];
    float2 vel[kTileCap];
    float dens[kTileCap];
    uint32_t flags[kTileCap];
    int aidx[kTileCap];
    int lidx[kTileCap];
    uint16_t perm[kTileCap];        // slot (bin order, then rank) -> entry
    int hbStart[kMaxHaloBins + 1];  // first slot of every halo bin (exclusive scan)
    union {
        struct {
            int hbGlob[kMaxHaloBins];  // position of the bin's first entry in `sorted`
            uint16_t ebin[kTileCap];   // halo bin of every entry
        } st;
        struct {
            unsigned long long mask[kTileThreads];   // neighbours among the stencil candidates
            unsigned long long bmask[kTileThreads];  // ... that are bond partners
            int pbase[kTileThreads + 1];             // first pair of every thread
            int owner[kTileThreads];                 // entry of the thread's cell
            uint32_t rows[3][kTileThreads];          // stencil rows: first slot | count << 16
            float term[kPairTerms][kTileThreads];
            uint32_t pflag[kTileThreads];
        } cp;
    } u;
};

__device__ __forceinline__ Nb lds_nb(const TileShared& S, int e) {
    Nb o;
    o.pos = to_f2(S.pos[e]);
    o.vel = to_f2(S.vel[e]);
    o.dens = S.dens[e];
    o.flags = S.flags[e];
    o.aidx = S.aidx[e];
    o.lidx = S.lidx[e];
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

// position of the r-th set bit of m (m has more than r set bits)
__device__ __forceinline__ int nth_bit(unsigned long long m, int r) {
    for (int q = 0; q < r; ++q) m &= m - 1ull;
    return __builtin_ctzll(m);
}

// interior slot q of the tile (slots of halo rows 1..IY, columns 1..IX, row by row) and its bin
__device__ __forceinline__ void interior_slot(const TileShared& S, int q, int HX, int IX, int IY, int& slot, int& hb) {
    int acc = 0, row = 1;
    slot = -1;
    for (int rr = 1; rr <= IY; ++rr) {
        const int n = S.hbStart[rr * HX + 1 + IX] - S.hbStart[rr * HX + 1];
        if (q < acc + n) {
            slot = S.hbStart[rr * HX + 1] + (q - acc);
            row = rr;
            break;
        }
        acc += n;
    }
    int lo = row * HX + 1, hi = row * HX + 1 + IX;  // hbStart[lo] <= slot < hbStart[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (S.hbStart[mid] <= slot)
            lo = mid;
        else
            hi = mid;
    }
    hb = lo;
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
    const float Wf = float(w.W), Hf = float(w.H);
    constexpr int kPer = kMaxHaloBins / kTileThreads;
    constexpr int kEnt = kTileCap / kTileThreads;

    // 1a. halo bins: entry counts and slot offsets
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
            S.u.st.hbGlob[hb] = w.cellOff + int(s - base);
        }
        cnt[q] = n;
        local += uint32_t(n);
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan(local, total);
    const bool staged = total <= uint32_t(kTileCap);
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int hb = tid * kPer + q;
        if (hb < nh) {
            S.hbStart[hb] = int(off);
            if (staged)
                for (int t = 0; t < cnt[q]; ++t) S.u.st.ebin[off + t] = uint16_t(hb);
        }
        off += uint32_t(cnt[q]);
    }
    if (tid == 0) S.hbStart[nh] = int(total);
    __syncthreads();

    int interior = 0;
    for (int rr = 1; rr <= IY; ++rr) interior += S.hbStart[rr * HX + 1 + IX] - S.hbStart[rr * HX + 1];

    if (!staged) {
        // dense tile: every thread walks the stencil of its cells straight from global memory;
        // entries of a bin are visited by increasing previous index (selection over the bin)
        for (int q = tid; q < interior; q += kTileThreads) {
            int lq, hb;
            interior_slot(S, q, HX, IX, IY, lq, hb);
            const int m0 = S.u.st.hbGlob[hb];
            const int n0 = S.hbStart[hb + 1] - S.hbStart[hb];
            const int i = c.sorted[m0 + (lq - S.hbStart[hb])];
            const Nb self = global_nb(c, i, m0 + rank_in_bin(c.sorted, m0, m0 + n0, i));
            cell_direct<kFluid>(c, w, P, wi, self, [&](auto&& visit) {
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int nb = hb + dy * HX + dx;
                        const int s0 = S.u.st.hbGlob[nb];
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
        return;
    }

    // 1b. entries: one thread per entry, all loads of a thread in flight together
    {
        int ii[kEnt];
#pragma unroll
        for (int q = 0; q < kEnt; ++q) {
            const int e = tid + q * kTileThreads;
            ii[q] = -1;
            if (e < int(total)) {
                const int hb = S.u.st.ebin[e];
                ii[q] = c.sorted[S.u.st.hbGlob[hb] + (e - S.hbStart[hb])];
            }
        }
#pragma unroll
        for (int q = 0; q < kEnt; ++q) {
            const int e = tid + q * kTileThreads;
            const int i = ii[q];
            if (i >= 0) {
                S.pos[e] = c.A.pos[i];
                S.vel[e] = c.A.vel[i];
                S.dens[e] = c.A.dens[i];
                S.flags[e] = c.A.flags[i];
                S.aidx[e] = i;
            }
        }
    }
    __syncthreads();
    // 1c. rank inside the bin by previous index (R1)
    for (int e = tid; e < int(total); e += kTileThreads) {
        const int hb = S.u.st.ebin[e];
        const int s0 = S.hbStart[hb], s1 = S.hbStart[hb + 1];
        const int a = S.aidx[e];
        int r = 0;
        for (int u = s0; u < s1; ++u) r += S.aidx[u] < a ? 1 : 0;
        S.perm[s0 + r] = uint16_t(e);
        S.lidx[e] = S.u.st.hbGlob[hb] + r;
    }
    __syncthreads();

    for (int q0 = 0; q0 < interior; q0 += kTileThreads) {
        // 2. stencil candidates of this thread's cell
        const int q = q0 + tid;
        const bool active = q < interior;
        Nb self;
        int bt[kMaxBonds];
        int own = -1;
        bool direct = false;
        uint32_t rows[3] = {0u, 0u, 0u};
        unsigned long long mask = 0ull, bmask = 0ull;
        if (active) {
            int lq, hb;
            interior_slot(S, q, HX, IX, IY, lq, hb);
            own = S.perm[lq];
            self = lds_nb(S, own);
            load_bonds(c, self, bt);
            int ncand = 0;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                const int rb = hb + (r - 1) * HX;
                const int a0 = S.hbStart[rb - 1], a1 = S.hbStart[rb + 2];
                rows[r] = uint32_t(a0) | (uint32_t(a1 - a0) << 16);
                ncand += a1 - a0;
            }
            if (ncand > 64) {
                direct = true;
            } else {
                const bool hasBonds = flags_num_conn(self.flags) > 0;
                int cidx = 0;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const int a0 = int(rows[r] & 0xFFFFu), n = int(rows[r] >> 16);
                    for (int sl = a0; sl < a0 + n; ++sl, ++cidx) {
                        const int e = S.perm[sl];
                        Nb o;
                        o.pos = to_f2(S.pos[e]);
                        o.flags = S.flags[e];
                        o.aidx = S.aidx[e];
                        if (is_neighbor<kFluid>(self, o, Wf, Hf, P.reachSq)) {
                            mask |= 1ull << cidx;
                            if (hasBonds && is_bonded(bt, o.aidx - w.cellOff)) bmask |= 1ull << cidx;
                        }
                    }
                }
            }
        }
        const int count = __builtin_popcountll(mask);
        uint32_t numPairs;
        const int pbase = int(block_exclusive_scan(uint32_t(count), numPairs));
        S.u.cp.mask[tid] = mask;
        S.u.cp.bmask[tid] = bmask;
        S.u.cp.pbase[tid] = pbase;
        S.u.cp.owner[tid] = own;
#pragma unroll
        for (int r = 0; r < 3; ++r) S.u.cp.rows[r][tid] = rows[r];
        if (tid == 0) S.u.cp.pbase[kTileThreads] = int(numPairs);
        __syncthreads();

        Acc a;
        if (active) acc_init(a, self);
        for (int off2 = 0; off2 < int(numPairs); off2 += kTileThreads) {
            // 3. pair terms, one pair per thread
            const int p = off2 + tid;
            if (p < int(numPairs)) {
                int lo = 0, hi = kTileThreads;  // owner: last thread with pbase <= p
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (S.u.cp.pbase[mid] <= p)
                        lo = mid;
                    else
                        hi = mid;
                }
                const int ci = nth_bit(S.u.cp.mask[lo], p - S.u.cp.pbase[lo]);
                int sl = -1, rest = ci;
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    const uint32_t rw = S.u.cp.rows[r][lo];
                    const int n = int(rw >> 16);
                    if (sl < 0) {
                        if (rest < n)
                            sl = int(rw & 0xFFFFu) + rest;
                        else
                            rest -= n;
                    }
                }
                const Nb s = lds_nb(S, S.u.cp.owner[lo]);
                const Nb o = lds_nb(S, S.perm[sl]);
                float t[kPairTerms];
                const bool bonded = ((S.u.cp.bmask[lo] >> ci) & 1ull) != 0ull;
                const uint32_t f = pair_term<kFluid>(c, w, P, wi, s, o, bonded, t);
#pragma unroll
                for (int v = 0; v < kPairTerms; ++v) S.u.cp.term[v][tid] = t[v];
                S.u.cp.pflag[tid] = f;
            }
            __syncthreads();
            // 4. the owner adds its pairs of this round in order
            if (count > 0) {
                const int lo = max(pbase, off2), hi = min(pbase + count, off2 + kTileThreads);
                for (int p2 = lo; p2 < hi; ++p2) {
                    const int slotp = p2 - off2;
                    float t[kPairTerms];
#pragma unroll
                    for (int v = 0; v < kPairTerms; ++v) t[v] = S.u.cp.term[v][slotp];
                    acc_add<kFluid>(a, S.u.cp.pflag[slotp], t);
                }
            }
            __syncthreads();
        }
        if (active) {
            if (direct) {
                cell_direct<kFluid>(c, w, P, wi, self, [&](auto&& visit) {
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        const int a0 = int(rows[r] & 0xFFFFu), n = int(rows[r] >> 16);
                        for (int sl = a0; sl < a0 + n; ++sl) visit(lds_nb(S, S.perm[sl]));
                    }
                });
            } else {
                cell_finish<kFluid>(c, w, P, wi, self, bt, a);
            }
        }
        __syncthreads();  // the pair arrays are reused by the next group of cells
    }
}

void launch_neighbor(const Ctx& c, const LaunchShape& sh, hipStream_t s, bool fluid, bool collision) {
    if (!sh.tileBlocks) return;
    if (fluid) hipLaunchKernelGGL(k_neighbor_tile<true>, dim3(sh.tileBlocks), dim3(kTileThreads), 0, s, c);
    if (collision) hipLaunchKernelGGL(k_neighbor_tile<false>, dim3(sh.tileBlocks), dim3(kTileThreads), 0, s, c);
}

}  // namespace alien
