// k_neighbor.hip — the neighbour pass of the physics timestep.
//
// ALIEN: calcCollisions_reconnectCells_correctOverlap / calcFluidForces_reconnectCells_correctOverlap
// (source/EngineGpuKernels/CellProcessor.cuh:122-365), then checkForces/applyForces (:367-406); for
// cells without bonds also the Verlet updates and friction of the same step (:525-559, :707-723),
// which reduce to local operations when the connection forces are zero.
//
// One wave owns a strip of bin columns and walks it row by row (see "Row streaming" below).  Per
// row it (1) tests the 3x3 bin stencil of every cell with the neighbour predicate on squared
// lengths and records the neighbours as a bit mask in visiting order, (2) evaluates all neighbour
// pairs of the row one pair per lane, and (3) lets every cell add its pairs' terms in visiting
// order, which gives the float sums of a sequential stencil walk bit for bit.  A cell with more
// than 64 stencil candidates, or a strip whose rows do not fit LDS, walks its stencil
// sequentially instead (same order, same results).  The kernel is instantiated per motion mode; a
// workgroup of a world in the other mode exits at once.
//
// Divisions by per-world constants (1/h, 1/h^2, 1/h^3), by a cell's density, by |d| and by the
// collision and viscosity denominators are products with the correctly rounded reciprocal (R16
// of the oracle), and the length comparisons |v| > c are tests of |v|^2 against the exact
// threshold of the correctly rounded square root.
#include <type_traits>

#include "kernels_common.hpp"
#include "kernels.hpp"

namespace alien {

struct Nb {
    f2 pos, vel;
    float dens;  // fluid: 1 / density; collision: density
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
// CellProcessor.cuh:290-296): |d| <= R tested as |d|^2 <= reachSq, which is exactly equivalent.
// Cells are never detached here (detached cells come from cell functions), so the detached rule
// never applies.
template <bool kFluid>
__device__ __forceinline__ bool is_neighbor(const Nb& s, const Nb& o, float Wf, float Hf, float reachSq) {
    if constexpr (kFluid)
        return lensq(torus_delta(s.pos, o.pos, Wf, Hf)) <= reachSq;
    else
        return o.aidx != s.aidx && lensq(sub(o.pos, s.pos)) <= reachSq;
}

// The contribution of neighbour o to cell s, for a pair that passed is_neighbor.  The logical
// indices of the two cells are needed only for a fusion candidate: lidxOf() returns them as
// (s, o) when asked.
template <bool kFluid, class LidxOf>
__device__ __forceinline__ uint32_t pair_term(const Ctx& c, const DevWorld& w, const DevParams& P, int wi, const Nb& s,
                                              const Nb& o, bool bonded, float* t, bool seam, LidxOf&& lidxOf) {
    const float Wf = float(w.W), Hf = float(w.H);
    const float cmd = P.minDistance;
    const bool bari = flags_barrier(s.flags);
    const f2 pd = stencil_delta(s.pos, o.pos, Wf, Hf, seam);
    const float d = len2(pd);
    uint32_t f = 0u;
    if constexpr (kFluid) {
        // CellProcessor.cuh:144-231
        if (flags_barrier(o.flags)) {
            t[0] = __int_as_float(o.aidx);
            t[1] = d;
            return kPairBarrier;
        }
        t[0] = sph_w(d * P.invH) * P.invH2;
        if (o.aidx == s.aidx) return kPairSelf;
        if (!bari && d < cmd) {
            const f2 D = dvd(mul(pd, cmd), 5.0f);
            t[1] = D.x;
            t[2] = D.y;
            f |= kPairOverlap;
        }
        if (bonded) return f;
        const f2 vd = sub(s.vel, o.vel);
        const float factor = s.dens + o.dens;  // 1/density_i + 1/density_j
        if (fabsf(d) > kNearZero) {
            const float kd = sph_dw(d * P.invH) * P.invH3;
            const f2 Fp = mul(mul(mul(pd, 1.0f / (-d)), factor), kd);
            const f2 Fv = mul(mul(mul(mul(vd, o.dens), d), kd), 1.0f / (d * d + 0.25f));
            t[3] = Fp.x;
            t[4] = Fp.y;
            t[5] = Fv.x;
            t[6] = Fv.y;
            f |= kPairForce;
        }
        if (flags_num_conn(s.flags) < flags_max_conn(s.flags) && flags_num_conn(o.flags) < flags_max_conn(o.flags) &&
            !bari && lensq(vd) > P.fusionBelowSq) {
            const int2 L = lidxOf();
            if (L.x < L.y) {
                const float ei = c.energy[w.cellOff + c.A.slot[s.aidx]];
                const float ej = c.energy[w.cellOff + c.A.slot[o.aidx]];
                if (ei <= P.maxBindingEnergy && ej <= P.maxBindingEnergy) push_candidate(c, L.x, L.y, wi);
            }
        }
    } else {
        // CellProcessor.cuh:297-361.  f(j,i) mirrors f(i,j): unit(-d) = -unit(d) and
        // (-dv).(-d) = dv.d exactly, so only the barrier factor and the branch differ.
        if (!bari && d < cmd) {
            const f2 D = dvd(mul(pd, cmd), 5.0f);
            t[0] = D.x;
            t[1] = D.y;
            f |= kPairOverlap;
        }
        if (bonded) return f;
        const f2 vd = sub(s.vel, o.vel);
        const float dvp = dot2(vd, pd);
        const bool appr = dot2(pd, vd) < 0.0f;
        const bool barj = flags_barrier(o.flags);
        const bool fastI = appr && lensq(s.vel) > P.fastSq;
        const bool fastJ = appr && lensq(o.vel) > P.fastSq;
        f2 fastT = {0.0f, 0.0f}, repT = {0.0f, 0.0f};
        if (fastI || fastJ) fastT = mul(mul(pd, dvp), 1.0f / (-2.0f * (d * d + 0.25f)));
        if (!(fastI && fastJ)) repT = mul(mul(unit(pd), P.collisionDistance - d), P.repulsion);
        const f2 fij = mul(fastI ? fastT : repT, bari ? 2.0f : 1.0f);
        // -f(j,i); unit() of a (near) zero vector is (1,0) in both directions, so there the
        // repulsion is not mirrored
        const f2 fjRep = mul(repT, barj ? 2.0f : 1.0f);
        const f2 nfji = fastJ ? mul(fastT, barj ? 2.0f : 1.0f) : (d > kNearZero ? fjRep : neg(fjRep));
        t[2] = fij.x;
        t[3] = fij.y;
        t[4] = nfji.x;
        t[5] = nfji.y;
        f |= kPairForce;
        if (flags_num_conn(s.flags) < flags_max_conn(s.flags) && flags_num_conn(o.flags) < flags_max_conn(o.flags) &&
            appr && !bari && !barj && lensq(vd) > P.fusionBelowSq) {
            const int2 L = lidxOf();
            if (L.x < L.y) {
                const float ei = c.energy[w.cellOff + c.A.slot[s.aidx]];
                const float ej = c.energy[w.cellOff + c.A.slot[o.aidx]];
                if (ei <= P.maxBindingEnergy && ej <= P.maxBindingEnergy) push_candidate(c, L.x, L.y, wi);
            }
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

// acc_add without branches, for terms zeroed before pair_term wrote them (lanes with !on add
// nothing).  Adding a zero term leaves a sum that started at +0 unchanged bit for bit (such a sum
// is never -0), so the flags select only what is not a plain zero: a barrier neighbour's data, and
// the collision position sum, which starts at the cell's position.
template <bool kFluid>
__device__ __forceinline__ void acc_add_sel(Acc& a, bool on, uint32_t f, const float* t) {
    if constexpr (kFluid) {
        const bool bar = on && (f & kPairBarrier) != 0u;
        const bool sum = on && !bar;
        const bool near = bar && a.numBarrier < 10 && (a.cb < 0 || t[1] < a.cbDist);
        a.cb = near ? __float_as_int(t[0]) : a.cb;
        a.cbDist = near ? t[1] : a.cbDist;
        a.numBarrier += bar ? 1 : 0;
        a.rho = sum ? a.rho + t[0] : a.rho;
        a.D.x = sum ? a.D.x + t[1] : a.D.x;
        a.D.y = sum ? a.D.y + t[2] : a.D.y;
        a.Fp.x = on ? a.Fp.x + t[3] : a.Fp.x;
        a.Fp.y = on ? a.Fp.y + t[4] : a.Fp.y;
        a.Fv.x = on ? a.Fv.x + t[5] : a.Fv.x;
        a.Fv.y = on ? a.Fv.y + t[6] : a.Fv.y;
    } else {
        const bool ov = on && (f & kPairOverlap) != 0u;
        a.pc.x = ov ? a.pc.x + t[0] : a.pc.x;
        a.pc.y = ov ? a.pc.y + t[1] : a.pc.y;
        f2 F = add(a.F, f2{t[2], t[3]});
        F = add(F, f2{t[4], t[5]});
        a.F.x = on ? F.x : a.F.x;
        a.F.y = on ? F.y : a.F.y;
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

// Barrier reflection, applyForces, and the end-of-step writes of one cell.  A bonded or barrier
// cell is marked staged for the bond kernels of the step.
template <bool kFluid>
__device__ __forceinline__ void cell_finish(const Ctx& c, const DevWorld& w, const DevParams& P, int wi, const Nb& s,
                                            int slot, const int bt[kMaxBonds], Acc& a, int gbOld = -1) {
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
    c.B.slot[k] = slot;
    c.B.dens[k] = kFluid ? a.rho : s.dens;
    if constexpr (kFluid) {
        // (unrolled: a loop to nci would index bt at run time and place it in scratch memory)
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b)
            if (b < nci) {
                c.B.btgt[b * cap + k] = bt[b];
                c.B.brest[b * cap + k] = c.A.brest[b * cap + i];
                c.B.bang[b * cap + k] = c.A.bang[b * cap + i];
            }
        if (nci > 0) c.bstiff[k] = c.stiff[w.cellOff + slot];
    } else {
        // (collision worlds are mostly bonded cells: their bond data is loaded before any of it
        // is stored, so that the loads are in flight together)
        float brest[kMaxBonds], bang[kMaxBonds];
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b) {
            brest[b] = b < nci ? c.A.brest[b * cap + i] : 0.0f;
            bang[b] = b < nci ? c.A.bang[b * cap + i] : 0.0f;
        }
        const float stiff = nci > 0 ? c.stiff[w.cellOff + slot] : 0.0f;
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b)
            if (b < nci) {
                c.B.btgt[b * cap + k] = bt[b];
                c.B.brest[b * cap + k] = brest[b];
                c.B.bang[b * cap + k] = bang[b];
            }
        if (nci > 0) c.bstiff[k] = stiff;
    }
    if (nci > 0) c.newidx[i] = k;
    // checkForces (CellProcessor.cuh:367-386): a bonded cell whose force is above cellMaxForce loses
    // all its bonds with probability cellMaxForceDecayProb (deleting the bonds of an unbonded cell
    // would change nothing)
    // (the per-colour threshold is read from memory: indexing the register copy of the
    // parameters by a run-time colour would place the whole copy in scratch memory)
    if (P.decayProb > 0.0f && nci > 0 && !bari && lensq(F) > c.params[wi].maxForceSq[flags_color(fi)]) {
        const unsigned long long id = c.cid[w.cellOff + slot];
        if (uniform_draw(w.seed, c.ctr[wi].timestep, id, kStreamForceDecay) < P.decayProb) c.delFlag[k] = 1;
    }
    f2 v = vi;
    if (!bari) {  // applyForces
        v = add(v, F);
        if (lensq(v) > P.maxVelSq) v = mul(unit(v), P.maxVelocity);
    }
    if (!bari && nci == 0) {
        // no bonds: both connection-force passes yield zero, so the Verlet position and velocity
        // updates and friction of this step are applied here with the same operations
        const f2 zero = {0.0f, 0.0f};
        const float dt = P.dt;
        const f2 p2 = torus_pos(add(pc, add(mul(v, dt), mul(mul(mul(zero, dt), dt), 0.5f))), w.W, w.H);
        const f2 acc = mul(add(zero, zero), 0.5f);
        v = add(v, mul(acc, dt));
        v = mul(v, 1.0f - P.friction);
        c.B.pos[k] = to_float2(p2);
        c.B.vel[k] = to_float2(v);
        bin_next(c, w, k, p2, gbOld >= 0 ? gbOld : cell_bin_of(w, pi));
    } else {
        c.B.pos[k] = to_float2(pc);
        c.B.vel[k] = to_float2(v);
    }
    if (w.bondWork) c.stagedMark[k] = uint8_t(bari || nci != 0);
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
        const uint32_t f = pair_term<kFluid>(c, w, P, wi, s, o, is_bonded(bt, o.aidx - w.cellOff), t, true,
                                             [&] { return make_int2(s.lidx, o.lidx); });
        acc_add<kFluid>(a, f, t);
    });
    cell_finish<kFluid>(c, w, P, wi, s, c.A.slot[s.aidx], bt, a);
}


// ---------------------------------------------------------------------------------------------
// Row streaming.  A workgroup (one wave) owns a strip of up to kStripMax bin columns and a chunk of
// staged rows, and walks them top to bottom with staged rows Y-1, Y, Y+1 in LDS.  A staged row is
// kSub consecutive bin rows (1, or 2 for sparse worlds, so that one pass has enough cells for the
// 64 lanes); each bin row covers the strip plus one halo column on each side, its entries stored
// bin by bin in rank order (R1), so the three stencil bins of a cell in one bin row are one
// contiguous range.  While row Y is computed, the entries of row Y+2 and the bin counters of row
// Y+3 are on their way into registers; row Y+2 is committed once row Y-1 is no longer needed.  The
// three staged rows share one ring pool of entries (positions counted up without wrapping, stored
// modulo kPool): a row takes the pool space right after the previous one, so any three
// consecutive rows of up to kPool entries together fit, however they are distributed.
// ---------------------------------------------------------------------------------------------
constexpr int kRowBins = kStripMax + 2;            // halo columns of one bin row
constexpr int kSubMax = 2;                         // bin rows per staged row, at most
constexpr int kPool = kRowPool;
constexpr int kPoolMask = kPool - 1;
constexpr int kPairCap = 128;  // pair records per evaluation window
constexpr int kPosMirror = 64;
static_assert((kPool & kPoolMask) == 0 && kPool <= 256, "pair records keep an 8-bit pool position");
static_assert(kRowCap <= 2 * kRowThreads, "a row in flight holds at most two entries per lane");
static_assert(kSubMax * kRowBins <= 3 * kRowThreads && kSubMax * kRowBins < 256,
              "bin slots: three per lane, 8-bit ids");

// LDS of one wave, for staged rows of up to kMaxSub bin rows
template <int kMaxSub>
struct RowShared {
    static constexpr int kStageBins = kMaxSub * kRowBins;  // bin slots of one staged row (row major)
    // per ring row: pool position of the first entry of every bin slot, then the end; per ring
    // row and bin row, map positions of the first entries of the left halo bin, the first interior
    // bin and the right halo bin (the interior bins of a bin row are contiguous in the map)
    int rs[3][kStageBins + 1];
    int rm[3][kMaxSub][3];
    // entries of the staged rows, by pool position modulo kPool; positions 0..kPosMirror-1 are
    // repeated after the end so that a stencil range of up to 64 entries reads without wrapping
    float2 pos[kPool + kPosMirror];
    float2 vel[kPool];
    float dens[kPool];  // fluid: 1 / density; collision: density
    uint16_t flags[kPool];  // the bits the pass reads of an entry's flags (flags16)
    int aidx[kPool];
    uint8_t sbin[kPool];  // bin slot: bin row * HX + halo column
    // the row in flight
    int nrs[kStageBins + 1];
    int nrm[kMaxSub][3];
    uint8_t nebin[kRowCap];  // entry of the row -> bin slot
    union {
        struct {
            // while a row is computed, the pair records of the current window: the cell's pool
            // position, and the neighbour as bonded << 15 | ring row << 8 | pool position
            uint8_t pself[kPairCap];
            uint16_t pnb[kPairCap];
        } pr;
    } u;
};

// registers of a row in flight: bin counters of the bin slots lane, lane + 64, lane + 128, then
// the entries lane and lane + 64 in map order
template <int kSub>
struct RowFlight {
    static constexpr int kSlots = kSub == 2 ? 3 : 2;  // bin slots per lane
    uint32_t s0[kSlots], s1[kSlots];
    int aidx[2];
    float2 pos[2], vel[2];
    float dens[2];
    uint32_t flags[2];
    int total;
};

struct StripGeom {
    int x0, IX, HX;  // first interior bin column, interior columns, halo columns (IX + 2)
    int nbx, nby, binOff, cellOff;
    uint32_t base;   // start[] value of the world's first bin
    bool seam;       // the staged rows or columns wrap around the torus
};

// map segment of halo column hx: 0 the left halo bin, 1 the interior, 2 the right halo bin; and
// the segment's first halo column
__device__ __forceinline__ int seg_of(int hx, int HX) { return hx == 0 ? 0 : (hx == HX - 1 ? 2 : 1); }
__device__ __forceinline__ int seg_first(int seg, int HX) { return seg == 0 ? 0 : (seg == 2 ? HX - 1 : 1); }

// bin row (0 or 1) of bin slot j
template <int kSub>
__device__ __forceinline__ int slot_row(int j, int HX) { return kSub == 2 && j >= HX ? 1 : 0; }

// An entry's flags in LDS: bond counts (bits 0..7) and the barrier bit (moved to bit 8), the only
// bits the pass reads of a neighbour; a cell's own full flags are read again before its writes.
__device__ __forceinline__ uint16_t flags16(uint32_t f) { return uint16_t((f & 0xFFu) | ((f & kBarrierBit) >> 8)); }
__device__ __forceinline__ uint32_t flags32(uint16_t p) { return (uint32_t(p) & 0xFFu) | ((uint32_t(p) & 0x100u) << 8); }

// logical index of the staged entry at pool position pp (already reduced modulo kPool) of ring row r
template <int kSub, class RS>
__device__ __forceinline__ int row_lidx(const RS& S, const StripGeom& g, int r, int pp) {
    const int j = S.sbin[pp];
    const int t = slot_row<kSub>(j, g.HX);
    const int seg = seg_of(j - t * g.HX, g.HX);
    return S.rm[r][t][seg] + ((pp - S.rs[r][t * g.HX + seg_first(seg, g.HX)]) & kPoolMask);
}

// the staged entry at pool position pp of ring row r; without its logical index if !kLidx
template <int kSub, class RS, bool kLidx = true>
__device__ __forceinline__ Nb row_nb(const RS& S, const StripGeom& g, int r, int pp) {
    Nb o;
    o.pos = to_f2(S.pos[pp]);
    o.vel = to_f2(S.vel[pp]);
    o.dens = S.dens[pp];
    o.flags = flags32(S.flags[pp]);
    o.aidx = S.aidx[pp];
    o.lidx = kLidx ? row_lidx<kSub>(S, g, r, pp) : -1;
    return o;
}

template <bool kFluid>
__device__ __forceinline__ Nb global_nb(const Ctx& c, int a, int l) {
    Nb o;
    o.pos = to_f2(c.A.pos[a]);
    o.vel = to_f2(c.A.vel[a]);
    o.dens = kFluid ? 1.0f / c.A.dens[a] : c.A.dens[a];
    o.flags = c.A.flags[a];
    o.aidx = a;
    o.lidx = l;
    return o;
}

// bin of halo column hx in bin row yy (both wrapped around the torus)
__device__ __forceinline__ int strip_bin(const StripGeom& g, int yy, int hx) {
    return g.binOff + wrap_small(yy, g.nby) * g.nbx + wrap_small(g.x0 - 1 + hx, g.nbx);
}

// A: bin counters of staged row Y
template <int kSub>
__device__ __forceinline__ void row_issue_counts(const Ctx& c, const StripGeom& g, int Y, RowFlight<kSub>& F) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int q = 0; q < RowFlight<kSub>::kSlots; ++q) {
        const int j = lane + q * kRowThreads;
        F.s0[q] = F.s1[q] = 0u;
        if (j < kSub * g.HX) {
            const int t = slot_row<kSub>(j, g.HX);
            const int gb = strip_bin(g, kSub * Y + t, j - t * g.HX);
            F.s0[q] = c.start[gb];
            F.s1[q] = c.start[gb + 1];
        }
    }
}

// B: slots of the row's bins, then the entry loads (map-order copy, contiguous per bin)
template <int kSub, class RS>
__device__ __forceinline__ void row_issue_entries(const Ctx& c, const StripGeom& g, RS& S, RowFlight<kSub>& F) {
    constexpr int kSlots = RowFlight<kSub>::kSlots;
    const int lane = threadIdx.x;
    const int SB = kSub * g.HX;
    int n[kSlots], inc[kSlots];
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        n[q] = lane + q * kRowThreads < SB ? int(F.s1[q] - F.s0[q]) : 0;
        inc[q] = int(wave_inclusive_scan(uint32_t(n[q])));
    }
    const int tot0 = wave_last(inc[0]), tot1 = wave_last(inc[1]);
    F.total = tot0 + tot1 + (kSlots > 2 ? wave_last(inc[kSlots - 1]) : 0);
#pragma unroll
    for (int q = 0; q < kSlots; ++q) {
        const int j = lane + q * kRowThreads;
        if (j >= SB) continue;
        const int off = inc[q] - n[q] + (q >= 1 ? tot0 : 0) + (q >= 2 ? tot1 : 0);
        const int t = slot_row<kSub>(j, g.HX);
        const int hx = j - t * g.HX;
        S.nrs[j] = off;
        if (hx <= 1 || hx == g.HX - 1) S.nrm[t][seg_of(hx, g.HX)] = g.cellOff + int(F.s0[q] - g.base);
        for (int e = 0; e < n[q]; ++e) S.nebin[off + e] = uint8_t(j);
    }
    if (lane == 0) S.nrs[SB] = F.total;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int e = lane + q * kRowThreads;
        if (e < F.total) {
            const int j = S.nebin[e];
            const int t = slot_row<kSub>(j, g.HX);
            const int seg = seg_of(j - t * g.HX, g.HX);
            const int m = S.nrm[t][seg] + (e - S.nrs[t * g.HX + seg_first(seg, g.HX)]);
            F.aidx[q] = c.sorted[m];
            F.pos[q] = c.mpos[m];
            F.vel[q] = c.mvel[m];
            F.dens[q] = c.mdens[m];
            F.flags[q] = c.mflags[m];
        }
    }
}

// C: the row into ring row r at pool position head.  The row was read from the map copy bin by
// bin, and the map holds every bin's cells by previous index (R1): an entry's place in the row is
// its rank in its bin.
template <int kSub, class RS>
__device__ __forceinline__ void row_commit(const StripGeom& g, RS& S, const RowFlight<kSub>& F, int r, int head) {
    const int lane = threadIdx.x;
    const int SB = kSub * g.HX;
    for (int i = lane; i <= SB; i += kRowThreads) S.rs[r][i] = head + S.nrs[i];
    if (lane < 3 * kSub) S.rm[r][lane / 3][lane % 3] = S.nrm[lane / 3][lane % 3];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int e = lane + q * kRowThreads;
        if (e < F.total) {
            const int j = S.nebin[e];
            const int a = F.aidx[q];
            const int pp = (head + e) & kPoolMask;
            S.pos[pp] = F.pos[q];
            if (pp < kPosMirror) S.pos[kPool + pp] = F.pos[q];
            S.vel[pp] = F.vel[q];
            S.dens[pp] = F.dens[q];
            S.flags[pp] = flags16(F.flags[q]);
            S.aidx[pp] = a;
            S.sbin[pp] = uint8_t(j);
        }
    }
    __syncthreads();
}

// element r (0..2) of a three-element register array, without indexing it at run time (which
// would place the array in scratch memory)
__device__ __forceinline__ int pick3(const int (&a)[3], int r) { return r == 0 ? a[0] : (r == 1 ? a[1] : a[2]); }

// The cells of staged row Y (ring row rc; rows Y-1 and Y+1 in rm_ and rp).
template <bool kFluid, int kSub, class RS>
__device__ __forceinline__ void row_compute(const Ctx& c, const DevWorld& w, const DevParams& P, int wi,
                                            const StripGeom& g, RS& S, int rm_, int rc, int rp) {
    const int lane = threadIdx.x;
    const float Wf = float(w.W), Hf = float(w.H);
    // the cells: the interior bins of each bin row of the staged row
    const int cs0 = S.rs[rc][1];
    const int n0 = S.rs[rc][g.IX + 1] - cs0;
    const int cs1 = kSub == 2 ? S.rs[rc][g.HX + 1] : 0;
    const int ncell = n0 + (kSub == 2 ? S.rs[rc][g.HX + g.IX + 1] - cs1 : 0);
    const float w0 = kFluid ? sph_w(0.0f) * P.invH2 : 0.0f;  // a fluid cell's own density term
    for (int q0 = 0; q0 < ncell; q0 += kRowThreads) {
        // candidates of this lane's cell: three contiguous pool ranges, one per stencil bin row
        const int q = q0 + lane;
        const bool active = q < ncell;
        Nb self;
        int bt[kMaxBonds];
        int a0s[3] = {0, 0, 0}, ns[3] = {0, 0, 0}, ringOf[3] = {rm_, rc, rp};
        unsigned long long mask = 0ull, bmask = 0ull;
        bool direct = false;
        int selfIdx = -1;  // fluid: number of the cell's pairs that come before its own density term
        const int t = kSub == 2 && q >= n0 ? 1 : 0;    // the cell's bin row in the staged row
        const int su = t == 0 ? cs0 + q : cs1 + (q - n0);  // pool position, not wrapped
        const int sp = su & kPoolMask;
        int slot = 0;
        if (active) {
            self = row_nb<kSub>(S, g, rc, sp);
            slot = c.A.slot[self.aidx];  // needed only at the end: loaded early
            load_bonds(c, self, bt);
            const int hx = S.sbin[sp] - t * g.HX;
            // stencil bin rows (ring row, bin row): above, own, below
            int subOf[3] = {0, 0, 0};
            if (kSub == 2) {
                ringOf[0] = t == 0 ? rm_ : rc;
                subOf[0] = t == 0 ? 1 : 0;
                ringOf[1] = rc;
                subOf[1] = t;
                ringOf[2] = t == 0 ? rc : rp;
                subOf[2] = t == 0 ? 1 : 0;
            }
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int b0 = subOf[k] * g.HX + hx;
                a0s[k] = S.rs[ringOf[k]][b0 - 1];
                ns[k] = S.rs[ringOf[k]][b0 + 2] - a0s[k];
            }
            if (ns[0] + ns[1] + ns[2] > 64) {
                direct = true;
            } else {
                const bool hasBonds = flags_num_conn(self.flags) > 0;
                // fluid: the cell's own density term needs no pair evaluation; only its place in
                // the visiting order is kept (a barrier cell meets itself as a barrier instead);
                // collision: the cell is not its own neighbour
                const bool keepSelf = kFluid && flags_barrier(self.flags);
                const int selfFlat = ns[0] + (su - a0s[1]);  // the cell's own visiting index
                // predicated tests, one mask bit per candidate in visiting order (one copy of the
                // loop for chunks on the torus seam, one for the rest)
                auto test = [&](auto seamTag) {
                    constexpr bool kSeam = decltype(seamTag)::value;
                    auto dsq = [&](f2 o) {
                        return kFluid ? lensq(kSeam ? torus_delta(self.pos, o, Wf, Hf) : sub(self.pos, o))
                                      : lensq(sub(o, self.pos));  // R13
                    };
                    int cidx = 0;
#pragma unroll
                    for (int r = 0; r < 3; ++r) {
                        const int pb = a0s[r] & kPoolMask, n = ns[r];
                        unsigned long long rowMask = 0ull;
                        for (int u = 0; u < n; u += 2) {
                            // two candidates per iteration; the second is read even past the
                            // range end (within the mirrored pool) and trimmed below
                            const f2 p0 = to_f2(S.pos[pb + u]);
                            const f2 p1 = to_f2(S.pos[pb + u + 1]);
                            rowMask |= dsq(p0) <= P.reachSq ? (1ull << u) : 0ull;
                            rowMask |= dsq(p1) <= P.reachSq ? (2ull << u) : 0ull;
                        }
                        rowMask &= n >= 64 ? ~0ull : (1ull << n) - 1ull;
                        mask |= rowMask << cidx;
                        cidx += n;
                    }
                };
                if (g.seam)
                    test(std::true_type{});
                else
                    test(std::false_type{});
                if (!keepSelf) mask &= ~(1ull << selfFlat);
                if (kFluid && !keepSelf) selfIdx = __builtin_popcountll(mask & ((1ull << selfFlat) - 1ull));
                if (hasBonds) {
                    // which of the neighbours are bonded
                    const bool bari = flags_barrier(self.flags);
                    unsigned long long m = mask;
                    while (m) {
                        const int ci = __builtin_ctzll(m);
                        m &= m - 1ull;
                        const int r = ci < ns[0] ? 0 : (ci < ns[0] + ns[1] ? 1 : 2);
                        const int rest = ci - (r == 0 ? 0 : (r == 1 ? ns[0] : ns[0] + ns[1]));
                        const int np = (pick3(a0s, r) + rest) & kPoolMask;
                        if (!is_bonded(bt, S.aidx[np] - w.cellOff)) continue;
                        if constexpr (!kFluid) {
                            // a bonded collision pair adds only its overlap push (closer than
                            // minDistance, not for a barrier cell): otherwise it adds nothing and
                            // leaves the list (the distance as the pair term takes it)
                            const f2 o = to_f2(S.pos[np]);
                            const float d2 = lensq(g.seam ? torus_delta(self.pos, o, Wf, Hf) : sub(self.pos, o));
                            if (bari || d2 > P.overlapBelowSq) {
                                mask &= ~(1ull << ci);
                                continue;
                            }
                        }
                        bmask |= 1ull << ci;
                    }
                }
            }
        }
        const int count = __builtin_popcountll(mask);
        const int inc = int(wave_inclusive_scan(uint32_t(count)));
        const int numPairs = wave_last(inc);
        const int pbase = inc - count;

        Acc a;
        if (active) acc_init(a, self);
        for (int W0 = 0; W0 < numPairs; W0 += kPairCap) {
            const int wend = min(numPairs, W0 + kPairCap);
            // owners write the records of their pairs inside this window
            if (count > 0 && pbase < wend && pbase + count > W0) {
                unsigned long long m = mask;
                for (int p = pbase; m; ++p) {
                    const int ci = __builtin_ctzll(m);
                    m &= m - 1ull;
                    if (p < W0 || p >= wend) continue;
                    int r = 0, rest = ci;
                    if (rest >= ns[0]) {
                        rest -= ns[0];
                        r = 1;
                        if (rest >= ns[1]) {
                            rest -= ns[1];
                            r = 2;
                        }
                    }
                    const int np = (pick3(a0s, r) + rest) & kPoolMask;
                    const uint32_t bonded = uint32_t((bmask >> ci) & 1ull);
                    S.u.pr.pself[p - W0] = uint8_t(sp);
                    S.u.pr.pnb[p - W0] = uint16_t((bonded << 15) | (uint32_t(pick3(ringOf, r)) << 8) | uint32_t(np));
                }
            }
            __syncthreads();
            for (int off = W0; off < wend; off += kRowThreads) {
                // pair terms, one pair per lane
                const int p = off + lane;
                float tm[kPairTerms];
#pragma unroll
                for (int v = 0; v < kPairTerms; ++v) tm[v] = 0.0f;
                uint32_t f = 0u;
                if (p < wend) {
                    const uint32_t rec = S.u.pr.pnb[p - W0];
                    const int ps = S.u.pr.pself[p - W0], ro = int((rec >> 8) & 3u), po = int(rec & 0xFFu);
                    const Nb s = row_nb<kSub, RS, false>(S, g, rc, ps);
                    const Nb o = row_nb<kSub, RS, false>(S, g, ro, po);
                    f = pair_term<kFluid>(c, w, P, wi, s, o, (rec >> 15) != 0u, tm, g.seam, [&] {
                        return make_int2(row_lidx<kSub>(S, g, rc, ps), row_lidx<kSub>(S, g, ro, po));
                    });
                }
                // the owner adds its pairs of this round in visiting order, reading each from the
                // lane that evaluated it (every lane takes part in every exchange)
                const int lo = max(pbase, off), hi = min(pbase + count, min(off + kRowThreads, wend));
                const int n = count > 0 ? max(0, hi - lo) : 0;
                for (int j = 0; __any(j < n); ++j) {
                    const int src = (lo + j - off) & (kRowThreads - 1);
                    float tt[kPairTerms];
#pragma unroll
                    for (int v = 0; v < kPairTerms; ++v) tt[v] = __shfl(tm[v], src, kRowThreads);
                    const uint32_t ff = uint32_t(__shfl(int(f), src, kRowThreads));
                    const bool on = j < n;
                    if constexpr (kFluid) a.rho = on && lo + j - pbase == selfIdx ? a.rho + w0 : a.rho;
                    acc_add_sel<kFluid>(a, on, ff, tt);
                }
            }
            __syncthreads();
        }
        if (active && selfIdx == count) a.rho = a.rho + w0;  // the own term after all pairs
        if (active) {
            // the cell's full flags, from the start-of-step arrays (the map copy holds a bin's
            // entries in arrival order, not at their logical indices)
            self.flags = c.A.flags[self.aidx];
            if (direct) {
                atomicAdd(&c.sc->directCells, 1ull);
                cell_direct<kFluid>(c, w, P, wi, self, [&](auto&& visit) {
#pragma unroll
                    for (int r = 0; r < 3; ++r)
                        for (int u = a0s[r]; u < a0s[r] + ns[r]; ++u)
                            visit(row_nb<kSub>(S, g, ringOf[r], u & kPoolMask));
                });
            } else {
                // the bond targets are loaded again rather than kept in registers through the pair
                // work (few cells have bonds in a fluid world; the loads hit the cache)
                int bt2[kMaxBonds];
                load_bonds(c, self, bt2);
                cell_finish<kFluid>(c, w, P, wi, self, slot, bt2, a);
            }
        }
    }
}

// A chunk with a row too dense for LDS: every cell walks its stencil in global memory, entries of
// a bin by increasing previous index (selection over the bin), same order and results.
template <bool kFluid>
__device__ __forceinline__ void rows_direct(const Ctx& c, const DevWorld& w, const DevParams& P, int wi, const StripGeom& g, int y0,
                            int IY) {
    const uint32_t base = g.base;
    auto binRange = [&](int gb, int& s0, int& n) {
        s0 = g.cellOff + int(c.start[gb] - base);
        n = int(c.start[gb + 1] - c.start[gb]);
    };
    for (int y = y0; y < y0 + IY; ++y) {
        const int rowFirst = g.binOff + y * g.nbx + g.x0;
        const int m0 = g.cellOff + int(c.start[rowFirst] - base);
        const int m1 = g.cellOff + int(c.start[rowFirst + g.IX] - base);
        for (int m = m0 + int(threadIdx.x); m < m1; m += kRowThreads) {
            // the bin of map position m: the last bin of the row segment starting at or before m
            int lo = 0, hi = g.IX;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (g.cellOff + int(c.start[rowFirst + mid] - base) <= m)
                    lo = mid;
                else
                    hi = mid;
            }
            const int hx = lo + 1;
            // (the map holds every bin's cells by previous index: a cell's logical index is its map
            // position)
            const int i = c.sorted[m];
            const Nb self = global_nb<kFluid>(c, i, m);
            cell_direct<kFluid>(c, w, P, wi, self, [&](auto&& visit) {
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        int s0, n;
                        binRange(strip_bin(g, y + dy, hx + dx), s0, n);
                        for (int t = 0; t < n; ++t) visit(global_nb<kFluid>(c, c.sorted[s0 + t], s0 + t));
                    }
            });
        }
    }
}

// Whether staged rows y0-1 .. y0+IY (kSub bin rows each) each fit a row in flight and every three
// consecutive ones the pool (the same answer on every lane).
template <int kSub>
__device__ __forceinline__ bool stage_fits(const Ctx& c, const StripGeom& g, int y0, int IY) {
    const int lane = threadIdx.x;
    int total = 0;
    if (lane < IY + 2) {
        const int Y = y0 - 1 + lane;
#pragma unroll
        for (int t = 0; t < kSub; ++t) {
            const int yy = kSub * Y + t;
            const int rowFirst = g.binOff + wrap_small(yy, g.nby) * g.nbx + g.x0;
            total += int(c.start[rowFirst + g.IX] - c.start[rowFirst]);
            const int gl = strip_bin(g, yy, 0), gr = strip_bin(g, yy, g.IX + 1);
            total += int(c.start[gl + 1] - c.start[gl]) + int(c.start[gr + 1] - c.start[gr]);
        }
    }
    const int three = total + __shfl_down(total, 1, kRowThreads) + __shfl_down(total, 2, kRowThreads);
    return !__any(total > kRowCap || (lane < IY && three > kPool));
}

// One chunk of IY staged rows from staged row y0 (kSub bin rows each); false (and nothing done)
// when its staged rows do not fit the LDS.
template <bool kFluid, int kSub, class RS>
__device__ __forceinline__ bool chunk_rows(const Ctx& c, const DevWorld& w, const DevParams& P, int wi,
                                           StripGeom& g, RS& S, int y0, int IY) {
    const int nbyS = g.nby / kSub;
    // Away from the seam no staged bin wraps, and two stencil neighbours lie less than three bin
    // edges apart (the last column and row absorb less than one edge of remainder) while half the
    // world is at least four: the torus difference is the plain difference there, bit for bit.
    g.seam = g.x0 == 0 || g.x0 + g.IX >= g.nbx || y0 == 0 || y0 + IY >= nbyS || g.nbx < 8 || g.nby < 8;
    if (!stage_fits<kSub>(c, g, y0, IY)) return false;

    RowFlight<kSub> F, G;
    int head = 0;  // pool position of the next committed row
    // rows y0-1, y0, y0+1 into ring rows 0, 1, 2
    for (int r = 0; r < 3; ++r) {
        row_issue_counts<kSub>(c, g, y0 - 1 + r, F);
        row_issue_entries<kSub>(c, g, S, F);
        row_commit<kSub>(g, S, F, r, head);
        head += F.total;
    }
    if (IY >= 2) row_issue_counts<kSub>(c, g, y0 + 2, F);
    for (int y = y0; y < y0 + IY; ++y) {
        const int rc = (y - y0 + 1) % 3;
        const int rm_ = (rc + 2) % 3, rp = (rc + 1) % 3;
        const bool next = y + 2 <= y0 + IY;  // row y+2 is the lower halo of row y+1
        if (next) {
            row_issue_entries<kSub>(c, g, S, F);
            if (y + 3 <= y0 + IY) row_issue_counts<kSub>(c, g, y + 3, G);
        }
        row_compute<kFluid, kSub>(c, w, P, wi, g, S, rm_, rc, rp);
        if (next) {
            row_commit<kSub>(g, S, F, rm_, head);  // row y+2 replaces row y-1
            head += F.total;
#pragma unroll
            for (int q = 0; q < RowFlight<kSub>::kSlots; ++q) {
                F.s0[q] = G.s0[q];
                F.s1[q] = G.s1[q];
            }
        }
    }
    return true;
}

// A collision world staged two bin rows at a time (sparse rows) runs a chunk whose two-row stages
// do not fit as single bin rows over the same bins.
// The fluid kernel is held to 128 VGPRs, four waves per SIMD (its 10.2 KB of LDS allow 16
// workgroups per CU); it ran 4 % faster than at 135 VGPRs and three waves.  The collision kernel
// needs ~146 and would spill more than it gains.
template <bool kFluid>
__global__ void __launch_bounds__(kRowThreads) __attribute__((amdgpu_waves_per_eu(kFluid ? 4 : 1)))
k_neighbor_rows(Ctx c) {
    __shared__ RowShared<kFluid ? 1 : kSubMax> S;
    const int2 tb = c.tileBlocks[blockIdx.x];
    if (tb.x < 0) return;
    const int wi = tb.x;
    // the world and its parameters are read once, before any store: read through references,
    // every use after a global store would load them again
    const DevWorld w = c.worlds[wi];
    const DevParams P = c.params[wi];
    if ((P.motion == 0) != kFluid) return;
    const int strip = tb.y % w.ntx, chunk = tb.y / w.ntx;
    StripGeom g;
    g.x0 = strip * w.tileX;
    g.IX = min(w.tileX, w.nbx - g.x0);
    g.HX = g.IX + 2;
    g.nbx = w.nbx;
    g.nby = w.nby;
    g.binOff = w.binOff;
    g.cellOff = w.cellOff;
    g.base = c.start[w.binOff];
    g.seam = true;
    const int y0 = chunk * w.tileY;  // in staged rows of w.rowsPerStage bin rows
    const int IY = min(w.tileY, w.nby / w.rowsPerStage - y0);
    const int by0 = w.rowsPerStage * y0, bIY = w.rowsPerStage * IY;  // the chunk in bin rows
    bool done = false;
    // (fluid worlds are always staged one bin row at a time: the two-row variant would cost the
    // fluid kernel a third of its resident waves in registers)
    if constexpr (!kFluid)
        if (w.rowsPerStage == 2) done = chunk_rows<kFluid, 2>(c, w, P, wi, g, S, y0, IY);
    if (!done) done = chunk_rows<kFluid, 1>(c, w, P, wi, g, S, by0, bIY);
    if (!done) {
        // too crowded for the LDS staging: every cell walks its stencil in global memory, and the
        // owner-counted part of the region's bins is empty
        if (threadIdx.x == 0) atomicAdd(&c.sc->fallbackChunks, 1ull);
        rows_direct<kFluid>(c, w, P, wi, g, by0, bIY);
    }
}

// ---------------------------------------------------------------------------------------------
// Cell-parallel fluid pass.  The map is in logical order (cells by bin, inside a bin by previous
// index), so the three stencil bins of a cell in one bin row are one contiguous range of the map,
// and a wave of 64 consecutive map positions reads its candidates through it (the ranges of
// neighbouring lanes overlap and stay in the caches; the start-of-step state is nearly in map
// order, so the reads through the map are nearly sequential).  Per wave: every lane tests its
// candidates into a mask in visiting order; the wave's neighbour pairs are listed back to back in
// LDS and evaluated one pair per lane; every lane then adds its own pairs' terms in visiting order
// (the float sums of a sequential stencil walk, bit for bit) and finishes its cell.  A cell in the
// first or last bin column (its stencil wraps around the torus there), one with more than 64
// candidates, and every cell of a grid narrower than three bins walks its stencil alone
// (cell_direct, same order and results).  Workgroups of 1..4 waves (the cell block size), each
// wave on its own with its own LDS (wave-level synchronisation only); block b
// runs table entry (b mod 8) * ceil(T / 8) + b / 8, so each XCD sweeps one contiguous band of the
// map and the stencil rows its waves share stay in its L2.
// ---------------------------------------------------------------------------------------------
constexpr int kCellWaves = 4;  // waves per workgroup at most (256-thread groups: 64- and 128-thread
                               // groups ran 2.3x and 1.4x slower, 512-thread ones 9 % slower)
constexpr int kCellRec = 256;  // pair records of one wave at a time

constexpr int kCellStage = 192;  // staged candidate positions per stencil row (three per lane)

struct CellShared {
    // the wave's own cells (the self side of its pairs)
    float2 pos[64];
    float2 vel[64];
    float dens[64];
    uint32_t flags[64];
    int aidx[64];
    union {
        // while candidates are tested: positions of the union of the lanes' ranges, per stencil
        // row, and one entry past the end (read by the second test of an iteration that passes a
        // list's end; its result is trimmed)
        float2 cpos[3 * kCellStage + 1];
        struct {
            // pair records: neighbour map position, owner lane | bonded << 7
            int rnb[kCellRec];
            uint8_t rown[kCellRec];
            // results of the pairs evaluated in one round: terms and the term flags
            float res[kPairTerms + 1][64];
        } pr;
    } u;
};

// the cell at map position m (fluid: 1 / density) as a neighbour
__device__ __forceinline__ Nb map_nb(const Ctx& c, int m) {
    Nb o;
    o.aidx = c.sorted[m];
    o.pos = to_f2(c.A.pos[o.aidx]);
    o.vel = to_f2(c.A.vel[o.aidx]);
    o.dens = 1.0f / c.A.dens[o.aidx];
    o.flags = c.A.flags[o.aidx];
    o.lidx = m;
    return o;
}

// (at most 80 VGPRs: six waves per SIMD)
__global__ void __launch_bounds__(64 * kCellWaves) __attribute__((amdgpu_waves_per_eu(6)))
k_neighbor_cells(Ctx c, int tableBlocks) {
    __shared__ CellShared SS[kCellWaves];
    const int per = (tableBlocks + kXcdCount - 1) / kXcdCount;
    const int e = int(blockIdx.x % kXcdCount) * per + int(blockIdx.x / kXcdCount);
    if (e >= tableBlocks) return;
    const int2 blk = c.cellBlocks[e];
    const int wi = blk.x;
    const DevWorld w = c.worlds[wi];
    const DevParams P = c.params[wi];
    if (P.motion != 0) return;
    const int lane = int(threadIdx.x) & 63, wave = int(threadIdx.x) >> 6;
    CellShared& S = SS[wave];
    const int m = blk.y + int(threadIdx.x);
    const int mEnd = w.cellOff + w.numCells;
    if (__builtin_amdgcn_readfirstlane(blk.y + wave * 64) >= mEnd) return;  // the whole wave is past the world
    const bool active = m < mEnd;
    const float Wf = float(w.W), Hf = float(w.H);
    const float w0 = sph_w(0.0f) * P.invH2;  // a cell's own density term
    const uint32_t base = c.start[w.binOff];
    auto S0 = [&](int b) { return w.cellOff + int(c.start[w.binOff + b] - base); };

    Nb self;
    self.pos = {0.0f, 0.0f};
    self.vel = {0.0f, 0.0f};
    self.dens = 0.0f;
    self.flags = 0u;
    self.aidx = 0;
    self.lidx = m;
    int bx = 0, by = 0;
    if (active) {
        self = map_nb(c, m);
        cell_bin_xy(w, self.pos, bx, by);
    }
    S.pos[lane] = to_float2(self.pos);
    S.vel[lane] = to_float2(self.vel);
    S.dens[lane] = self.dens;
    S.flags[lane] = self.flags;
    S.aidx[lane] = self.aidx;
    // stencil rows above, own, below: one map range each (bins bx-1 .. bx+1)
    int lo[3] = {0, 0, 0}, ns[3] = {0, 0, 0};
    bool direct = active && (bx == 0 || bx == w.nbx - 1);
    if (active && !direct) {
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int rb = wrap_small(by + r - 1, w.nby) * w.nbx;
            lo[r] = S0(rb + bx - 1);
            ns[r] = S0(rb + bx + 2) - lo[r];
        }
        if (ns[0] + ns[1] + ns[2] > 64) direct = true;
    }
    if (direct) ns[0] = ns[1] = ns[2] = 0;
    // away from the seam rows (and in grids of at least eight bins each way) the plain difference
    // equals the torus difference bit for bit: two stencil neighbours lie less than three bin
    // edges apart, less than half the world
    const bool seam = __any(active && (by == 0 || by == w.nby - 1)) || w.nbx < 8 || w.nby < 8;
    unsigned long long mask = 0ull, bmask = 0ull;
    int selfIdx = -1;  // pairs before the cell's own density term
    // The union of the lanes' ranges of each stencil row is staged in LDS (all its loads in flight
    // together) when it fits; a wave whose cells span two bin rows tests from global memory.
    const bool tests = active && !direct;
    // The staged range of each stencil row, [u0, u0 + un): from the first testing lane's range start
    // to the last one's end (in one bin row both grow with the lane), used when every testing
    // lane's range lies inside (a wave across two bin rows fails that).
    int u0[3], un[3];
    const unsigned long long testers = __ballot(tests);
    bool staged = testers != 0ull;
    if (staged) {
        const int first = __builtin_ctzll(testers), last = 63 - __builtin_clzll(testers);
        bool inside = true;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            u0[r] = __builtin_amdgcn_readlane(lo[r], first);
            un[r] = __builtin_amdgcn_readlane(lo[r] + ns[r], last) - u0[r];
            inside = inside && (!tests || (lo[r] >= u0[r] && lo[r] + ns[r] <= u0[r] + un[r]));
            staged = staged && un[r] >= 0 && un[r] <= kCellStage;
        }
        staged = staged && __all(inside);
    }
    if (!staged) u0[0] = u0[1] = u0[2] = un[0] = un[1] = un[2] = 0;
    if (staged) {
        int a[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) a[r][q] = lane + 64 * q < un[r] ? c.sorted[u0[r] + lane + 64 * q] : -1;
        float2 v[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) v[r][q] = a[r][q] >= 0 ? c.A.pos[a[r][q]] : float2{0.0f, 0.0f};
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) S.u.cpos[r * kCellStage + lane + 64 * q] = v[r][q];
    }
    __syncwarp();
    if (tests) {
        // the three ranges as one list in visiting order: candidate t lies in row 0 below n0, in
        // row 1 below n01, else in row 2 (the wave runs its longest list, not its longest row)
        const int n0 = ns[0], n01 = ns[0] + ns[1], nt = n01 + ns[2];
        // candidate t lies at t plus its row's offset: in the staged rows (read as one array of
        // three rows) or in the map
        const int so0 = lo[0] - u0[0], so1 = kCellStage + lo[1] - u0[1] - n0, so2 = 2 * kCellStage + lo[2] - u0[2] - n01;
        const int go0 = lo[0], go1 = lo[1] - n0, go2 = lo[2] - n01;
        const int sd1 = so1 - so0, sd2 = so2 - so1, gd1 = go1 - go0, gd2 = go2 - go1;
        auto test = [&](auto seamTag, auto stagedTag) {
            constexpr bool kSeam = decltype(seamTag)::value;
            constexpr bool kStaged = decltype(stagedTag)::value;
            auto cand = [&](int t) -> f2 {
                if constexpr (kStaged) {
                    // (t <= nt: at most one entry past the last row's range, inside cpos; the
                    // offsets in bytes, so that a row select is not scaled per candidate)
                    int xb = 8 * t + 8 * so0;
                    xb += t >= n0 ? 8 * sd1 : 0;
                    xb += t >= n01 ? 8 * sd2 : 0;
                    return to_f2(*reinterpret_cast<const float2*>(reinterpret_cast<const char*>(S.u.cpos) + xb));
                } else {
                    int x = t + go0;
                    x += t >= n0 ? gd1 : 0;
                    x += t >= n01 ? gd2 : 0;
                    return to_f2(c.A.pos[c.sorted[min(x, mEnd - 1)]]);
                }
            };
            // two candidates per iteration (the second may pass the list's end: trimmed below),
            // their bits set in the half of the mask that holds t (a uniform loop bound per half:
            // a wave's lists rarely pass 32 candidates)
            auto half = [&](int t0, int t1, uint32_t& m32) {
                for (int t = t0; t < t1; t += 2) {
                    const f2 p0 = cand(t), p1 = cand(t + 1);
                    const float d0 = lensq(kSeam ? torus_delta(self.pos, p0, Wf, Hf) : sub(self.pos, p0));
                    const float d1 = lensq(kSeam ? torus_delta(self.pos, p1, Wf, Hf) : sub(self.pos, p1));
                    const uint32_t bit = 1u << (t & 31);
                    m32 |= (d0 <= P.reachSq ? bit : 0u) | (d1 <= P.reachSq ? bit << 1 : 0u);
                }
            };
            uint32_t lo32 = 0u, hi32 = 0u;
            half(0, min(nt, 32), lo32);
            half(32, nt, hi32);
            mask = (static_cast<unsigned long long>(hi32) << 32) | lo32;
            mask &= nt >= 64 ? ~0ull : (1ull << nt) - 1ull;
        };
        if (staged) {
            if (seam)
                test(std::true_type{}, std::true_type{});
            else
                test(std::false_type{}, std::true_type{});
        } else {
            if (seam)
                test(std::true_type{}, std::false_type{});
            else
                test(std::false_type{}, std::false_type{});
        }
        // a barrier cell meets itself as a barrier neighbour; any other cell adds its own density
        // term at its place in the visiting order, without a pair
        const int selfFlat = ns[0] + (m - lo[1]);
        if (!flags_barrier(self.flags)) {
            mask &= ~(1ull << selfFlat);
            selfIdx = __builtin_popcountll(mask & ((1ull << selfFlat) - 1ull));
        }
        if (flags_num_conn(self.flags) > 0) {
            // (the bond targets are loaded here and again at the end rather than held through the
            // pair work: few cells of a fluid world have bonds, and the loads hit the cache)
            int bt[kMaxBonds];
            load_bonds(c, self, bt);
            unsigned long long mm = mask;
            while (mm) {
                const int ci = __builtin_ctzll(mm);
                mm &= mm - 1ull;
                const int r = ci < ns[0] ? 0 : (ci < ns[0] + ns[1] ? 1 : 2);
                const int rest = ci - (r == 0 ? 0 : (r == 1 ? ns[0] : ns[0] + ns[1]));
                const int jm = (r == 0 ? lo[0] : (r == 1 ? lo[1] : lo[2])) + rest;
                if (is_bonded(bt, c.sorted[jm] - w.cellOff)) bmask |= 1ull << ci;
            }
        }
    }
    const int count = __builtin_popcountll(mask);
    const int inc = int(wave_inclusive_scan(uint32_t(count)));
    const int numPairs = wave_last(inc);
    const int pbase = inc - count;
    __syncwarp();  // (the pair records reuse the staged positions' LDS)

    Acc a;
    acc_init(a, self);
    for (int R0 = 0; R0 < numPairs; R0 += kCellRec) {
        const int rend = min(numPairs, R0 + kCellRec);
        // owners list their pairs of this window, in visiting order
        if (count > 0 && pbase < rend && pbase + count > R0) {
            unsigned long long mm = mask;
            for (int p = pbase; mm; ++p) {
                const int ci = __builtin_ctzll(mm);
                mm &= mm - 1ull;
                if (p < R0 || p >= rend) continue;
                const int r = ci < ns[0] ? 0 : (ci < ns[0] + ns[1] ? 1 : 2);
                const int rest = ci - (r == 0 ? 0 : (r == 1 ? ns[0] : ns[0] + ns[1]));
                S.u.pr.rnb[p - R0] = (r == 0 ? lo[0] : (r == 1 ? lo[1] : lo[2])) + rest;
                S.u.pr.rown[p - R0] = uint8_t(lane | int((bmask >> ci) & 1ull) << 7);
            }
        }
        __syncwarp();
        for (int off = R0; off < rend; off += 64) {
            // pair terms, one pair per lane
            const int p = off + lane;
            float tm[kPairTerms];
#pragma unroll
            for (int v = 0; v < kPairTerms; ++v) tm[v] = 0.0f;
            uint32_t f = 0u;
            if (p < rend) {
                const int jm = S.u.pr.rnb[p - R0];
                const int ow = S.u.pr.rown[p - R0];
                const int ol = ow & 63;
                Nb s;
                s.pos = to_f2(S.pos[ol]);
                s.vel = to_f2(S.vel[ol]);
                s.dens = S.dens[ol];
                s.flags = S.flags[ol];
                s.aidx = S.aidx[ol];
                s.lidx = blk.y + wave * 64 + ol;
                const Nb o = map_nb(c, jm);
                f = pair_term<true>(c, w, P, wi, s, o, (ow >> 7) != 0, tm, seam,
                                    [&] { return make_int2(s.lidx, o.lidx); });
            }
#pragma unroll
            for (int v = 0; v < kPairTerms; ++v) S.u.pr.res[v][lane] = tm[v];
            S.u.pr.res[kPairTerms][lane] = __uint_as_float(f);
            __syncwarp();
            // every owner adds its pairs of this round in visiting order
            const int jlo = max(pbase, off), jhi = min(pbase + count, min(off + 64, rend));
            for (int j = jlo; j < jhi; ++j) {
                const int q = j - off;
                float tt[kPairTerms];
#pragma unroll
                for (int v = 0; v < kPairTerms; ++v) tt[v] = S.u.pr.res[v][q];
                const uint32_t ff = __float_as_uint(S.u.pr.res[kPairTerms][q]);
                if (j - pbase == selfIdx) a.rho = a.rho + w0;
                acc_add<true>(a, ff, tt);
            }
            __syncwarp();
        }
    }
    if (!active) return;
    if (direct) {
        atomicAdd(&c.sc->directCells, 1ull);
        cell_direct<true>(c, w, P, wi, self, [&](auto&& visit) {
            for (int dy = -1; dy <= 1; ++dy) {
                const int rb = wrap_small(by + dy, w.nby) * w.nbx;
                for (int dx = -1; dx <= 1; ++dx) {
                    const int b = rb + wrap_small(bx + dx, w.nbx);
                    const int s1 = S0(b + 1);
                    for (int j = S0(b); j < s1; ++j) visit(map_nb(c, j));
                }
            }
        });
        return;
    }
    if (selfIdx == count) a.rho = a.rho + w0;  // the own term after all pairs
    int bt[kMaxBonds];
    load_bonds(c, self, bt);
    cell_finish<true>(c, w, P, wi, self, c.A.slot[self.aidx], bt, a, w.binOff + by * w.nbx + bx);
}

int neighbor_blocks_per_cu() {
    int n = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_neighbor_rows<false>, kRowThreads, 0);
    return e == hipSuccess ? std::max(1, n) : 1;
}

void launch_neighbor(const Ctx& c, const LaunchShape& sh, hipStream_t s, bool fluid, bool collision) {
    if (fluid && sh.cellBlocks) {
        const int grid = (sh.cellBlocks + kXcdCount - 1) / kXcdCount * kXcdCount;
        hipLaunchKernelGGL(k_neighbor_cells, dim3(grid), dim3(sh.cellThreads), 0, s, c, sh.cellBlocks);
    }
    if (collision && sh.tileBlocks)
        hipLaunchKernelGGL(k_neighbor_rows<false>, dim3(sh.tileBlocks), dim3(kRowThreads), 0, s, c);
}

}  // namespace alien
