// k_bonds.hip — bonded cells and structural operations.
//
// Connection forces (CellProcessor::calcConnectionForces, source/EngineGpuKernels/CellProcessor.cuh:
// 408-491) run twice per step around the Verlet position update (:525-544) and before the velocity
// update (:546-559), over the staged cells (cells with bonds, and barrier cells).  Angle forces that
// the reference adds to the neighbours with float atomics are written by the owner into per-bond
// slots and gathered by the receiver in its own bond order (R7).  Structural operations follow
// CellConnectionProcessor (source/EngineGpuKernels/CellConnectionProcessor.cuh): fusion pairs
// claim their cells and bond neighbours and apply only when every claim holds (R5); bond deletion
// works from per-cell flags (R6).
#include "kernels_common.hpp"
#include "kernels.hpp"

namespace alien {

// The kernels below run one thread per cell of the worlds that can have bonded or barrier cells
// (bond block table) and act on the staged cells: those the neighbour pass found bonded or barrier
// (stagedMark), which is the set CellProcessor's bonded-cell passes cover.  Cells bonded by this
// step's fusion are not among them.
__device__ __forceinline__ bool staged_cell(const Ctx& c, int& x, int& wi) {
    const int2 blk = c.bondBlocks[blockIdx.x];
    wi = blk.x;
    x = blk.y + int(threadIdx.x);
    const DevWorld& w = c.worlds[wi];
    return x < w.cellOff + w.numCells && c.stagedMark[x] != 0;
}

// bond targets of the staged cells from start-of-step to end-of-step indices
__global__ void k_remap(Ctx c) {
    const int cap = c.cellCap;
    int x, wi;
    if (!staged_cell(c, x, wi)) return;
    {
        const DevWorld& w = c.worlds[wi];
        const int nc = flags_num_conn(c.B.flags[x]);
        // all targets first, then all lookups, then the stores (so that the loads overlap)
        int t[kMaxBonds], ni[kMaxBonds];
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b) t[b] = b < nc ? c.B.btgt[b * cap + x] : 0;
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b) ni[b] = b < nc ? c.newidx[w.cellOff + t[b]] : 0;
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b)
            if (b < nc) c.B.btgt[b * cap + x] = ni[b] - w.cellOff;
    }
}

// bond vector from x to o.  In the second pass the cells stand at pos + shift; the vector is taken
// as the start-of-pass difference plus the difference of the shifts, which keeps the FP32 result
// close to the exact one (rounding pos + shift first would cost up to ulp(world size)).
__device__ __forceinline__ f2 bond_vector(const Ctx& c, int o, f2 px, f2 sx, int pass, float Wf, float Hf) {
    const f2 d = torus_delta(to_f2(c.B.pos[o]), px, Wf, Hf);
    if (pass == 0) return d;
    return add(d, sub(to_f2(c.shift[o]), sx));
}

// connection forces of the staged cells: own force plus per-bond angle forces for the partners;
// pass 1 also runs checkConnections (CellProcessor.cuh:493-523) on the moved positions.  The
// partners' data of all the cell's bonds is loaded up front, so the loads are in flight together
// (lanes skip the slots beyond their cell's bonds: 7 % faster than loading all six).
// (at most 80 VGPRs, six waves per SIMD: the partner loads are latency-bound)
__global__ void __attribute__((amdgpu_waves_per_eu(6))) k_conn_own(Ctx c, int pass) {
    const int cap = c.cellCap;
    int x, wi;
    if (!staged_cell(c, x, wi)) return;
    {
        const DevWorld& w = c.worlds[wi];
        const DevParams& P = c.params[wi];
        const uint32_t fx = c.B.flags[x];
        const int nc = flags_num_conn(fx);
        if (nc == 0 || flags_barrier(fx)) {
            c.valid[x] = 0u;
            if (pass == 1) c.stretched[x] = 0;  // checkConnections skips barrier cells
            return;
        }
        const float Wf = float(w.W), Hf = float(w.H);
        const bool considerAngles = (c.ctr[wi].timestep % 3ull) == 0ull;
        const f2 px = to_f2(c.B.pos[x]);
        const f2 shx = pass == 0 ? f2{0.0f, 0.0f} : to_f2(c.shift[x]);
        const float sx = c.bstiff[x];
        const float kx2 = sx * sx;
        int tg[kMaxBonds];
        f2 disp[kMaxBonds];
        float kc2[kMaxBonds], rest[kMaxBonds];
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b) tg[b] = b < nc ? w.cellOff + c.B.btgt[b * cap + x] : x;
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b) {
            const int o = tg[b];
            kc2[b] = 0.0f;
            rest[b] = 0.0f;
            disp[b] = f2{0.0f, 0.0f};
            if (b < nc) {
                const float so = c.bstiff[o];
                kc2[b] = so * so;
                rest[b] = c.B.brest[b * cap + x];
                disp[b] = bond_vector(c, o, px, shx, pass, Wf, Hf);
            }
        }
        // the bond before the first is the last one
        f2 prevDisp = disp[0];
        int lastTarget = tg[0];
#pragma unroll
        for (int b = 1; b < kMaxBonds; ++b)
            if (b == nc - 1) {
                prevDisp = disp[b];
                lastTarget = tg[b];
            }
        f2 force = {0.0f, 0.0f};
        uint32_t v = 0u;
        bool tooLong = false;
        const float maxBind = P.maxBindingDistance[flags_color(fx) % kColors];
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b) {
            if (b >= nc) break;
            const int o = tg[b];
            const f2 d = disp[b];
            const float ad = len2(d);
            if (pass == 1 && ad > maxBind) tooLong = true;  // checkConnections on new positions
            const float dev = ad - rest[b];
            force = add(force, dvd(mul(mul(unit(d), dev), kx2 + kc2[b]), 6.0f));
            bool skipped = false;  // a bond whose actual angle is negative keeps the previous vector
            if (considerAngles && (nc > 2 || (nc == 2 && b == 0))) {
                const int lc = b == 0 ? lastTarget : tg[b > 0 ? b - 1 : 0];
                const uint32_t fo = c.B.flags[o];
                const int no = flags_num_conn(fo);
                bool tri = false;
                for (int j = 0; j < no; ++j)
                    if (w.cellOff + c.B.btgt[j * cap + o] == lc) {
                        tri = true;
                        break;
                    }
                if (!tri) {
                    const float angle = angle_of(d);
                    const float prevAngle = angle_of(prevDisp);
                    const float act = sub_angle(angle, prevAngle);
                    if (act < 0.0f) {
                        skipped = true;
                    } else {
                        const float ref = c.B.bang[b * cap + x];
                        const float strength = fabsf(ref - act) / 2000.0f * kx2;
                        f2 f1 = mul(dvd(unit(d), fmaxf(len2(d), P.minDistance)), strength);
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
            }
            if (!skipped) prevDisp = d;
        }
        c.own[x] = to_float2(force);
        c.valid[x] = v;
        if (tooLong) c.delFlag[x] = 1;
        if (pass == 1) c.stretched[x] = tooLong ? 1 : 0;
    }
}

// own force plus the angle forces the partners wrote for this cell, in its bond order (R7); the
// partners' markers are loaded up front
__device__ __forceinline__ f2 gather_conn(const Ctx& c, const DevWorld& w, int x, int nc) {
    const int cap = c.cellCap;
    f2 acc = {0.0f, 0.0f};
    acc = add(acc, to_f2(c.own[x]));
    const int xl = x - w.cellOff;
    int pt[kMaxBonds];
    uint32_t vps[kMaxBonds];
#pragma unroll
    for (int q = 0; q < kMaxBonds; ++q) pt[q] = q < nc ? w.cellOff + c.B.btgt[q * cap + x] : x;
#pragma unroll
    for (int q = 0; q < kMaxBonds; ++q) vps[q] = q < nc ? c.valid[pt[q]] : 0u;
#pragma unroll
    for (int q = 0; q < kMaxBonds; ++q) {
        const uint32_t vp = vps[q];
        if (vp == 0u) continue;  // the partner wrote no angle force this pass (2 of 3 steps)
        const int p = pt[q];
        const int np = flags_num_conn(c.B.flags[p]);
        int b = -1;
        for (int kk = 0; kk < np; ++kk)
            if (c.B.btgt[kk * cap + p] == xl) {
                b = kk;
                break;
            }
        if (b < 0) continue;
        if (vp & (1u << b)) acc = add(acc, to_f2(c.out1[b * cap + p]));
        const int b2 = (b + 1) % np;
        if (vp & (1u << (6 + b2))) acc = add(acc, to_f2(c.out2[b2 * cap + p]));
    }
    return acc;
}

// Verlet position step (CellProcessor.cuh:525-544): the displacement is stored and applied after
// the second force pass
__global__ void k_conn_verlet_pos(Ctx c) {
    int x, wi;
    if (!staged_cell(c, x, wi)) return;
    {
        const DevWorld& w = c.worlds[wi];
        const DevParams& P = c.params[wi];
        const uint32_t fx = c.B.flags[x];
        const float dt = P.dt;
        const f2 v = to_f2(c.B.vel[x]);
        if (flags_barrier(fx)) {
            c.shift[x] = to_float2(mul(v, dt));
            return;
        }
        const f2 s1 = gather_conn(c, w, x, flags_num_conn(fx));
        c.shift[x] = to_float2(add(mul(v, dt), mul(mul(mul(s1, dt), dt), 0.5f)));  // x/2 == x*0.5 exactly
        c.s2[x] = to_float2(s1);
    }
}

// Verlet velocity step, friction and decay flag (CellProcessor.cuh:546-559, 707-723, 786-790); the
// moved position is written here, and its bin recorded for the next step's map
__global__ void k_conn_verlet_vel(Ctx c) {
    int x, wi;
    const bool st = staged_cell(c, x, wi);
    const DevWorld& w = c.worlds[wi];
    const DevParams& P = c.params[wi];
    const float dt = P.dt, friction = P.friction, maxEnergy = P.maxBindingEnergy;
    // on an inner-friction step the velocity is left before friction for k_inner_friction
    const bool inner = P.innerFriction > 0.0f && (c.ctr[wi].timestep % 3ull) == 0ull;
    // every load before the first store, so that they are in flight together
    uint32_t fx = 0u;
    f2 p = {0.0f, 0.0f}, sh = {0.0f, 0.0f}, v0 = {0.0f, 0.0f}, s2 = {0.0f, 0.0f}, s1 = {0.0f, 0.0f};
    float energy = 0.0f;
    if (st) {
        fx = c.B.flags[x];
        p = to_f2(c.B.pos[x]);
        sh = to_f2(c.shift[x]);
        v0 = to_f2(c.B.vel[x]);
        s2 = to_f2(c.s2[x]);
        energy = c.energy[w.cellOff + c.B.slot[x]];
        if (!flags_barrier(fx)) s1 = gather_conn(c, w, x, flags_num_conn(fx));
    }
    if (!st) return;
    const f2 p2 = torus_pos(add(p, sh), w.W, w.H);
    c.B.pos[x] = to_float2(p2);
    // (the map gives the cell's start-of-step index at its end-of-step index)
    bin_next(c, w, x, p2, cell_bin_of(w, to_f2(c.A.pos[c.sorted[x]])));
    if (flags_barrier(fx)) return;
    const f2 acc = mul(add(s1, s2), 0.5f);
    f2 v = add(v0, mul(acc, dt));
    if (inner) {
        c.s2[x] = to_float2(v);  // the start-of-pass velocity of the inner friction
    } else {
        v = mul(v, 1.0f - friction);
    }
    c.B.vel[x] = to_float2(v);
    if (energy > maxEnergy) c.delFlag[x] = 1;
}

// applyInnerFriction (CellProcessor.cuh:678-705) every third step, then applyFriction (:707-723).
// Jacobi (rule R11): each bond is relaxed twice for the cell, once for its own visit and once for
// the partner's visit of the bond, toward the partner's start-of-pass velocity, in bond order;
// barrier partners are skipped.
__global__ void k_inner_friction(Ctx c) {
    const int cap = c.cellCap;
    int x, wi;
    if (!staged_cell(c, x, wi)) return;
    const DevParams& P = c.params[wi];
    if (!(P.innerFriction > 0.0f) || (c.ctr[wi].timestep % 3ull) != 0ull) return;
    const uint32_t fx = c.B.flags[x];
    if (flags_barrier(fx)) return;
    const DevWorld& w = c.worlds[wi];
    const int nc = flags_num_conn(fx);
    const float f = P.innerFriction;
    f2 vp[kMaxBonds];
    bool use[kMaxBonds];
#pragma unroll
    for (int b = 0; b < kMaxBonds; ++b) {
        use[b] = false;
        vp[b] = f2{0.0f, 0.0f};
        if (b < nc) {
            const int p = w.cellOff + c.B.btgt[b * cap + x];
            use[b] = !flags_barrier(c.B.flags[p]);
            if (use[b]) vp[b] = to_f2(c.s2[p]);
        }
    }
    f2 v = to_f2(c.s2[x]);
    for (int visit = 0; visit < 2; ++visit)
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b)
            if (b < nc && use[b]) {
                const f2 avg = dvd(add(v, vp[b]), 2.0f);
                v = add(mul(v, 1.0f - f), mul(avg, f));
            }
    v = mul(v, 1.0f - P.friction);
    c.B.vel[x] = to_float2(v);
}

// ---------------------------------------------------------------------------------------------
// living states of the worlds that track them (rule R17, cellDeathConsequences None), one thread
// per end-of-step cell.  The three phases alternate between the two state buffers (living at the
// start of the step -> livingNext -> living -> livingNext, the next step's start).
// ---------------------------------------------------------------------------------------------
constexpr uint8_t kReady = 0, kUnderConstruction = 1, kActivating = 2, kDetaching = 3, kReviving = 4, kDying = 5;

__device__ __forceinline__ bool living_cell(const Ctx& c, int& x, int& wi) {
    const int2 blk = c.livingBlocks[blockIdx.x];
    wi = blk.x;
    x = blk.y + int(threadIdx.x);
    const DevWorld& w = c.worlds[wi];
    return x < w.cellOff + w.numCells;
}

// checkConnections' Detaching marks on the partners of cells with an overstretched bond
// (CellProcessor.cuh:515-520), then aging (:561-594)
__global__ void k_living_mark_age(Ctx c) {
    const int cap = c.cellCap;
    int x, wi;
    if (!living_cell(c, x, wi)) return;
    const DevWorld& w = c.worlds[wi];
    const uint32_t fx = c.B.flags[x];
    const int sl = w.cellOff + c.B.slot[x];
    uint8_t st = c.living[sl];
    if (w.bondWork) {  // (stretched marks exist for the staged cells of worlds with bond work)
        const int nc = flags_num_conn(fx);
        for (int b = 0; b < nc; ++b)
            if (c.stretched[w.cellOff + c.B.btgt[b * cap + x]]) st = kDetaching;
    }
    if (!flags_barrier(fx)) {
        const uint32_t act = c.activation[sl];
        if (st == kReady && act > 0u) c.activation[sl] = act - 1u;
    }
    c.livingNext[sl] = st;
}

// the state transitions (CellProcessor.cuh:596-676) from the states after the marks
__global__ void k_living_transition(Ctx c) {
    const int cap = c.cellCap;
    int x, wi;
    if (!living_cell(c, x, wi)) return;
    const DevWorld& w = c.worlds[wi];
    const uint32_t fx = c.B.flags[x];
    const int sl = w.cellOff + c.B.slot[x];
    const uint8_t st = c.livingNext[sl];
    uint8_t next = st;
    if (flags_barrier(fx)) {
        next = kReady;
    } else if (st == kActivating || st == kReviving || st == kDetaching) {
        next = kReady;  // Detaching -> Ready under cellDeathConsequences None
    } else if (st == kUnderConstruction) {
        const uint32_t cr = c.creature[sl];
        const int nc = flags_num_conn(fx);
        for (int b = 0; b < nc; ++b) {
            const int ps = w.cellOff + c.B.slot[w.cellOff + c.B.btgt[b * cap + x]];
            if (c.creature[ps] == cr && c.livingNext[ps] == kActivating) next = kActivating;
        }
    }
    c.living[sl] = next;
}

// decay (CellProcessor.cuh:775-858): energy below cellMinEnergy -> Dying; the partners of a cell
// that becomes Dying -> Detaching unless they are or become Dying (cellDeathProbability 0)
__device__ __forceinline__ bool below_min_energy(const Ctx& c, const DevWorld& w, const DevParams& P, uint32_t f, int sl) {
    return !flags_barrier(f) && c.energy[sl] < P.minEnergy[flags_color(f)];
}

__global__ void k_living_decay(Ctx c) {
    const int cap = c.cellCap;
    int x, wi;
    if (!living_cell(c, x, wi)) return;
    const DevWorld& w = c.worlds[wi];
    const DevParams& P = c.params[wi];
    const uint32_t fx = c.B.flags[x];
    const int sl = w.cellOff + c.B.slot[x];
    const uint8_t st = c.living[sl];
    uint8_t next = st;
    if (below_min_energy(c, w, P, fx, sl)) {
        next = kDying;
    } else if (st != kDying) {
        const int nc = flags_num_conn(fx);
        for (int b = 0; b < nc; ++b) {
            const int p = w.cellOff + c.B.btgt[b * cap + x];
            const int ps = w.cellOff + c.B.slot[p];
            if (below_min_energy(c, w, P, c.B.flags[p], ps) && c.living[ps] != kDying) next = kDetaching;
        }
    }
    c.livingNext[sl] = next;
}

// A set without bond work has no bonded or barrier cell: aging with the activation count
// (CellProcessor.cuh:561-594), the transitions (:596-676; UnderConstruction stays without an
// Activating partner) and the decay to Dying (:775-858; no partner to mark Detaching) of the three
// kernels above need no partner, and one pass per cell gives their end-of-step state.
__global__ void k_living_unbonded(Ctx c) {
    int x, wi;
    if (!living_cell(c, x, wi)) return;
    const DevWorld& w = c.worlds[wi];
    const DevParams& P = c.params[wi];
    const uint32_t fx = c.B.flags[x];
    const int sl = w.cellOff + c.B.slot[x];
    uint8_t st = c.living[sl];
    if (!flags_barrier(fx)) {
        const uint32_t act = c.activation[sl];
        if (st == kReady && act > 0u) c.activation[sl] = act - 1u;
    }
    if (flags_barrier(fx) || st == kActivating || st == kReviving || st == kDetaching) st = kReady;
    if (below_min_energy(c, w, P, fx, sl)) st = kDying;
    c.livingNext[sl] = st;
}

void launch_living_transitions(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec) {
    if (!sh.livingBlocks) return;
    if (rec) rec->mark("living_mark_age");
    hipLaunchKernelGGL(k_living_mark_age, dim3(sh.livingBlocks), dim3(sh.cellThreads), 0, s, c);
    if (rec) rec->mark("living_transition");
    hipLaunchKernelGGL(k_living_transition, dim3(sh.livingBlocks), dim3(sh.cellThreads), 0, s, c);
}

void launch_living_decay(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec) {
    if (!sh.livingBlocks) return;
    if (rec) rec->mark("living_decay");
    hipLaunchKernelGGL(k_living_decay, dim3(sh.livingBlocks), dim3(sh.cellThreads), 0, s, c);
}

// ---------------------------------------------------------------------------------------------
// structural operations (CellConnectionProcessor.cuh), R5/R6
// ---------------------------------------------------------------------------------------------
__global__ void k_mark(Ctx c) {
    const int cap = c.cellCap;
    int x, wi;
    if (!staged_cell(c, x, wi)) return;
    {
        // the cell or a bond partner flagged for deletion; the targets first, then their flags, so
        // that the loads of all bonds are in flight together
        const DevWorld& w = c.worlds[wi];
        const int nc = flags_num_conn(c.B.flags[x]);
        int t[kMaxBonds];
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b) t[b] = b < nc ? c.B.btgt[b * cap + x] : 0;
        uint32_t m = c.delFlag[x];
#pragma unroll
        for (int b = 0; b < kMaxBonds; ++b)
            if (b < nc) m |= c.delFlag[w.cellOff + t[b]];
        c.mark[x] = m ? 1 : 0;
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
        const f2 cd = torus_delta(to_f2(c.B.pos[w.cellOff + c.B.btgt[c1]]), p1, Wf, Hf);
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
        prevAngle = angle_of(torus_delta(to_f2(c.B.pos[w.cellOff + c.B.btgt[prevIndex * cap + c1]]), p1, Wf, Hf));
        nextAngle = angle_of(torus_delta(to_f2(c.B.pos[w.cellOff + c.B.btgt[index * cap + c1]]), p1, Wf, Hf));
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
        const f2 posDelta = torus_delta(to_f2(c.B.pos[b]), to_f2(c.B.pos[a]), float(w.W), float(w.H));
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
    const int cap = c.cellCap;
    int x, wi;
    if (!staged_cell(c, x, wi)) return;
    {
        if (!c.mark[x]) return;
        const DevWorld& w = c.worlds[wi];
        const int nc = flags_num_conn(c.B.flags[x]);
        int snap[kMaxBonds];
        for (int b = 0; b < nc; ++b) snap[b] = c.B.btgt[b * cap + x];
        const bool self = c.delFlag[x] != 0;
        for (int b = 0; b < nc; ++b)
            if (self || c.delFlag[w.cellOff + snap[b]]) delete_one_way(c, x, snap[b]);
    }
}

// last kernel of the step: clears the flags of the staged cells and advances every world's
// timestep (SimulationCudaFacade.cu:140-143)
__global__ void k_clear_flags(Ctx c) {
    if (blockIdx.x == 0)
        for (int wi = threadIdx.x; wi < c.numWorlds; wi += blockDim.x) c.ctr[wi].timestep += 1ull;
    int x, wi;
    if (!staged_cell(c, x, wi)) return;
    c.mark[x] = 0;
    c.delFlag[x] = 0;
}

void launch_bonds(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec) {
    const dim3 sg(sh.bondBlocks), cg(sh.candGrid);
#define L(name, kernel, grid, threads, ...)                           \
    do {                                                              \
        if (rec) rec->mark(name);                                     \
        hipLaunchKernelGGL(kernel, grid, threads, 0, s, __VA_ARGS__); \
    } while (0)
    // no world can have a bonded or barrier cell this step (none has one and none can fuse): no
    // cell is staged, and the scan at the start of the step advanced the timesteps
    if (!sh.hasBondWork || sh.bondBlocks == 0) {
        if (sh.livingBlocks) L("living_unbonded", k_living_unbonded, dim3(sh.livingBlocks), dim3(sh.cellThreads), c);
        return;
    }
    const dim3 b(sh.cellThreads), cb(256);
    L("remap", k_remap, sg, b, c);
    L("conn_own", k_conn_own, sg, b, c, 0);
    L("conn_verlet_pos", k_conn_verlet_pos, sg, b, c);
    L("conn_own", k_conn_own, sg, b, c, 1);
    L("conn_verlet_vel", k_conn_verlet_vel, sg, b, c);
    // SimulationKernelsLauncher.cu:52-71: aging and living states, inner friction, friction and decay
    launch_living_transitions(c, sh, s, rec);
    if (sh.hasInnerFriction) L("inner_friction", k_inner_friction, sg, b, c);
    launch_living_decay(c, sh, s, rec);
    L("mark", k_mark, sg, b, c);
    L("fuse_claim", k_fuse_claim, cg, cb, c);
    L("fuse_apply", k_fuse_apply, cg, cb, c);
    L("fuse_reset", k_fuse_reset, cg, cb, c);
    L("delete", k_delete, sg, b, c);
    L("clear_flags", k_clear_flags, sg, b, c);
#undef L
}

}  // namespace alien

