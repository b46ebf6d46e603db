// k_particles.hip — energy particles: movement, particle-particle fusion, absorption into cells and
// reflection at barrier cells (ALIEN RadiationProcessor::movement/collision,
// source/EngineGpuKernels/RadiationProcessor.cuh:61-170).
//
// The reference resolves concurrent fusions and absorptions with tryLock and skips on contention.
// Here the outcome is fixed (R8): a particle targets the winner of its 1x1 slot in the particle map
// (lowest index, pre-move position); it fuses only into a target that is not fusing itself, and a
// sink accepts the lowest-index fusing particle; a cell accepts the lowest-index absorbing
// particle.  Claims are atomicMin on indices, which is order independent.
#include "kernels_common.hpp"
#include "kernels.hpp"

namespace alien {

// start-of-step position of particle map entry src (a live particle or a staging slot)
__device__ __forceinline__ f2 pre_move_pos(const Ctx& c, int src) {
    if (c.partStage > 0 && src >= c.partStage) return to_f2(c.staged[src - c.partStage].pos);
    if (c.splitStage > 0 && src >= c.splitStage) return to_f2(c.splitStaged[src - c.splitStage].pos);
    return to_f2(c.A.ppos[src]);
}

// The particle map winner of the slot of position p (R2): the first particle of the bin whose
// pre-move slot matches; particle k fuses into it when it is another particle closer than sqrt(0.5)
// after the move.  Returns the target or -1.
__device__ __forceinline__ int particle_target(const Ctx& c, const DevWorld& w, int k, f2 p) {
    int sx, sy;
    slot_of(p, w.W, w.H, sx, sy);
    int s, e;
    part_bin_range(c, w, part_bin_local(w, sx, sy), s, e);
    int t = -1;
    for (int m = s; m < e; ++m) {
        int qx, qy;
        slot_of(pre_move_pos(c, c.psorted[m]), w.W, w.H, qx, qy);
        if (qx == sx && qy == sy) {
            t = m;
            break;
        }
    }
    return t >= 0 && t != k && lensq(sub(p, to_f2(c.B.ppos[t]))) < 0.5f ? t : -1;
}

// Targets and fusion claims in one pass: a particle claims its target when the target does not
// fuse itself, which it finds by computing the target's own target (the same function of the same
// moved positions; fusing particles are rare).
__global__ void k_part_target(Ctx c) {
    for_live_particles(c, [&](int, const DevWorld& w, int k) {
        const f2 p = to_f2(c.B.ppos[k]);
        const int want = particle_target(c, w, k, p);
        c.want[k] = want;
        if (want >= 0 && particle_target(c, w, want, to_f2(c.B.ppos[want])) < 0) atomicMin(&c.fclaim[want], k);
        // cell hit this step: lowest-index cell of the slot at map-build time (cellMap.getFirst),
        // the first of the bin's cells in map order (a bin holds its cells by previous index)
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
    });
}

__global__ void k_part_apply(Ctx c) {
    for_live_particles(c, [&](int wi, const DevWorld& w, int k) {
        const DevParams& P = c.params[wi];
        const int q = c.fclaim[k];
        if (q != kClaimFree) {
            // sink: merge the lowest-index particle fusing into it (RadiationProcessor.cuh:86-93)
            c.fclaim[k] = kClaimFree;
            const float eq = c.B.pen[q], ek = c.B.pen[k];
            if (eq > kNearZero && ek > kNearZero) {
                const float f1 = eq / (eq + ek);
                const f2 vq = to_f2(c.B.pvel[q]), vk = to_f2(c.B.pvel[k]);
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
        const int cm = c.absorbCell[k];
        if (cm < 0) return;
        // the cell as the neighbour pass found it (through the map: a barrier cell does not move
        // or change speed in that pass, and no cell changes its flags or slot there)
        const int ca = c.sorted[cm];
        const uint32_t fc = c.A.flags[ca];
        if (flags_barrier(fc)) {  // reflection (:99-106)
            const f2 vb = to_f2(c.A.vel[ca]);
            const f2 vr = sub(to_f2(c.B.pvel[k]), vb);
            const f2 rr = torus_delta(to_f2(c.B.ppos[k]), to_f2(c.A.pos[ca]), float(w.W), float(w.H));
            const float d = dot2(vr, rr);
            if (d < 0.0f) {
                const float trs = fmaxf(0.1f, lensq(rr));
                c.B.pvel[k] = to_float2(add(sub(vr, dvd(mul(mul(rr, 2.0f), d), trs)), vb));
            }
            c.absorbCell[k] = -1;
            return;
        }
        if (c.B.plast[k] == c.A.slot[ca] || P.absorption[flags_color(fc) % kColors] < kNearZero) {
            c.absorbCell[k] = -1;
            return;
        }
        atomicMin(&c.cclaim[cm], k);
    });
}

// absorption (:108-165); also bins the surviving particles for the next step
// RadiationProcessor::splitting (RadiationProcessor.cuh:172-204), after absorption: with
// probability 0.01 (R9 draw keyed by the particle id) a particle above particleSplitEnergy of its
// colour halves its energy and splits in two, 1/5 of a random unit vector u either side of its
// position, their velocities its own -/+ u radiationVelocityPerturbation / (E/2 + 1).  The new half
// goes to the split staging slot of the particle's index and into the next step's histogram, so
// that the next map sorts it behind the particles that lived, in index order (R18); its id is
// kCreatedParticleId | kSplitParticleId | timestep << 32 | local index.
__device__ __forceinline__ void split_particle(const Ctx& c, const DevWorld& w, int wi, int k) {
    const DevParams& P = c.params[wi];
    const unsigned long long id = c.B.pid[k];
    const unsigned long long ts = c.ctr[wi].timestep - (c.timestepAdvanced ? 1ull : 0ull);
    if (!(uniform_draw(w.seed, ts, id, kStreamSplit) < 0.01f)) return;
    const int col = c.B.pflags[k] & kParticleColorMask;
    float e = c.B.pen[k];
    if (!(e > P.splitEnergy[col % kColors])) return;
    e = e * 0.5f;
    f2 u = to_f2(unit_vector_of_angle(uniform_draw(w.seed, ts, id, kStreamSplitAngle) * 360.0f));
    const f2 p = to_f2(c.B.ppos[k]);
    const f2 d = {u.x / 5.0f, u.y / 5.0f};
    const f2 other = torus_pos(add(p, d), w.W, w.H);
    const f2 self = torus_pos(sub(p, d), w.W, w.H);
    u = mul(u, P.radVelPert / (e + 1.0f));
    const f2 v = to_f2(c.B.pvel[k]);
    c.B.pen[k] = e;
    c.B.ppos[k] = to_float2(self);
    c.B.pvel[k] = to_float2(sub(v, u));
    StagedParticle r;
    r.pos = to_float2(other);
    r.vel = to_float2(add(v, u));
    r.energy = e;
    r.color = uint32_t(col) | kStagedPending;
    r.id = kCreatedParticleId | kSplitParticleId | ((ts & 0x3fffffffull) << 32) | (unsigned long long)uint32_t(k - w.partOff);
    c.splitStaged[k] = r;
    pbin_for_next_step(c, w, c.splitStage + k, other, c.countNext);
}

__global__ void k_part_absorb(Ctx c) {
    for_live_particles(c, [&](int wi, const DevWorld& w, int k) {
        const int cm = c.absorbCell[k];
        if (cm >= 0 && c.cclaim[cm] == k) {
            c.cclaim[cm] = kClaimFree;
            const DevParams& P = c.params[wi];
            const int ca = c.sorted[cm];
            const uint32_t fc = c.A.flags[ca];
            const float ra = P.absorption[flags_color(fc) % kColors];
            float e = c.B.pen[k];
            float transfer = e * ra;
            if (e < 0.01f) transfer = e;
            const int cslot = c.A.slot[ca];
            float* ce = &c.energy[w.cellOff + cslot];
            *ce = *ce + transfer;
            e = e - transfer;
            c.B.pen[k] = e;
            if (e < kNearZero)
                c.B.pflags[k] = uint8_t(c.B.pflags[k] | kParticleDead);
            else
                c.B.plast[k] = cslot;
        }
        if (c.B.pflags[k] & kParticleDead) {
            c.pbinOf[k] = -1;
            return;
        }
        if (c.splitStage > 0 && c.params[wi].splits) split_particle(c, w, wi, k);
        pbin_for_next_step(c, w, k, to_f2(c.B.ppos[k]), c.countNext);
    });
}

// Cell radiation at the start of a step (CellProcessor::radiation, CellProcessor.cuh:725-773, run
// by the reference after the cell map update and before the forces): a non-barrier cell radiates
// with probability radiationProb (R9 draw); its factor is highRadiationFactor above
// highRadiationMinCellEnergy plus radiationCellAgeStrength above radiationMinCellAge; it loses
// 2 u E factor / radiationProb (at most E - 1, only if E > 1), emitted as an energy particle with
// the cell's velocity times radiationVelocityMultiplier plus a unit vector of random angle times
// radiationVelocityPerturbation, 1.5 ahead of the cell along it and one velocity behind (the move
// of this step brings it forward), unless the energy is below NEAR_ZERO
// (RadiationProcessor::createEnergyParticle, RadiationProcessor.cuh:248-364, without radiation
// sources and external energy).  The particle of cell i goes to the staging slot partStage + i and
// into this step's particle histogram, so that the map sorts it behind the bin's live particles,
// in cell order (R18); its id is kCreatedParticleId | timestep << 32 | local cell index (R18).
__global__ void k_radiation(Ctx c) {
    const int2 blk = c.cellBlocks[blockIdx.x];
    const int wi = blk.x;
    const DevWorld& w = c.worlds[wi];
    const int i = blk.y + int(threadIdx.x);
    if (i >= w.cellOff + w.numCells) return;
    const DevParams& P = c.params[wi];
    if (!P.radiates) return;
    const uint32_t f = c.A.flags[i];
    if (flags_barrier(f)) return;
    const int sl = w.cellOff + c.A.slot[i];
    const unsigned long long id = c.cid[sl];
    const unsigned long long ts = c.ctr[wi].timestep;
    if (!(uniform_draw(w.seed, ts, id, kStreamRadiation) < P.radProb)) return;
    const int col = flags_color(f) % kColors;
    const float e = c.energy[sl];
    float factor = 0.0f;
    if (e > P.highRadMinEnergy[col]) factor = factor + P.highRadFactor[col];
    const long long age = (long long)(c.ageBase[sl] + uint32_t(ts - w.ageTs));
    if (age > (long long)P.radMinAge[col]) factor = factor + P.radAgeStrength[col];
    if (!(factor > 0.0f)) return;
    float loss = e * factor;
    loss = loss / P.radProb;
    loss = 2.0f * loss * uniform_draw(w.seed, ts, id, kStreamRadiationEnergy);
    if (!(e > 1.0f)) return;
    const f2 u = to_f2(unit_vector_of_angle(uniform_draw(w.seed, ts, id, kStreamRadiationAngle) * 360.0f));
    const f2 v = add(mul(to_f2(c.A.vel[i]), P.radVelMul), mul(u, P.radVelPert));
    const f2 p = torus_pos(sub(add(to_f2(c.A.pos[i]), mul(unit(v), 1.5f)), v), w.W, w.H);
    if (loss > e - 1.0f) loss = e - 1.0f;
    if (loss > kNearZero) {
        StagedParticle r;
        r.pos = to_float2(p);
        r.vel = to_float2(v);
        r.energy = loss;
        r.color = uint32_t(col) & kParticleColorMask;
        r.id = kCreatedParticleId | ((ts & 0x3fffffffull) << 32) | (unsigned long long)uint32_t(i - w.cellOff);
        c.staged[i] = r;
        pbin_for_next_step(c, w, c.partStage + i, p, c.count);
    }
    c.energy[sl] = e - loss;
}

void launch_radiation(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec) {
    if (!sh.hasRadiation || !sh.cellBlocks) return;
    if (rec) rec->mark("radiation");
    hipLaunchKernelGGL(k_radiation, dim3(sh.cellBlocks), dim3(sh.cellThreads), 0, s, c);
}

// targets and fusion claims: the moved particles and the map
void launch_particles_pre(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec) {
    if (!sh.partBlocks || !sh.hasParticles) return;
    if (rec) rec->mark("part_target");
    hipLaunchKernelGGL(k_part_target, dim3(sh.partGrid), dim3(sh.partThreads), 0, s, c);
}

// fusions, reflections and absorptions: also the cells' end-of-pass state
void launch_particles_post(const Ctx& c, const LaunchShape& sh, hipStream_t s, Recorder* rec) {
    if (!sh.partBlocks || !sh.hasParticles) return;
    const dim3 g(sh.partGrid), b(sh.partThreads);
    if (rec) rec->mark("part_apply");
    hipLaunchKernelGGL(k_part_apply, g, b, 0, s, c);
    if (rec) rec->mark("part_absorb");
    hipLaunchKernelGGL(k_part_absorb, g, b, 0, s, c);
}

}  // namespace alien
