// alien_oracle.hpp — CPU oracle for the ALIEN physics timestep.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness reference ("oracle") for the HIP
// engine under alien_amd/.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it, and only as the checker or the timed CPU baseline.  The product never includes,
// links or calls anything under oracle/.
//
// It is a plain, sequential-semantics C++ restatement of ALIEN's step
// (source/EngineGpuKernels/SimulationKernelsLauncher.cu:25-91) templated on the real type, so the
// same code builds in FP32 (bit-level reference for the GPU) and FP64 (accuracy reference).
// Build with -ffp-contract=off -fno-fast-math: every floating-point operation below is a single
// IEEE operation in the order written.
//
// The reference is racy by design (float atomics, tryLock skips, atomic insertion order).  Every
// race is resolved by a fixed rule, each marked "R<n>" below:
//   R1  cell map: cells stably sorted by spatial bin at the start of every step; within a bin,
//       ascending index.  The sorted order becomes the engine's cell order for the step.
//   R2  particle map: winner of a 1x1 slot = lowest-index particle whose pre-move position is in it.
//   R3  collision forces gathered: F_i = sum over neighbours j of [f(i,j) - f(j,i)], visiting order.
//   R4  overlap correction is Jacobi: all distances from start-of-pass positions.
//   R5  fusion: candidate pairs (i<j) claim their cells and bond neighbours; a pair that does not
//       hold every claim is dropped for the step (the reference's tryLock skip), no retry.
//   R6  bond deletion: per-cell delete flags; a cell drops every bond to a flagged partner (or all
//       bonds if itself flagged) in bond-list order; fusion pairs touching such cells are dropped.
//   R7  angle forces: written by the owner into per-bond slots, gathered by the receiver in its
//       bond order after its own force.
//   R8  absorption: one absorbing particle per cell per step (lowest index), one fusing particle
//       per sink particle per step (lowest index); a sink does not absorb in the same step.
//   R9  random draws are counter-based (seed, timestep, cell id, purpose) instead of a shared
//       table cursor (CudaNumberGenerator.cuh:99-103); see uniformDraw.
//   R11 inner friction (CellProcessor.cuh:678-705) is Jacobi: every bond is relaxed twice per
//       pass for each of its cells (the visit of the cell and the visit of its partner, each of
//       which averages both velocities in the reference), each time toward the partner's velocity
//       at the start of the pass, in bond order; the reference's tryLock skips do not happen.
//   R17 living states under cellDeathConsequences = None (CellProcessor.cuh:493-523, 561-676,
//       775-858): a cell whose energy is below cellMinEnergy becomes Dying; the partners of a cell
//       that becomes Dying become Detaching unless they are or become Dying themselves (the
//       reference's atomicExch order is resolved that way).
//   R10 dead particles are removed with a stable compaction.
//   R18 cell radiation (CellProcessor.cuh:725-773) runs over the cells in the order of the last
//       step's map (the cells' order at the start of the step), before the step's maps are built;
//       the particles it creates join the step's particle map behind the live ones, in that order
//       (the reference appends them with atomic slot grabs).  A created particle's id is
//       kCreatedParticleId | timestep << 32 | index of the emitting cell (the reference takes ids
//       from a shared atomic counter).  Particle splitting (RadiationProcessor.cuh:172-204) runs
//       over the particles in map order after absorption; the halves it creates join the next
//       step's particle map behind the particles that lived, in that order, before that step's
//       radiation; their ids are kCreatedParticleId | kSplitParticleId | timestep << 32 | index.
//   R12 SPH sums gathered per cell in visiting order.
//   R13 collision neighbours use the uncorrected (seam-blind) distance like Map.cuh:189.
// Three rules change only the rounding of the FP32 build (in exact arithmetic they are the
// reference's formulas); R14 and R15 keep FP32 within 1e-5 of FP64 also on large worlds, R16 saves
// divisions:
//   R14 torus differences subtract the period from the operand beyond it first (exact), instead
//       of rounding a - b and then taking remainderf (Map.cuh:35-39).
//   R15 the second connection-force pass and checkConnections evaluate bond vectors at the moved
//       positions as (start difference) + (shift difference); positions are moved after it.
//   R16 unit vectors, the SPH normalisations (1/h, 1/h^2, 1/h^3), the density terms
//       (density/density^2 = 1/density) and the collision and viscosity denominators are applied
//       as products with the correctly rounded reciprocal (one division instead of several).
// Cells are never detached here (detached cells come from cell functions), so the detached rule
// of the neighbour predicate never applies and is not restated.
// Parity status: pinned by tests/test_oracle.py and tests/test_oracle_pinning.py against closed
// forms written from the reference formulas, invariants, brute force and the reference's own test
// scenarios (their docstrings say which test pins what).  Not pinned beyond its invariants: which
// of two conflicting fusions or absorptions wins (R5, R8), a definition of this build.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <vector>

#include "../include/alien_hip.h"

namespace oracle {

template <class R> struct K;
template <> struct K<float> {
    static constexpr float PI = 3.1415926535897932384626433832795f;  // Math.cuh:9
    static constexpr float RAD_TO_DEG = 180.0f / PI;                  // Math.cuh:11
    static constexpr float KNORM = 3.0f / (2.0f * PI);                // CellProcessor.cuh:103
    static constexpr float NEAR_ZERO = 0.00001f;                      // Macros.cuh:63
    static constexpr float TWO_THIRDS = 2.0f / 3.0f;                  // CellProcessor.cuh:96
    static constexpr float PIO2 = 1.57079637050628662109375f;         // nearest float to pi/2
};
template <> struct K<double> {
    static constexpr double PI = 3.14159265358979323846;
    static constexpr double RAD_TO_DEG = 180.0 / PI;
    static constexpr double KNORM = 3.0 / (2.0 * PI);
    static constexpr double NEAR_ZERO = double(0.00001f);
    static constexpr double TWO_THIRDS = 2.0 / 3.0;
    static constexpr double PIO2 = 1.57079632679489661923;
};

// ---------------------------------------------------------------------------------------------
// scalar / vector math (Math.cuh, Base.cuh)
// ---------------------------------------------------------------------------------------------
template <class R> struct V2 {
    R x, y;
};
template <class R> inline V2<R> operator+(V2<R> a, V2<R> b) { return {a.x + b.x, a.y + b.y}; }
template <class R> inline V2<R> operator-(V2<R> a, V2<R> b) { return {a.x - b.x, a.y - b.y}; }
template <class R> inline V2<R> operator*(V2<R> a, R m) { return {a.x * m, a.y * m}; }
template <class R> inline V2<R> operator/(V2<R> a, R m) { return {a.x / m, a.y / m}; }

// Base.cuh:83-90
template <class R> inline int floorInt(R v) {
    int r = static_cast<int>(v);
    if (R(r) > v) --r;
    return r;
}
template <class R> inline R dot(V2<R> a, V2<R> b) { return a.x * b.x + a.y * b.y; }           // Math.cuh:263
template <class R> inline R lengthSquared(V2<R> v) { return v.x * v.x + v.y * v.y; }          // Math.cuh:288
template <class R> inline R length(V2<R> v) { return std::sqrt(v.x * v.x + v.y * v.y); }     // Math.cuh:273
// Math.cuh:246-261: unit vector, (1,0) below NEAR_ZERO
template <class R> inline V2<R> normalized(V2<R> v) {  // R16: times the reciprocal length
    R len = std::sqrt(v.x * v.x + v.y * v.y);
    if (len > K<R>::NEAR_ZERO) {
        const R inv = R(1) / len;
        return {v.x * inv, v.y * inv};
    }
    return {R(1), R(0)};
}
template <class R> inline void rotateQuarterClockwise(V2<R>& v) {  // Math.cuh:157-162
    R t = v.x;
    v.x = -v.y;
    v.y = t;
}
template <class R> inline void rotateQuarterCounterClockwise(V2<R>& v) {  // Math.cuh:164-169
    R t = v.x;
    v.x = v.y;
    v.y = -t;
}

// arcsine.  FP32: argument reduction asin(a) = pi/2 - 2 asin(sqrt((1-a)/2)) for a > 1/2 and an
// odd polynomial x + x^3 P(x^2) on [0, 1/2] whose coefficients were least-squares fitted for this
// project (max abs error 1.7e-7 rad over [-1,1]).  FP64: std::asin.
inline float asin_r(float v) {
    float a = v < 0.0f ? -v : v;
    bool big = a > 0.5f;
    float z = big ? (1.0f - a) * 0.5f : a * a;
    float s = big ? std::sqrt(z) : a;
    float p = 0.037105727940797806f;
    p = p * z + 0.014683568850159645f;
    p = p * z + 0.0317583866417408f;
    p = p * z + 0.044520482420921326f;
    p = p * z + 0.07500503212213516f;
    p = p * z + 0.16666659712791443f;
    float r = s + (s * z) * p;
    if (big) r = K<float>::PIO2 - (r + r);
    return v < 0.0f ? -r : r;
}
inline double asin_r(double v) { return std::asin(v); }

// Math.cuh:140-155 (0 deg corresponds to (0,-1))
template <class R> inline R angleOfVector(V2<R> v) {
    if (length(v) < K<R>::NEAR_ZERO) return R(0);
    R nvy = -v.y / length(v);
    nvy = std::max(R(-1), std::min(R(1), nvy));
    R angleSin = asin_r(nvy) * K<R>::RAD_TO_DEG;
    if (v.x >= R(0)) return R(90) - angleSin;
    return angleSin + R(270);
}
// Math.cuh:300-310
template <class R> inline R subtractAngle(R a, R b) {
    R d = a - b;
    if (d > R(360)) d -= R(360);
    if (d < R(0)) d += R(360);
    return d;
}
// Math.cuh:223-236
template <class R> inline bool isAngleInBetween(R a1, R a2, R c) {
    if (a1 == a2 && a1 != c) return false;
    if (c < a1) {
        c += R(360);
        a2 += R(360);
    }
    if (a2 < c) a2 += R(360);
    return a2 - a1 < R(360);
}
// Math.cuh:363-395
template <class R> inline bool crossing(V2<R> s, V2<R> e, V2<R> os, V2<R> oe) {
    if ((s.x == os.x && s.y == os.y) || (s.x == oe.x && s.y == oe.y) || (e.x == os.x && e.y == os.y) ||
        (e.x == oe.x && e.y == oe.y))
        return false;
    V2<R> p1 = s, v1 = e - s, p2 = os, v2 = oe - os;
    R divisor = v2.x * v1.y - v2.y * v1.x;
    if (std::abs(divisor) < K<R>::NEAR_ZERO) return false;
    R mue = (v1.x * (p2.y - p1.y) - v1.y * (p2.x - p1.x)) / divisor;
    if (mue < -K<R>::NEAR_ZERO || mue > R(1) + K<R>::NEAR_ZERO) return false;
    R lambda;
    if (std::abs(v1.x) > K<R>::NEAR_ZERO)
        lambda = (p2.x - p1.x + mue * v2.x) / v1.x;
    else if (std::abs(v1.y) > K<R>::NEAR_ZERO)
        lambda = (p2.y - p1.y + mue * v2.y) / v1.y;
    else
        return false;
    return lambda >= K<R>::NEAR_ZERO && lambda <= R(1) - K<R>::NEAR_ZERO;
}

// Counter-based random numbers (R9).  The reference draws from a pre-generated table through a
// shared atomic cursor (CudaNumberGenerator.cuh:12-104), so its draws depend on thread timing; here
// a draw is a pure function of (world seed, timestep, cell id, purpose): Philox4x32-10 of Salmon,
// Moraes, Dror and Shaw ("Parallel random numbers: as easy as 1, 2, 3", SC 2011), counter
// {id lo, id hi, timestep lo, purpose}, key {seed lo, seed hi}; the first output word's upper 24
// bits give a float in [0, 1).
inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c[0];
        const uint64_t p1 = uint64_t(0xCD9E8D57u) * c[2];
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c[1] ^ k0, n1 = uint32_t(p1);
        const uint32_t n2 = uint32_t(p0 >> 32) ^ c[3] ^ k1, n3 = uint32_t(p0);
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}
constexpr uint32_t kStreamForceDecay = 1;  // CellProcessor.cuh:367-386
// CellProcessor::radiation (CellProcessor.cuh:725-773): the radiation draw, the energy fraction,
// the perturbation angle
constexpr uint32_t kStreamRadiation = 2, kStreamRadiationEnergy = 3, kStreamRadiationAngle = 4;
// RadiationProcessor::splitting (RadiationProcessor.cuh:172-204): the split draw, the split direction
constexpr uint32_t kStreamSplit = 5, kStreamSplitAngle = 6;
constexpr uint64_t kCreatedParticleId = uint64_t(1) << 63, kSplitParticleId = uint64_t(1) << 62;

// Math::unitVectorOfAngle (Math.cuh:134-138): {sin a, -cos a}, a in degrees in [0, 360).  The
// reference calls sinf/cosf; here the sine and cosine are their Taylor series (to x^13 and x^14,
// Horner form, coefficients 1/k! rounded to float) of the angle reduced to a quarter turn, so that
// the FP32 build and the GPU compute the same bits.
template <class R> inline V2<R> unitVectorOfAngle(R deg) {
    int q = int(deg / R(90));
    q = q < 0 ? 0 : (q > 3 ? 3 : q);
    const R x = (deg - R(90) * R(q)) * R(0.017453292f);
    const R x2 = x * x;
    const R sn = x * (R(1) + x2 * (R(-0.16666667f) +
                                   x2 * (R(0.008333334f) +
                                         x2 * (R(-0.0001984127f) +
                                               x2 * (R(2.7557319e-06f) + x2 * (R(-2.5052108e-08f) + x2 * R(1.6059044e-10f)))))));
    const R cs = R(1) + x2 * (R(-0.5f) +
                              x2 * (R(0.041666668f) +
                                    x2 * (R(-0.0013888889f) +
                                          x2 * (R(2.4801588e-05f) +
                                                x2 * (R(-2.7557319e-07f) + x2 * (R(2.0876757e-09f) + x2 * R(-1.1470746e-11f)))))));
    R s = sn, c = cs;
    if (q == 1) {
        s = cs;
        c = -sn;
    } else if (q == 2) {
        s = -sn;
        c = -cs;
    } else if (q == 3) {
        s = -cs;
        c = sn;
    }
    return {s, -c};
}
inline float uniformDraw(uint64_t seed, uint64_t timestep, uint64_t id, uint32_t purpose) {
    uint32_t c[4] = {uint32_t(id), uint32_t(id >> 32), uint32_t(timestep), purpose};
    philox4x32_10(c, uint32_t(seed), uint32_t(seed >> 32));
    return float(c[0] >> 8) * (1.0f / 16777216.0f);
}

// SPH cubic spline (CellProcessor.cuh:92-119)
template <class R> inline R calcKernel(R q) {
    R r;
    if (q < R(1))
        r = K<R>::TWO_THIRDS - q * q + R(0.5f) * q * q * q;
    else if (q < R(2)) {
        r = R(2) - q;
        r = r * r * r / R(6);
    } else
        r = R(0);
    r *= K<R>::KNORM;
    return r;
}
template <class R> inline R calcKernel_d(R q) {
    R r;
    if (q < R(1))
        r = R(-2) * q + R(1.5f) * q * q;
    else if (q < R(2))
        r = R(-0.5f) * (R(2) - q) * (R(2) - q);
    else
        r = R(0);
    r *= K<R>::KNORM;
    return r;
}

// Torus (Map.cuh:10-63).  correctDirection is IEEE remainder(x, size), exact for |x| < 1.5 size,
// which holds for differences of wrapped positions.
struct Torus {
    int W = 0, H = 0;
    void wrapInt(int& x, int& y) const {
        x = ((x % W) + W) % W;
        y = ((y % H) + H) % H;
    }
    template <class R> void correctPosition(V2<R>& p) const {  // Map.cuh:20-26
        int ix = floorInt(p.x), iy = floorInt(p.y);
        R fx = p.x - R(ix), fy = p.y - R(iy);
        wrapInt(ix, iy);
        p = {R(ix) + fx, R(iy) + fy};
    }
    // a - b on the torus (Map.cuh:35-39 applies remainderf to the rounded difference).  Here the
    // period is subtracted from the operand beyond it first: for wrapped coordinates that
    // subtraction is exact (Sterbenz), so the result is the correctly rounded torus difference,
    // also across the seam, where rounding a - b first loses up to ulp(size) / 2.
    template <class R> static R delta1(R a, R b, R size) {
        const R d = a - b;
        const R half = size / R(2);
        if (d > half) return (a - size) - b;
        if (d < -half) return a - (b - size);
        return d;
    }
    template <class R> V2<R> delta(V2<R> a, V2<R> b) const { return {delta1(a.x, b.x, R(W)), delta1(a.y, b.y, R(H))}; }
    template <class R> void slotOf(V2<R> p, int& sx, int& sy) const {  // Map.cuh:132-137
        sx = floorInt(p.x);
        sy = floorInt(p.y);
        wrapInt(sx, sy);
    }
};

// ---------------------------------------------------------------------------------------------
// state
// ---------------------------------------------------------------------------------------------
template <class R> struct Params {
    R timestepSize;
    int motionType;
    R smoothingLength, viscosityStrength, pressureStrength;
    R cellMaxCollisionDistance, cellRepulsionStrength;
    R friction, cellMaxVelocity, cellMinDistance, cellFusionVelocity, cellMaxBindingEnergy;
    R innerFriction, cellMaxForceDecayProb;
    R cellMaxBindingDistance[ALIEN_MAX_COLORS];
    R radiationAbsorption[ALIEN_MAX_COLORS];
    R cellMaxForce[ALIEN_MAX_COLORS];
    R cellMinEnergy[ALIEN_MAX_COLORS];
    R radiationProb, radiationVelocityMultiplier, radiationVelocityPerturbation;
    int radiationMinCellAge[ALIEN_MAX_COLORS];
    R highRadiationFactor[ALIEN_MAX_COLORS], highRadiationMinCellEnergy[ALIEN_MAX_COLORS];
    R radiationCellAgeStrength[ALIEN_MAX_COLORS];
    R particleSplitEnergy[ALIEN_MAX_COLORS];
    static Params from(const alien_params& p) {
        Params q;
        q.timestepSize = R(p.timestepSize);
        q.motionType = p.motionType;
        q.smoothingLength = R(p.smoothingLength);
        q.viscosityStrength = R(p.viscosityStrength);
        q.pressureStrength = R(p.pressureStrength);
        q.cellMaxCollisionDistance = R(p.cellMaxCollisionDistance);
        q.cellRepulsionStrength = R(p.cellRepulsionStrength);
        q.friction = R(p.friction);
        q.cellMaxVelocity = R(p.cellMaxVelocity);
        q.cellMinDistance = R(p.cellMinDistance);
        q.cellFusionVelocity = R(p.cellFusionVelocity);
        q.cellMaxBindingEnergy = R(p.cellMaxBindingEnergy);
        q.innerFriction = R(p.innerFriction);
        q.cellMaxForceDecayProb = R(p.cellMaxForceDecayProb);
        for (int c = 0; c < ALIEN_MAX_COLORS; ++c) {
            q.cellMaxBindingDistance[c] = R(p.cellMaxBindingDistance[c]);
            q.radiationAbsorption[c] = R(p.radiationAbsorption[c]);
            q.cellMaxForce[c] = R(p.cellMaxForce[c]);
            q.cellMinEnergy[c] = R(p.cellMinEnergy[c]);
            q.radiationMinCellAge[c] = p.radiationMinCellAge[c];
            q.highRadiationFactor[c] = R(p.highRadiationFactor[c]);
            q.highRadiationMinCellEnergy[c] = R(p.highRadiationMinCellEnergy[c]);
            q.radiationCellAgeStrength[c] = R(p.radiationCellAgeStrength[c]);
            q.particleSplitEnergy[c] = R(p.particleSplitEnergy[c]);
        }
        q.radiationProb = R(p.radiationProb);
        q.radiationVelocityMultiplier = R(p.radiationVelocityMultiplier);
        q.radiationVelocityPerturbation = R(p.radiationVelocityPerturbation);
        return q;
    }
};

template <class R> struct Bond {
    int target;  // logical index of the partner cell
    R rest;      // CellConnection::distance
    R angle;     // CellConnection::angleFromPrevious (degrees)
};

template <class R> struct Cell {  // Object.cuh:201-263, physics subset
    uint64_t id;
    V2<R> pos, vel;
    R energy, stiffness, density;
    uint8_t color, maxConn, numConn, barrier, livingState, cellFunction, ancestorMutationId, selected;
    uint32_t age, creatureId, mutationId, activationTime;
    float genomeComplexity;
    Bond<R> bonds[ALIEN_MAX_CELL_BONDS];
};

template <class R> struct Particle {  // Object.cuh:8-37
    uint64_t id;
    V2<R> pos, vel;
    R energy;
    uint8_t color, selected;
    uint64_t lastAbsorbedCellId;  // 0 = none (ids are never 0 once set)
    bool hasLast;
    bool alive;
};

inline int binSizeFor(float radius) {
    int s = static_cast<int>(std::ceil(radius));
    return s < 2 ? 2 : s;
}

template <class R> class World {
public:
    Torus torus;
    Params<R> p;
    alien_params rawParams;
    uint64_t timestep = 0;
    uint64_t seed = 0;  // key of the counter-based random draws (R9)
    std::vector<Cell<R>> cells;
    std::vector<Particle<R>> parts;
    bool densityResetPending = false;
    int numThreads = 1;

    // map of the last step (R1): bin geometry, per-cell bin in current order.  Energy particles use
    // their own grid of kParticleBin-unit bins (the last bin of a row/column absorbs the rest).
    static constexpr int kParticleBin = 8;
    int binSize = 2, nbx = 0, nby = 0, pnbx = 1, pnby = 1;
    std::vector<int> cellBinStart, partBinStart;
    std::vector<int> cellBinOfIndex;
    std::vector<V2<R>> startPos;  // cell positions at map-build time (current order)

    World(int W, int H, const alien_params& prm) {
        torus.W = W;
        torus.H = H;
        setParams(prm);
    }

    void setParams(const alien_params& prm) {
        if (!(prm.innerFriction >= 0.0f && prm.innerFriction <= 1.0f))
            throw std::invalid_argument("innerFriction must lie in [0, 1]");
        if (!(prm.cellMaxForceDecayProb >= 0.0f && prm.cellMaxForceDecayProb <= 1.0f))
            throw std::invalid_argument("cellMaxForceDecayProb must lie in [0, 1]");
        if (prm.motionType != ALIEN_MOTION_FLUID && prm.motionType != ALIEN_MOTION_COLLISION)
            throw std::invalid_argument("bad motionType");
        if (!(prm.radiationProb >= 0.0f && prm.radiationProb <= 1.0f))
            throw std::invalid_argument("radiationProb must lie in [0, 1]");
        rawParams = prm;
        p = Params<R>::from(prm);
        densityResetPending = true;  // SimulationCudaFacade.cu:667-669
    }

    float interactionRadius() const {
        return rawParams.motionType == ALIEN_MOTION_FLUID ? rawParams.smoothingLength * 2.0f
                                                          : rawParams.cellMaxCollisionDistance;
    }

    void computeGeometry() {
        binSize = binSizeFor(interactionRadius());
        nbx = torus.W / binSize;
        nby = torus.H / binSize;
        if (nbx < 3 || nby < 3) throw std::invalid_argument("world smaller than three bins per axis");
        pnbx = std::max(1, torus.W / kParticleBin);
        pnby = std::max(1, torus.H / kParticleBin);
    }
    int binOfSlot(int sx, int sy) const {
        int bx = std::min(sx / binSize, nbx - 1), by = std::min(sy / binSize, nby - 1);
        return by * nbx + bx;
    }
    int binOfPos(V2<R> q) const {
        int sx, sy;
        torus.slotOf(q, sx, sy);
        return binOfSlot(sx, sy);
    }
    int pbinOfSlot(int sx, int sy) const {
        int bx = std::min(sx / kParticleBin, pnbx - 1), by = std::min(sy / kParticleBin, pnby - 1);
        return by * pnbx + bx;
    }
    int pbinOfPos(V2<R> q) const {
        int sx, sy;
        torus.slotOf(q, sx, sy);
        return pbinOfSlot(sx, sy);
    }

    // ---- boundary (DataAccessKernels.cu / ObjectFactory.cuh:46-86) -----------------------------
    void setData(const alien_cell_to* c, uint64_t n, const alien_particle_to* q, uint64_t np) {
        cells.assign(n, Cell<R>{});
        for (uint64_t i = 0; i < n; ++i) {
            auto& d = cells[i];
            const auto& s = c[i];
            d.id = s.id;
            d.pos = {R(s.pos[0]), R(s.pos[1])};
            torus.correctPosition(d.pos);
            d.vel = {R(s.vel[0]), R(s.vel[1])};
            d.energy = R(s.energy);
            d.stiffness = R(s.stiffness);
            d.density = R(1);
            d.color = s.color;
            d.maxConn = s.maxConnections;
            d.numConn = s.numConnections;
            d.barrier = s.barrier ? 1 : 0;
            d.livingState = s.livingState;
            d.cellFunction = s.cellFunction;
            d.ancestorMutationId = s.ancestorMutationId;
            d.selected = s.selected;
            d.age = s.age;
            d.creatureId = s.creatureId;
            d.mutationId = s.mutationId;
            d.activationTime = s.activationTime;
            d.genomeComplexity = s.genomeComplexity;
            if (d.numConn > ALIEN_MAX_CELL_BONDS) throw std::invalid_argument("numConnections > 6");
            if (d.color >= ALIEN_MAX_COLORS) throw std::invalid_argument("cell color must be below 7");
            for (int b = 0; b < d.numConn; ++b) {
                int t = s.connections[b].cellIndex;
                if (t < 0 || uint64_t(t) >= n) throw std::invalid_argument("bond index out of range");
                d.bonds[b] = {t, R(s.connections[b].distance), R(s.connections[b].angleFromPrevious)};
            }
        }
        for (uint64_t i = 0; i < n; ++i)
            for (int b = 0; b < cells[i].numConn; ++b) {
                const int t = cells[i].bonds[b].target;
                int back = 0, same = 0;
                for (int k = 0; k < cells[t].numConn; ++k) back += cells[t].bonds[k].target == int(i);
                for (int k = 0; k < cells[i].numConn; ++k) same += cells[i].bonds[k].target == t;
                if (t == int(i) || back != 1 || same != 1) throw std::invalid_argument("bonds must be symmetric");
            }
        parts.assign(np, Particle<R>{});
        for (uint64_t i = 0; i < np; ++i) {
            auto& d = parts[i];
            d.id = q[i].id;
            d.pos = {R(q[i].pos[0]), R(q[i].pos[1])};
            torus.correctPosition(d.pos);
            d.vel = {R(q[i].vel[0]), R(q[i].vel[1])};
            d.energy = R(q[i].energy);
            d.color = q[i].color;
            if (d.color >= ALIEN_MAX_COLORS) throw std::invalid_argument("particle color must be below 7");
            d.selected = q[i].selected;
            d.hasLast = false;
            d.lastAbsorbedCellId = 0;
            d.alive = true;
        }
    }

    // cells and particles in ascending id order, bonds as indices into the returned cells
    void getData(std::vector<alien_cell_to>& outC, std::vector<alien_particle_to>& outP) const {
        std::vector<int> order(cells.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells[a].id < cells[b].id; });
        std::vector<int> pos(cells.size());
        for (size_t k = 0; k < order.size(); ++k) pos[order[k]] = int(k);
        outC.assign(cells.size(), alien_cell_to{});
        for (size_t k = 0; k < order.size(); ++k) {
            const auto& s = cells[order[k]];
            auto& d = outC[k];
            d.id = s.id;
            d.pos[0] = float(s.pos.x);
            d.pos[1] = float(s.pos.y);
            d.vel[0] = float(s.vel.x);
            d.vel[1] = float(s.vel.y);
            d.energy = float(s.energy);
            d.stiffness = float(s.stiffness);
            d.color = s.color;
            d.maxConnections = s.maxConn;
            d.numConnections = s.numConn;
            d.barrier = s.barrier;
            d.age = s.age;
            d.livingState = s.livingState;
            d.cellFunction = s.cellFunction;
            d.ancestorMutationId = s.ancestorMutationId;
            d.selected = s.selected;
            d.creatureId = s.creatureId;
            d.mutationId = s.mutationId;
            d.genomeComplexity = s.genomeComplexity;
            d.activationTime = s.activationTime;
            for (int b = 0; b < s.numConn; ++b) {
                d.connections[b].cellIndex = pos[s.bonds[b].target];
                d.connections[b].distance = float(s.bonds[b].rest);
                d.connections[b].angleFromPrevious = float(s.bonds[b].angle);
            }
        }
        std::vector<int> po(parts.size());
        for (size_t i = 0; i < po.size(); ++i) po[i] = int(i);
        std::stable_sort(po.begin(), po.end(), [&](int a, int b) { return parts[a].id < parts[b].id; });
        outP.assign(parts.size(), alien_particle_to{});
        for (size_t k = 0; k < po.size(); ++k) {
            const auto& s = parts[po[k]];
            auto& d = outP[k];
            d.id = s.id;
            d.energy = float(s.energy);
            d.pos[0] = float(s.pos.x);
            d.pos[1] = float(s.pos.y);
            d.vel[0] = float(s.vel.x);
            d.vel[1] = float(s.vel.y);
            d.color = s.color;
            d.selected = s.selected;
        }
    }

    // ---- R1: spatial sort of cells -------------------------------------------------------------
    void sortCells() {
        const int nb = nbx * nby;
        const size_t n = cells.size();
        std::vector<int> key(n);
        std::vector<int> count(nb + 1, 0);
        for (size_t i = 0; i < n; ++i) {
            key[i] = binOfPos(cells[i].pos);
            ++count[key[i] + 1];
        }
        for (int b = 0; b < nb; ++b) count[b + 1] += count[b];
        cellBinStart = count;
        std::vector<int> cursor(count.begin(), count.end() - 1);
        std::vector<int> newIdx(n);
        for (size_t i = 0; i < n; ++i) newIdx[i] = cursor[key[i]]++;  // stable: ascending old index
        std::vector<Cell<R>> sorted(n);
        for (size_t i = 0; i < n; ++i) {
            Cell<R> c = cells[i];
            for (int b = 0; b < c.numConn; ++b) c.bonds[b].target = newIdx[c.bonds[b].target];
            sorted[newIdx[i]] = c;
        }
        cells.swap(sorted);
        cellBinOfIndex.assign(n, 0);
        startPos.resize(n);
        for (size_t i = 0; i < n; ++i) {
            cellBinOfIndex[newIdx[i]] = key[i];
            startPos[newIdx[i]] = cells[newIdx[i]].pos;
        }
    }

    // ---- R2/R10: spatial sort of live particles (pre-move positions) -----------------------------
    void sortParticles() {
        const int nb = pnbx * pnby;
        std::vector<Particle<R>> live;
        live.reserve(parts.size());
        for (auto& q : parts)
            if (q.alive) live.push_back(q);
        const size_t n = live.size();
        std::vector<int> key(n), count(nb + 1, 0);
        for (size_t i = 0; i < n; ++i) {
            key[i] = pbinOfPos(live[i].pos);
            ++count[key[i] + 1];
        }
        for (int b = 0; b < nb; ++b) count[b + 1] += count[b];
        partBinStart = count;
        std::vector<int> cursor(count.begin(), count.end() - 1);
        parts.assign(n, Particle<R>{});
        for (size_t i = 0; i < n; ++i) parts[cursor[key[i]]++] = live[i];
    }

    template <class F> void forStencil(int bin, F&& f) const {
        const int bx = bin % nbx, by = bin / nbx;
        for (int dy = -1; dy <= 1; ++dy) {
            int y = by + dy;
            y = y < 0 ? y + nby : (y >= nby ? y - nby : y);
            for (int dx = -1; dx <= 1; ++dx) {
                int x = bx + dx;
                x = x < 0 ? x + nbx : (x >= nbx ? x - nbx : x);
                int b = y * nbx + x;
                for (int j = cellBinStart[b]; j < cellBinStart[b + 1]; ++j) f(j);
            }
        }
    }

    bool bonded(int i, int j) const {
        const auto& c = cells[i];
        for (int b = 0; b < c.numConn; ++b)
            if (c.bonds[b].target == j) return true;
        return false;
    }

    // neighbour predicate of the step (Map.cuh:189 collision; CellProcessor.cuh:169-180 fluid)
    bool isNeighbor(int i, int j) const {
        if (i == j) return false;
        const auto& a = cells[i];
        const auto& b = cells[j];
        if (rawParams.motionType == ALIEN_MOTION_COLLISION) {
            return length(b.pos - a.pos) <= p.cellMaxCollisionDistance;
        }
        V2<R> d = torus.delta(a.pos, b.pos);
        return length(d) <= p.smoothingLength * R(2);
    }

    // ---- neighbour pass outputs ------------------------------------------------------------------
    struct PassOut {
        std::vector<V2<R>> force, posCorr;
        std::vector<R> rho;
        std::vector<std::vector<std::pair<int, int>>> fusion;  // per cell i: (i, j) with i < j
    };

    // The force the visit of cell `self` to a neighbour adds to `self` (its negation goes to the
    // neighbour): CellProcessor.cuh:319-339.  posDelta = self - other, velDelta = v_self - v_other.
    V2<R> collisionVisitForce(const Cell<R>& self, V2<R> posDelta, V2<R> velDelta) const {
        const R distance = length(posDelta);
        const bool approaching = dot(posDelta, velDelta) < R(0);
        const R barrierFactor = self.barrier ? R(2) : R(1);
        if (length(self.vel) > R(0.5f) && approaching) {
            R dsq = distance * distance + R(0.25f);
            return posDelta * dot(velDelta, posDelta) * (R(1) / (R(-2) * dsq)) * barrierFactor;  // R16
        }
        return normalized(posDelta) * (p.cellMaxCollisionDistance - length(posDelta)) * p.cellRepulsionStrength *
               barrierFactor;
    }

    // collision mode: CellProcessor.cuh:285-365 (R3, R4, R13)
    void collisionCell(int i, PassOut& out) const {
        const auto& c = cells[i];
        const R cmd = p.cellMinDistance;
        V2<R> F{R(0), R(0)};
        V2<R> pc = c.pos;
        forStencil(cellBinOfIndex[i], [&](int j) {
            if (!isNeighbor(i, j)) return;
            const auto& o = cells[j];
            V2<R> posDelta = torus.delta(c.pos, o.pos);
            R distance = length(posDelta);
            if (!c.barrier && distance < cmd) pc = pc + posDelta * cmd / R(5);
            if (bonded(i, j)) return;
            V2<R> velDelta = c.vel - o.vel;
            bool approaching = dot(posDelta, velDelta) < R(0);
            // f(i,j): force i's visit adds to i; f(j,i): force j's visit adds to j, i receives its
            // negation
            const V2<R> fij = collisionVisitForce(c, posDelta, velDelta);
            const V2<R> fji = collisionVisitForce(o, V2<R>{-posDelta.x, -posDelta.y}, V2<R>{-velDelta.x, -velDelta.y});
            F = F + fij;
            F = F + V2<R>{-fji.x, -fji.y};
            if (i < j && c.numConn < c.maxConn && o.numConn < o.maxConn && length(velDelta) >= p.cellFusionVelocity &&
                approaching && c.energy <= p.cellMaxBindingEnergy && o.energy <= p.cellMaxBindingEnergy && !c.barrier &&
                !o.barrier)
                out.fusion[i].push_back({i, j});
        });
        out.force[i] = F;
        out.posCorr[i] = pc;
    }

    // fluid mode: CellProcessor.cuh:122-283 (R4, R12)
    void fluidCell(int i, PassOut& out) const {
        const auto& c = cells[i];
        const R h = p.smoothingLength;
        const R cmd = p.cellMinDistance;
        // R16: divisions by h, h^2, h^3 and by the densities as products with reciprocals
        const R invH = R(1) / h, invH2 = R(1) / (h * h), invH3 = R(1) / (h * h * h);
        const R invDi = R(1) / c.density;
        V2<R> Fp{R(0), R(0)}, Fv{R(0), R(0)}, D{R(0), R(0)};
        R rho = R(0);
        int barrierList[10];
        int numBarrier = 0;
        forStencil(cellBinOfIndex[i], [&](int j) {
            const auto& o = cells[j];
            V2<R> posDelta = torus.delta(c.pos, o.pos);
            R distance = length(posDelta);
            if (!(distance <= h * R(2))) return;
            if (o.barrier) {
                if (numBarrier < 10) barrierList[numBarrier] = j;
                ++numBarrier;
                return;
            }
            rho = rho + calcKernel(distance * invH) * invH2;
            if (j == i) return;
            if (!c.barrier && distance < cmd) D = D + posDelta * cmd / R(5);
            if (bonded(i, j)) return;
            V2<R> velDelta = c.vel - o.vel;
            const R invDj = R(1) / o.density;
            R factor = invDi + invDj;
            if (std::abs(distance) > K<R>::NEAR_ZERO) {
                R kd = calcKernel_d(distance * invH) * invH3;
                Fp = Fp + posDelta * (R(1) / (-distance)) * factor * kd;
                Fv = Fv + velDelta * invDj * distance * kd * (R(1) / (distance * distance + R(0.25f)));
            }
            if (i < j && length(velDelta) >= p.cellFusionVelocity && c.numConn < c.maxConn && o.numConn < o.maxConn &&
                c.energy <= p.cellMaxBindingEnergy && o.energy <= p.cellMaxBindingEnergy && !c.barrier && !o.barrier)
                out.fusion[i].push_back({i, j});
        });
        V2<R> F{R(0), R(0)};
        numBarrier = std::min(numBarrier, 10);
        if (numBarrier > 0) {  // CellProcessor.cuh:234-274
            int cb = -1;
            R cbDist = R(0);
            for (int k = 0; k < numBarrier; ++k) {
                R d = length(torus.delta(c.pos, cells[barrierList[k]].pos));
                if (cb < 0 || d < cbDist) {
                    cb = barrierList[k];
                    cbDist = d;
                }
            }
            const auto& b = cells[cb];
            V2<R> rr{R(0), R(0)};
            if (b.numConn <= 1) {
                rr = torus.delta(c.pos, b.pos);
            } else {
                R angleToCell = angleOfVector(torus.delta(c.pos, b.pos));
                int nc = b.numConn;
                for (int k = 0; k < nc; ++k) {
                    const auto& o1 = cells[b.bonds[k].target];
                    const auto& o2 = cells[b.bonds[(k + 1) % nc].target];
                    R a1 = angleOfVector(torus.delta(o1.pos, b.pos));
                    R a2 = angleOfVector(torus.delta(o2.pos, b.pos));
                    if (isAngleInBetween(a1, a2, angleToCell)) {
                        rr = o2.pos - o1.pos;
                        rotateQuarterCounterClockwise(rr);
                        break;
                    }
                }
            }
            V2<R> vr = c.vel - b.vel;
            R d = dot(vr, rr);
            if (d < R(0)) {
                R trs = std::max(R(0.05f), lengthSquared(rr));
                R td = std::max(R(0.05f), cbDist);
                F = F + (vr - rr * R(2) * d / trs + b.vel - c.vel) / td;
            }
        }
        F = F + (Fp * p.pressureStrength + Fv * p.viscosityStrength);
        out.force[i] = F;
        out.posCorr[i] = c.pos + D;
        out.rho[i] = rho;
    }

    // ---- particles: RadiationProcessor.cuh:61-170 (R2, R8, R10) --------------------------------
    int firstCellInSlot(int sx, int sy) const {
        int b = binOfSlot(sx, sy);
        for (int j = cellBinStart[b]; j < cellBinStart[b + 1]; ++j) {
            int cx, cy;
            torus.slotOf(startPos[j], cx, cy);
            if (cx == sx && cy == sy) return j;
        }
        return -1;
    }

    void particlePhase() {
        const size_t np = parts.size();
        std::vector<V2<R>> prePos(np);
        for (size_t k = 0; k < np; ++k) {
            prePos[k] = parts[k].pos;
            parts[k].pos = parts[k].pos + parts[k].vel * p.timestepSize;  // movement
            torus.correctPosition(parts[k].pos);
        }
        auto winner = [&](int sx, int sy) -> int {
            int b = pbinOfSlot(sx, sy);
            for (int j = partBinStart[b]; j < partBinStart[b + 1]; ++j) {
                int px, py;
                torus.slotOf(prePos[j], px, py);
                if (px == sx && py == sy) return j;
            }
            return -1;
        };
        std::vector<int> target(np, -1);
        for (size_t k = 0; k < np; ++k) {
            int sx, sy;
            torus.slotOf(parts[k].pos, sx, sy);
            int t = winner(sx, sy);
            if (t >= 0 && t != int(k) && lengthSquared(parts[k].pos - parts[t].pos) < R(0.5f)) target[k] = int(t);
        }
        std::vector<int> claimant(np, -1);
        for (size_t k = 0; k < np; ++k) {
            int t = target[k];
            if (t >= 0 && target[t] < 0 && claimant[t] < 0) claimant[t] = int(k);  // lowest k first
        }
        // fusion (RadiationProcessor.cuh:86-93)
        for (size_t t = 0; t < np; ++t) {
            int q = claimant[t];
            if (q < 0) continue;
            auto& P = parts[q];
            auto& T = parts[t];
            if (P.energy > K<R>::NEAR_ZERO && T.energy > K<R>::NEAR_ZERO) {
                R f1 = P.energy / (P.energy + T.energy);
                T.vel = P.vel * f1 + T.vel * (R(1) - f1);
                T.energy = T.energy + P.energy;
                T.hasLast = false;
                P.energy = R(0);
                P.alive = false;
            }
        }
        // absorption / barrier reflection (RadiationProcessor.cuh:97-167)
        std::vector<int> absorbCell(np, -1);
        std::vector<int> cellClaim(cells.size(), -1);
        for (size_t k = 0; k < np; ++k) {
            if (target[k] >= 0 || claimant[k] >= 0) continue;
            auto& P = parts[k];
            int sx, sy;
            torus.slotOf(P.pos + P.vel, sx, sy);
            int c = firstCellInSlot(sx, sy);
            if (c < 0) continue;
            const auto& C = cells[c];
            if (C.barrier) {
                V2<R> vr = P.vel - C.vel;
                V2<R> rr = torus.delta(P.pos, C.pos);
                R d = dot(vr, rr);
                if (d < R(0)) {
                    R trs = std::max(R(0.1f), lengthSquared(rr));
                    P.vel = vr - rr * R(2) * d / trs + C.vel;
                }
                continue;
            }
            if (P.hasLast && P.lastAbsorbedCellId == C.id) continue;
            if (p.radiationAbsorption[C.color % ALIEN_MAX_COLORS] < K<R>::NEAR_ZERO) continue;
            absorbCell[k] = c;
            if (cellClaim[c] < 0) cellClaim[c] = int(k);
        }
        for (size_t k = 0; k < np; ++k) {
            int c = absorbCell[k];
            if (c < 0 || cellClaim[c] != int(k)) continue;
            auto& P = parts[k];
            auto& C = cells[c];
            R ra = p.radiationAbsorption[C.color % ALIEN_MAX_COLORS];
            R transfer = P.energy * ra;
            if (P.energy < R(0.01f)) transfer = P.energy;
            C.energy = C.energy + transfer;
            P.energy = P.energy - transfer;
            if (P.energy < K<R>::NEAR_ZERO) {
                P.alive = false;
            } else {
                P.hasLast = true;
                P.lastAbsorbedCellId = C.id;
            }
        }
        // splitting (RadiationProcessor.cuh:172-204): with probability 0.01 a particle above the
        // split energy of its colour halves its energy and splits, the halves 1/5 of a random unit
        // vector u either side, their velocities -/+ u perturbation / (E/2 + 1)
        for (size_t k = 0; k < np; ++k) {
            if (!parts[k].alive) continue;
            const uint64_t id = parts[k].id;
            if (!(R(uniformDraw(seed, timestep, id, kStreamSplit)) < R(0.01f))) continue;
            if (!(parts[k].energy > p.particleSplitEnergy[parts[k].color % ALIEN_MAX_COLORS])) continue;
            Particle<R> q = parts[k];
            q.energy = q.energy * R(0.5f);
            V2<R> u = unitVectorOfAngle(R(uniformDraw(seed, timestep, id, kStreamSplitAngle)) * R(360));
            const V2<R> d = u / R(5);
            V2<R> other = q.pos + d;
            torus.correctPosition(other);
            q.pos = q.pos - d;
            torus.correctPosition(q.pos);
            u = u * (p.radiationVelocityPerturbation / (q.energy + R(1)));
            Particle<R> half{};
            half.id = kCreatedParticleId | kSplitParticleId | ((timestep & 0x3fffffffull) << 32) | uint64_t(uint32_t(k));
            half.pos = other;
            half.vel = q.vel + u;
            half.energy = q.energy;
            half.color = q.color;
            half.selected = 0;
            half.lastAbsorbedCellId = 0;
            half.hasLast = false;
            half.alive = true;
            q.vel = q.vel - u;
            parts[k] = q;
            parts.push_back(half);
        }
    }

    // ---- connection forces: CellProcessor.cuh:408-491 (R7) -------------------------------------
    struct ConnOut {
        std::vector<V2<R>> own;
        std::vector<V2<R>> out1, out2;  // [cell*6 + bond]
        std::vector<uint16_t> valid;    // bit b: out1[b], bit 6+b: out2[b]
    };

    // bond vector from cell x to cell o; with `shift`, both cells stand at pos + shift and the
    // vector is the unshifted difference plus the difference of the shifts (R15)
    V2<R> bondVector(int o, int x, const std::vector<V2<R>>* shift) const {
        V2<R> d = torus.delta(cells[o].pos, cells[x].pos);
        if (shift) d = d + ((*shift)[o] - (*shift)[x]);
        return d;
    }

    void connectionForces(bool considerAngles, std::vector<V2<R>>& s1, const std::vector<V2<R>>* shift) const {
        const int n = int(cells.size());
        ConnOut co;
        co.own.assign(n, V2<R>{R(0), R(0)});
        co.out1.assign(size_t(n) * 6, V2<R>{R(0), R(0)});
        co.out2.assign(size_t(n) * 6, V2<R>{R(0), R(0)});
        co.valid.assign(n, 0);
#pragma omp parallel for schedule(static) num_threads(numThreads)
        for (int x = 0; x < n; ++x) {
            const auto& c = cells[x];
            if (c.numConn == 0 || c.barrier) continue;
            V2<R> force{R(0), R(0)};
            const int nc = c.numConn;
            V2<R> prevDisp = bondVector(c.bonds[nc - 1].target, x, shift);
            R kx2 = c.stiffness * c.stiffness;
            uint16_t v = 0;
            for (int b = 0; b < nc; ++b) {
                const auto& o = cells[c.bonds[b].target];
                R kc2 = o.stiffness * o.stiffness;
                V2<R> disp = bondVector(c.bonds[b].target, x, shift);
                R actualDistance = length(disp);
                R deviation = actualDistance - c.bonds[b].rest;
                force = force + normalized(disp) * deviation * (kx2 + kc2) / R(6);
                if (considerAngles && (nc > 2 || (nc == 2 && b == 0))) {
                    int last = (b + nc - 1) % nc;
                    int lc = c.bonds[last].target;
                    bool tri = false;
                    for (int k = 0; k < o.numConn; ++k)
                        if (o.bonds[k].target == lc) {
                            tri = true;
                            break;
                        }
                    if (!tri) {
                        R angle = angleOfVector(disp);
                        R prevAngle = angleOfVector(prevDisp);
                        R act = subtractAngle(angle, prevAngle);
                        if (act < R(0)) continue;
                        R ref = c.bonds[b].angle;
                        R strength = std::abs(ref - act) / R(2000) * kx2;
                        V2<R> f1 = normalized(disp) / std::max(length(disp), p.cellMinDistance) * strength;
                        rotateQuarterClockwise(f1);
                        V2<R> f2 = normalized(prevDisp) / std::max(length(prevDisp), p.cellMinDistance) * strength;
                        rotateQuarterCounterClockwise(f2);
                        if (ref < act) {
                            f1 = f1 * R(-1);
                            f2 = f2 * R(-1);
                        }
                        if (!o.barrier) {
                            co.out1[size_t(x) * 6 + b] = f1;
                            v |= uint16_t(1u << b);
                        }
                        if (!cells[lc].barrier) {
                            co.out2[size_t(x) * 6 + b] = f2;
                            v |= uint16_t(1u << (6 + b));
                        }
                        force = force - (f1 + f2);
                    }
                }
                prevDisp = disp;
            }
            co.own[x] = force;
            co.valid[x] = v;
        }
        s1.assign(n, V2<R>{R(0), R(0)});
#pragma omp parallel for schedule(static) num_threads(numThreads)
        for (int x = 0; x < n; ++x) {
            const auto& c = cells[x];
            if (c.numConn == 0 || c.barrier) continue;
            V2<R> acc{R(0), R(0)};
            acc = acc + co.own[x];
            for (int q = 0; q < c.numConn; ++q) {
                int pIdx = c.bonds[q].target;
                const auto& P = cells[pIdx];
                int b = -1;
                for (int k = 0; k < P.numConn; ++k)
                    if (P.bonds[k].target == x) {
                        b = k;
                        break;
                    }
                if (b < 0) continue;
                if (co.valid[pIdx] & (1u << b)) acc = acc + co.out1[size_t(pIdx) * 6 + b];
                int b2 = (b + 1) % P.numConn;
                if (co.valid[pIdx] & (1u << (6 + b2))) acc = acc + co.out2[size_t(pIdx) * 6 + b2];
            }
            s1[x] = acc;
        }
    }

    // ---- structural operations: CellConnectionProcessor.cuh (R5, R6) ---------------------------
    bool wouldResultInOverlappingConnection(int c1, V2<R> otherPos) const {  // :476-500
        const auto& c = cells[c1];
        const int n = c.numConn;
        if (n < 2) return false;
        for (int i = 0; i < n; ++i) {
            int cc = c.bonds[i].target;
            int nc = c.bonds[(i + 1) % n].target;
            bool both = false;
            for (int j = 0; j < cells[cc].numConn; ++j)
                if (cells[cc].bonds[j].target == nc) {
                    both = true;
                    break;
                }
            if (!both) continue;
            if (crossing(c.pos, otherPos, cells[cc].pos, cells[nc].pos)) return true;
        }
        return false;
    }

    bool tryAddConnectionOneWay(int c1, int c2, V2<R> posDelta) {  // :274-427 (no alignment)
        auto& c = cells[c1];
        if (c.numConn == ALIEN_MAX_CELL_BONDS) return false;
        if (wouldResultInOverlappingConnection(c1, cells[c2].pos)) return false;
        R newAngle = angleOfVector(posDelta);
        R desiredDistance = length(posDelta);
        if (c.numConn == 0) {
            c.bonds[0] = {c2, desiredDistance, R(360)};
            c.numConn = 1;
            return true;
        }
        if (c.numConn == 1) {
            V2<R> cd = torus.delta(cells[c.bonds[0].target].pos, c.pos);
            R prevAngle = angleOfVector(cd);
            R angleDiff = newAngle - prevAngle;
            if (angleDiff < R(0)) angleDiff += R(360);
            if (std::abs(angleDiff) < K<R>::NEAR_ZERO || std::abs(angleDiff - R(360)) < K<R>::NEAR_ZERO ||
                std::abs(angleDiff + R(360)) < K<R>::NEAR_ZERO)
                return false;
            c.bonds[1].angle = angleDiff;
            c.bonds[0].angle = R(360) - angleDiff;
            c.bonds[1].target = c2;
            c.bonds[1].rest = desiredDistance;
            c.numConn = 2;
            return true;
        }
        const int n = c.numConn;
        int index = 0;
        R prevAngle = R(0), nextAngle = R(0);
        bool found = false;
        for (; index < n; ++index) {
            int prevIndex = (index + n - 1) % n;
            prevAngle = angleOfVector(torus.delta(cells[c.bonds[prevIndex].target].pos, c.pos));
            nextAngle = angleOfVector(torus.delta(cells[c.bonds[index].target].pos, c.pos));
            if (isAngleInBetween(prevAngle, nextAngle, newAngle) || prevIndex == index) {
                found = true;
                break;
            }
        }
        if (!found) return false;  // no angular gap found: treated as a failed insertion
        R refAngle = c.bonds[index].angle;
        R angleDiff1 = subtractAngle(newAngle, prevAngle);
        R angleDiff2 = subtractAngle(nextAngle, prevAngle);
        R frac = angleDiff2 != R(0) ? angleDiff1 / angleDiff2 : R(0.5f);
        R afp = refAngle * frac;
        afp = std::min(afp, refAngle);
        if (afp < K<R>::NEAR_ZERO) return false;
        if (index == 0) index = n;
        for (int j = n; j > index; --j) c.bonds[j] = c.bonds[j - 1];
        c.bonds[index] = {c2, desiredDistance, afp};
        c.bonds[(index + 1) % (n + 1)].angle = refAngle - afp;
        c.numConn = uint8_t(n + 1);
        return true;
    }

    bool tryAddConnections(int a, int b) {  // :211-240
        V2<R> posDelta = torus.delta(cells[b].pos, cells[a].pos);
        Cell<R> saved = cells[a];
        if (!tryAddConnectionOneWay(a, b, posDelta)) return false;
        if (!tryAddConnectionOneWay(b, a, posDelta * R(-1))) {
            cells[a].numConn = saved.numConn;
            for (int i = 0; i < ALIEN_MAX_CELL_BONDS; ++i) cells[a].bonds[i] = saved.bonds[i];
            return false;
        }
        return true;
    }

    void deleteConnectionOneWay(int c1, int c2) {  // :429-448
        auto& c = cells[c1];
        for (int i = 0; i < c.numConn; ++i) {
            if (c.bonds[i].target == c2) {
                R angleToAdd = c.bonds[i].angle;
                for (int j = i; j < c.numConn - 1; ++j) c.bonds[j] = c.bonds[j + 1];
                if (i < c.numConn - 1)
                    c.bonds[i].angle += angleToAdd;
                else
                    c.bonds[0].angle += angleToAdd;
                --c.numConn;
                return;
            }
        }
    }

    void structuralOps(const std::vector<std::pair<int, int>>& cands, const std::vector<uint8_t>& delFlag) {
        const int n = int(cells.size());
        std::vector<uint8_t> mark(n, 0);
        for (int x = 0; x < n; ++x) {
            if (delFlag[x]) {
                mark[x] = 1;
                continue;
            }
            for (int b = 0; b < cells[x].numConn; ++b)
                if (delFlag[cells[x].bonds[b].target]) mark[x] = 1;
        }
        // R5: claims over {a, b} and their bond partners
        const uint64_t NONE = std::numeric_limits<uint64_t>::max();
        std::vector<uint64_t> claim(n, NONE);
        std::vector<uint8_t> live(cands.size(), 0);
        auto forClaimSet = [&](int a, int b, auto&& f) {
            f(a);
            for (int k = 0; k < cells[a].numConn; ++k) f(cells[a].bonds[k].target);
            f(b);
            for (int k = 0; k < cells[b].numConn; ++k) f(cells[b].bonds[k].target);
        };
        for (size_t q = 0; q < cands.size(); ++q) {
            int a = cands[q].first, b = cands[q].second;
            if (mark[a] || mark[b]) continue;
            live[q] = 1;
            uint64_t key = (uint64_t(uint32_t(a)) << 32) | uint32_t(b);
            forClaimSet(a, b, [&](int c) { claim[c] = std::min(claim[c], key); });
        }
        std::vector<uint8_t> win(cands.size(), 0);
        for (size_t q = 0; q < cands.size(); ++q) {
            if (!live[q]) continue;
            int a = cands[q].first, b = cands[q].second;
            uint64_t key = (uint64_t(uint32_t(a)) << 32) | uint32_t(b);
            bool ok = true;
            forClaimSet(a, b, [&](int c) { ok = ok && claim[c] == key; });
            win[q] = ok;
        }
        // winners touch disjoint neighbourhoods, so applying them in list order equals any order
        for (size_t q = 0; q < cands.size(); ++q) {
            if (!win[q]) continue;
            int a = cands[q].first, b = cands[q].second;
            // lockAndtryAddConnections (:249-272)
            if (bonded(a, b)) continue;
            if (!(cells[a].numConn < cells[a].maxConn && cells[b].numConn < cells[b].maxConn)) continue;
            tryAddConnections(a, b);
        }
        // R6: deletions, bond-list order of the schedule-time list
        for (int x = 0; x < n; ++x) {
            if (!mark[x]) continue;
            int snap[ALIEN_MAX_CELL_BONDS];
            int ns = cells[x].numConn;
            for (int b = 0; b < ns; ++b) snap[b] = cells[x].bonds[b].target;
            for (int b = 0; b < ns; ++b)
                if (delFlag[x] || delFlag[snap[b]]) deleteConnectionOneWay(x, snap[b]);
        }
    }

    // ---- cell radiation: CellProcessor.cuh:725-773, RadiationProcessor.cuh:248-364 (R9, R18) ----
    // A non-barrier cell radiates when its draw is below radiationProb: its factor is
    // highRadiationFactor above highRadiationMinCellEnergy plus radiationCellAgeStrength above
    // radiationMinCellAge; it loses 2 u E factor / radiationProb (capped at E - 1, and only if
    // E > 1), carried off by a particle with the cell's velocity times the multiplier plus a unit
    // vector of random angle times the perturbation, placed 1.5 ahead of the cell along that
    // velocity and one velocity back; a particle below NEAR_ZERO energy is not created (its energy
    // is lost).  No radiation sources, no external energy.
    void radiation() {
        if (!(p.radiationProb > R(0))) return;
        for (size_t i = 0; i < cells.size(); ++i) {
            auto& c = cells[i];
            if (c.barrier) continue;
            if (!(R(uniformDraw(seed, timestep, c.id, kStreamRadiation)) < p.radiationProb)) continue;
            const int col = c.color % ALIEN_MAX_COLORS;
            R factor = R(0);
            if (c.energy > p.highRadiationMinCellEnergy[col]) factor = factor + p.highRadiationFactor[col];
            if (int64_t(c.age) > int64_t(p.radiationMinCellAge[col])) factor = factor + p.radiationCellAgeStrength[col];
            if (!(factor > R(0))) continue;
            R loss = c.energy * factor;
            loss = loss / p.radiationProb;
            loss = R(2) * loss * R(uniformDraw(seed, timestep, c.id, kStreamRadiationEnergy));
            if (!(c.energy > R(1))) continue;
            const V2<R> u = unitVectorOfAngle(R(uniformDraw(seed, timestep, c.id, kStreamRadiationAngle)) * R(360));
            const V2<R> vel = c.vel * p.radiationVelocityMultiplier + u * p.radiationVelocityPerturbation;
            V2<R> pos = c.pos + normalized(vel) * R(1.5f) - vel;
            torus.correctPosition(pos);
            if (loss > c.energy - R(1)) loss = c.energy - R(1);
            if (loss > K<R>::NEAR_ZERO) {
                Particle<R> q{};
                q.id = kCreatedParticleId | ((timestep & 0x3fffffffull) << 32) | uint64_t(uint32_t(i));
                q.pos = pos;
                q.vel = vel;
                q.energy = loss;
                q.color = uint8_t(col);
                q.selected = 0;
                q.lastAbsorbedCellId = 0;
                q.hasLast = false;
                q.alive = true;
                parts.push_back(q);
            }
            c.energy = c.energy - loss;
        }
    }

    // ---- inner friction: CellProcessor.cuh:678-705 (R11) ----------------------------------------
    void innerFriction() {
        const R f = p.innerFriction;
        const int n = int(cells.size());
        std::vector<V2<R>> v0(n);
        for (int i = 0; i < n; ++i) v0[i] = cells[i].vel;
        for (int i = 0; i < n; ++i) {
            auto& c = cells[i];
            if (c.barrier || c.numConn == 0) continue;
            V2<R> v = v0[i];
            for (int visit = 0; visit < 2; ++visit)  // the cell's visit of the bond, then the partner's
                for (int b = 0; b < c.numConn; ++b) {
                    const int j = c.bonds[b].target;
                    if (cells[j].barrier) continue;
                    const V2<R> avg = (v + v0[j]) / R(2);
                    v = v * (R(1) - f) + avg * f;
                }
            c.vel = v;
        }
    }

    // ---- living states (R17), cellDeathConsequences = None ------------------------------------
    static constexpr uint8_t kReady = 0, kUnderConstruction = 1, kActivating = 2, kDetaching = 3, kReviving = 4,
                             kDying = 5;  // CellFunctionConstants.h:24-33

    // checkConnections' Detaching marks (CellProcessor.cuh:515-520), aging (:561-594) and the state
    // transitions (:596-676)
    void agingAndTransitions(const std::vector<uint8_t>& stretched) {
        const int n = int(cells.size());
        for (int i = 0; i < n; ++i)
            if (stretched[i])
                for (int b = 0; b < cells[i].numConn; ++b) cells[cells[i].bonds[b].target].livingState = kDetaching;
        for (auto& c : cells) {
            if (c.barrier) continue;
            ++c.age;
            if (c.livingState == kReady && c.activationTime > 0) --c.activationTime;
        }
        std::vector<uint8_t> st(n);
        for (int i = 0; i < n; ++i) st[i] = cells[i].livingState;
        for (int i = 0; i < n; ++i) {
            auto& c = cells[i];
            uint8_t next = st[i];
            if (c.barrier) {
                next = kReady;
            } else if (st[i] == kActivating || st[i] == kReviving || st[i] == kDetaching) {
                next = kReady;  // Detaching -> Ready under None
            } else if (st[i] == kUnderConstruction) {
                for (int b = 0; b < c.numConn; ++b) {
                    const int j = c.bonds[b].target;
                    if (cells[j].creatureId == c.creatureId && st[j] == kActivating) next = kActivating;
                }
            }
            c.livingState = next;
        }
    }

    // decay (CellProcessor.cuh:775-858): energy below cellMinEnergy -> Dying, the partners of a
    // newly dying cell -> Detaching (cellDeathProbability 0: no cell is deleted)
    void decayToDying() {
        const int n = int(cells.size());
        std::vector<uint8_t> low(n, 0), fresh(n, 0);
        for (int i = 0; i < n; ++i) {
            const auto& c = cells[i];
            if (c.barrier) continue;
            low[i] = c.energy < p.cellMinEnergy[c.color] ? 1 : 0;
            fresh[i] = low[i] && c.livingState != kDying;
        }
        std::vector<uint8_t> before(n);
        for (int i = 0; i < n; ++i) before[i] = cells[i].livingState;
        for (int i = 0; i < n; ++i) {
            if (low[i]) {
                cells[i].livingState = kDying;
                continue;
            }
            if (before[i] == kDying) continue;
            for (int b = 0; b < cells[i].numConn; ++b)
                if (fresh[cells[i].bonds[b].target]) cells[i].livingState = kDetaching;
        }
    }

    // ---- one timestep: SimulationKernelsLauncher.cu:25-91 ----------------------------------------
    void step() {
        computeGeometry();
        const bool fluid = rawParams.motionType == ALIEN_MOTION_FLUID;
        if (densityResetPending) {  // resetDensity, CellProcessor.cuh:860-870
            for (auto& c : cells) c.density = R(1);
            densityResetPending = false;
        }
        const bool considerAngles = (timestep % 3 == 0);  // SimulationKernelsLauncher.cu:31
        radiation();  // (R18: in the cells' order of the last map, before the maps)
        sortCells();
        sortParticles();
        const int n = int(cells.size());

        // neighbour pass (calcFluidForces / calcCollisions)
        PassOut po;
        po.force.assign(n, V2<R>{R(0), R(0)});
        po.posCorr.resize(n);
        po.rho.assign(n, R(0));
        po.fusion.assign(n, {});
#pragma omp parallel for schedule(dynamic, 256) num_threads(numThreads)
        for (int i = 0; i < n; ++i) {
            if (fluid)
                fluidCell(i, po);
            else
                collisionCell(i, po);
        }
        // checkForces + applyForces (CellProcessor.cuh:367-406)
        std::vector<uint8_t> delFlag(n, 0);
        for (int i = 0; i < n; ++i) {
            auto& c = cells[i];
            c.pos = po.posCorr[i];
            if (fluid) c.density = po.rho[i];
            if (c.barrier) continue;
            // a bond break by force (deleting all bonds of an unbonded cell does nothing)
            if (p.cellMaxForceDecayProb > R(0) && c.numConn > 0 && length(po.force[i]) > p.cellMaxForce[c.color] &&
                R(uniformDraw(seed, timestep, c.id, kStreamForceDecay)) < p.cellMaxForceDecayProb)
                delFlag[i] = 1;
            c.vel = c.vel + po.force[i];
            if (length(c.vel) > p.cellMaxVelocity) c.vel = normalized(c.vel) * p.cellMaxVelocity;
        }
        // energy particles: movement + collision
        particlePhase();

        // connection forces, Verlet (CellProcessor.cuh:408-559)
        std::vector<V2<R>> s1, s2(n, V2<R>{R(0), R(0)}), shift(n, V2<R>{R(0), R(0)});
        connectionForces(considerAngles, s1, nullptr);
        const R dt = p.timestepSize;
        for (int i = 0; i < n; ++i) {  // Verlet position step; the displacement is applied below (R15)
            auto& c = cells[i];
            if (c.barrier) {
                shift[i] = c.vel * dt;
            } else {
                shift[i] = c.vel * dt + s1[i] * dt * dt / R(2);
                s2[i] = s1[i];
            }
        }
        std::vector<uint8_t> stretched(n, 0);
        for (int i = 0; i < n; ++i) {  // checkConnections, CellProcessor.cuh:493-523
            const auto& c = cells[i];
            if (c.barrier) continue;
            for (int b = 0; b < c.numConn; ++b) {
                V2<R> disp = bondVector(c.bonds[b].target, i, &shift);
                if (length(disp) > p.cellMaxBindingDistance[c.color % ALIEN_MAX_COLORS]) stretched[i] = 1;
            }
            if (stretched[i]) delFlag[i] = 1;
        }
        connectionForces(considerAngles, s1, &shift);
        for (int i = 0; i < n; ++i) {
            auto& c = cells[i];
            c.pos = c.pos + shift[i];
            torus.correctPosition(c.pos);
        }
        for (int i = 0; i < n; ++i) {  // verletVelocityUpdate, CellProcessor.cuh:546-559
            auto& c = cells[i];
            if (c.barrier) continue;
            V2<R> acc = (s1[i] + s2[i]) / R(2);
            c.vel = c.vel + acc * dt;
        }
        agingAndTransitions(stretched);  // SimulationKernels.cu:81-88
        // applyInnerFriction every third step (SimulationKernelsLauncher.cu:32, 69-71)
        if (timestep % 3 == 0 && p.innerFriction != R(0)) innerFriction();
        for (int i = 0; i < n; ++i) {  // applyFriction + decay (SimulationKernels.cu:153-158)
            auto& c = cells[i];
            if (c.barrier) continue;
            c.vel = c.vel * (R(1) - p.friction);
            if (c.energy > p.cellMaxBindingEnergy) delFlag[i] = 1;
        }
        decayToDying();
        std::vector<std::pair<int, int>> cands;
        for (int i = 0; i < n; ++i)
            for (auto& pr : po.fusion[i]) cands.push_back(pr);
        structuralOps(cands, delFlag);
        // garbage collection of particles (R10)
        std::vector<Particle<R>> live;
        live.reserve(parts.size());
        for (auto& q : parts)
            if (q.alive) live.push_back(q);
        parts.swap(live);
        ++timestep;
    }

    // neighbour lists of the current state in ascending-id order (debug / parity)
    void neighborLists(std::vector<uint64_t>& offsets, std::vector<uint64_t>& ids) {
        computeGeometry();
        sortCells();
        const int n = int(cells.size());
        std::vector<int> order(n);
        for (int i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cells[a].id < cells[b].id; });
        offsets.assign(n + 1, 0);
        ids.clear();
        for (int k = 0; k < n; ++k) {
            int i = order[k];
            forStencil(cellBinOfIndex[i], [&](int j) {
                if (isNeighbor(i, j)) ids.push_back(cells[j].id);
            });
            offsets[k + 1] = ids.size();
        }
    }

    alien_totals totals() const {
        alien_totals t{};
        for (auto& c : cells) {
            t.cellEnergy += double(c.energy);
            t.momentum[0] += double(c.vel.x);
            t.momentum[1] += double(c.vel.y);
            t.kinetic += 0.5 * (double(c.vel.x) * double(c.vel.x) + double(c.vel.y) * double(c.vel.y));
        }
        for (auto& q : parts) t.particleEnergy += double(q.energy);
        t.numCells = cells.size();
        t.numParticles = parts.size();
        return t;
    }
};

}  // namespace oracle
