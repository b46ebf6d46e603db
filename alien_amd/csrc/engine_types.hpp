// engine_types.hpp — device-visible layout of a world set.
//
// A world set is any number of independent worlds stored back to back in one segmented SoA.  A
// plain simulation is a world set of one world; the batch and the server use the same kernels.
// Every index stored in a world's arrays (bond targets, cell slots, last absorbed cell, particle
// slots) is local to that world, so a world's ranges can be copied between sets unchanged.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace alien {

constexpr int kMaxBonds = 6;
constexpr int kColors = 7;
constexpr int kParticleBin = 8;  // edge of the energy-particle bins, in world units
constexpr int kNone = -1;
// neighbour pass: one wave per strip of bin columns, three staged bin rows (strip plus two halo
// columns each) in LDS; at most kRowCap entries per staged row (two per lane while in flight);
// widest strip
constexpr int kRowThreads = 64;
constexpr int kRowCap = 128;
constexpr int kRowPool = 256;  // entries of three consecutive staged rows
constexpr int kStripMax = 64;
constexpr int kChunkMax = 8;  // bin rows per neighbour-pass chunk
constexpr int kClaimFree = 0x7fffffff;
constexpr uint32_t kMoverBit = 0x80000000u;  // binOf mark of a cell that changed bins
constexpr int kXcdCount = 8;  // XCDs of an MI355X: workgroup b of a launch runs on XCD b % 8
// Size limits of one world set.  Cell indices, bin offsets and the bond planes ([kMaxBonds][cap],
// addressed as b * cap + x) are 32-bit ints in the kernels, so a set holds at most
// (2^31 - 1) / kMaxBonds cells; particle indices and the uint32 bin scan over cells plus particles
// stay in range with at most 2^30 particles.  At ~470 B per bonded cell the cell limit is ~168 GB
// of the 288 GB of HBM.
constexpr int64_t kMaxSetCells = (int64_t(0x7fffffff)) / kMaxBonds;  // 357,913,941
constexpr int64_t kMaxSetParticles = int64_t(1) << 30;

// cell flags word
constexpr uint32_t kNumConnMask = 0xFu;        // bits 0..3
constexpr uint32_t kMaxConnShift = 4;          // bits 4..7
constexpr uint32_t kColorShift = 8;            // bits 8..15
constexpr uint32_t kBarrierBit = 1u << 16;
constexpr uint32_t kDetachedBit = 1u << 17;
// particle flags byte
constexpr uint8_t kParticleColorMask = 0x7F;
constexpr uint8_t kParticleDead = 0x80;
// ids of the particles cell radiation creates: this bit, the timestep (bits 32-61) and the local
// index of the emitting cell (bits 0-31)
constexpr unsigned long long kCreatedParticleId = 1ull << 63;
// with this bit too: the particle came from a split (timestep in bits 32-61, index of the particle
// that split in bits 0-31)
constexpr unsigned long long kSplitParticleId = 1ull << 62;

__host__ __device__ inline int flags_num_conn(uint32_t f) { return int(f & kNumConnMask); }
__host__ __device__ inline int flags_max_conn(uint32_t f) { return int((f >> kMaxConnShift) & 0xFu); }
__host__ __device__ inline int flags_color(uint32_t f) { return int((f >> kColorShift) & 0xFFu); }
__host__ __device__ inline bool flags_barrier(uint32_t f) { return (f & kBarrierBit) != 0; }
__host__ __device__ inline int flags_detached(uint32_t f) { return (f & kDetachedBit) ? 1 : 0; }
__host__ __device__ inline uint32_t flags_with_num_conn(uint32_t f, int n) {
    return (f & ~kNumConnMask) | (uint32_t(n) & kNumConnMask);
}

// per-world simulation parameters (physics subset of SimulationParameters)
struct DevParams {
    float dt;
    int motion;  // 0 fluid, 1 collision
    float smoothing, viscosity, pressure;
    float collisionDistance, repulsion;
    float friction, maxVelocity, minDistance, fusionVelocity, maxBindingEnergy;
    float maxBindingDistance[kColors];
    float absorption[kColors];
    // neighbour predicate without the square root: sqrt(s) <= R  <=>  s <= reachSq, where R is the
    // interaction radius (2 h for fluid, cellMaxCollisionDistance for collision) and reachSq the
    // largest float whose correctly rounded square root does not exceed R
    float reachSq;
    // the same for the other length comparisons of the pass: |v| > maxVelocity <=> |v|^2 > maxVelSq,
    // |v| > 0.5 <=> |v|^2 > fastSq, |dv| < fusionVelocity <=> |dv|^2 <= fusionBelowSq,
    // |d| < minDistance <=> |d|^2 <= overlapBelowSq
    float maxVelSq, fastSq, fusionBelowSq, overlapBelowSq;
    float invH, invH2, invH3;  // 1/h, 1/h^2, 1/h^3 (fluid)
    float innerFriction;       // applied every third step to bonded cells (0: off)
    float decayProb;           // cellMaxForceDecayProb (0: off)
    float maxForceSq[kColors];  // |F| > cellMaxForce  <=>  |F|^2 > maxForceSq (as reachSq)
    float minEnergy[kColors];   // cellMinEnergy
    // cell radiation (CellProcessor.cuh:725-773); radiates: radiationProb > 0 and some colour has
    // a positive high-radiation factor or age strength
    int radiates;
    int splits;                  // some colour has a finite particleSplitEnergy
    float splitEnergy[kColors];  // particleSplitEnergy (RadiationProcessor::splitting)
    float radProb, radVelMul, radVelPert;
    int radMinAge[kColors];
    float highRadFactor[kColors], highRadMinEnergy[kColors], radAgeStrength[kColors];
};

// per-world geometry and placement in the set
struct DevWorld {
    int W, H;
    int binSize, nbx, nby;  // cell bins (row-major)
    int binShift;           // log2(binSize) when binSize is a power of two, else -1
    int pnbx, pnby;         // particle bins of kParticleBin units
    int cellOff, numCells;  // hot and cold cell arrays
    int binOff;             // first cell bin in the combined bin arrays
    int pbinOff;            // first particle bin in the combined bin arrays
    int partOff, partCap;   // particle arrays
    int tileX, tileY;       // neighbour pass: bins per strip, staged rows per chunk (launch shape only)
    int ntx, nty;           // strips, chunks
    int rowsPerStage;       // bin rows per staged row of the neighbour pass (1, or 2 for sparse rows)
    int bondWork;           // 1 if a bonded or barrier cell exists or can appear by fusion
    int livingWork;         // 1 if living states are tracked on the device (see WorldSet::build)
    unsigned long long seed;  // key of the world's random draws
    // timestep at which the cells' base ages (ageBase) were taken: a non-barrier cell ages by one
    // every step, so its age at the start of step t is ageBase + (t - ageTs)
    unsigned long long ageTs;
    int active;  // 0: left out of every launch (frozen)
};

// per-world device counters
struct DevCounters {
    unsigned long long timestep;
    int numParticles;  // live particles at [partOff, partOff + numParticles)
    int pad;
};

// global counters of one step (reset at the start of every step)
struct StepCounters {
    int candidates;  // fusion candidate pairs
    int errors;      // bit 0: candidate list overflow; bit 1: map-build look-back gave up;
                     // bit 2: a world's particles exceeded its particle capacity
    int liveParticles;  // energy particles alive after this step's particle sort, all worlds
    // neighbour-pass paths since the data was set (not reset per step): chunks that walked their
    // stencils in global memory, and cells of staged rows that walked theirs alone
    unsigned long long fallbackChunks;
    unsigned long long directCells;
};

// hot per-cell / per-particle state; two of these exist and alternate every step
struct HotBuffers {
    float2* pos;
    float2* vel;
    float* dens;
    uint32_t* flags;
    int* slot;    // local cold slot of the cell (stable identity)
    int* btgt;    // [kMaxBonds][cellCap] local logical index of the bond partner
    float* brest; // [kMaxBonds][cellCap]
    float* bang;  // [kMaxBonds][cellCap]
    float2* ppos;
    float2* pvel;
    float* pen;
    int* plast;   // local cold slot of the last absorbing cell, kNone if none
    int* pslot;   // local particle slot (host-side selection state), kNone for created particles
    uint8_t* pflags;
    unsigned long long* pid;  // particle id
};

// a particle emitted by cell radiation, in its staging slot (one record: one cache line touched to
// write it and one to read it)
struct __align__(16) StagedParticle {
    float2 pos;  // pre-move position
    float2 vel;
    float energy;
    uint32_t color;  // | kStagedPending for a split half not yet in the particle arrays
    unsigned long long id;
};
// a split half waits in its record from the end of one step to the movement of the next (and is
// part of the world's state in between: get_data, statistics and rebuilt sets include it)
constexpr uint32_t kStagedPending = 0x100u;

struct Ctx {
    int numWorlds;
    int cellCap;  // total cells (= sum of numCells)
    int partCap;  // total particle slots
    int numBins;  // cell bins + particle bins of all worlds
    const DevWorld* worlds;
    const DevParams* params;
    DevCounters* ctr;
    StepCounters* sc;
    HotBuffers A;  // state at the start of the step (read)
    HotBuffers B;  // state at the end of the step (written)
    float* energy;  // cold [cellCap], indexed by cellOff + slot
    float* stiff;   // cold [cellCap]
    unsigned long long* cid;  // cold [cellCap]: cell id (key of the cell's random draws)
    uint32_t* ageBase;        // cold [cellCap]: cell age at the world's ageTs
    // Particles emitted by cell radiation during a step are written at partStage + (cell index),
    // behind the worlds' particle ranges (partStage = 0: no staging area), and enter the step's
    // particle map like the live particles.
    int partStage;
    StagedParticle* staged;  // [cellCap]: the particle of staging slot partStage + i at staged[i]
    // Particles split off at the end of a step: staging slot splitStage + k for the particle at k
    // (0: no split staging area), record splitStaged[k]; they enter the next step's map behind the
    // particles that lived and before that step's radiation
    int splitStage;
    StagedParticle* splitStaged;
    int timestepAdvanced;  // 1 if the step's map scan advanced the worlds' timesteps
    // living states of the worlds with livingWork, cold by slot: `living` holds the state at the
    // start of the step, `livingNext` the other buffer (the step's phases alternate between the
    // two and leave the end-of-step state in `livingNext`, the next step's `living`)
    uint8_t* living;
    uint8_t* livingNext;
    uint32_t* activation;     // cold [cellCap]: activationTime (worlds with livingWork)
    uint32_t* creature;       // cold [cellCap]: creatureId (worlds with livingWork)
    uint8_t* stretched;       // [cellCap] end-of-step staged cell had a bond beyond cellMaxBindingDistance
    int* binOf;     // [cellCap] global bin of A-cell i, | kMoverBit if it changed bins
    int* pbinOf;    // [partCap] global bin of A-particle i, -1 if dead
    // Bin sizes.  Cell bins: written by the map count at the start of every step (the cells that
    // land in the bin); particle bins: counted with atomics by the kernels that produce the
    // end-of-step particle positions (and by seeding), brought back to zero by the particle scatter.
    uint32_t* count;      // [numBins] start-of-step state
    uint32_t* countNext;  // [numBins] end-of-step state (next step's `count`)
    int numCellBins;
    uint32_t* start;      // [numBins + 1] exclusive scan of this step's map
    const uint32_t* startPrev;  // [numBins + 1] the last step's map (all zero after seeding)
    // Cells that change bins ("movers"; after seeding every cell) are listed per new cell bin:
    // moverHead[bin] is one of them (A index), moverNext[i] the next (kNone ends a list).
    // moverHeadNext: the lists of the next step (built during this step).
    int* moverHead;
    int* moverHeadNext;
    int* moverNext;  // [cellCap]
    // per cell bin: movers in minus movers out (cell-bin size = last size + delta)
    int* delta;
    int* deltaNext;
    // one-pass map scan: per tile of bins a status word (flag, launch epoch, sum), the launch epoch
    // and the tile ticket
    unsigned long long* mapStatus;
    uint32_t* mapEpoch;
    uint32_t* mapTicket;
    int* sorted;   // [cellCap] map order (bin, then previous index) -> A index
    // [cellCap] the start-of-step state of the cells of collision worlds in map order, written next
    // to `sorted`, so that the row-staged neighbour pass reads each row of bins as one contiguous
    // range (the fluid pass reads the state through `sorted`)
    float2* mpos;
    float2* mvel;
    float* mdens;  // fluid: 1 / density
    uint32_t* mflags;
    int* newidx;   // [cellCap] A index -> new index
    int* psorted;  // [partCap]
    uint8_t* stagedMark;  // [cellCap] end-of-step cell is staged (bonded or barrier at the neighbour pass);
                          // written for the cells of bond-work worlds
    float2* own;    // [cellCap]
    float2* shift;  // [cellCap] Verlet displacement of a staged cell, applied after the second force pass
    float* bstiff;  // [cellCap] stiffness of a bonded cell in end-of-step order (written by the neighbour pass)
    float2* s2;     // [cellCap] connection force of the first pass
    float2* out1;  // [kMaxBonds][cellCap]
    float2* out2;
    uint32_t* valid;  // [cellCap]
    uint8_t* delFlag;
    uint8_t* mark;
    unsigned long long* claim;  // [cellCap]
    int4* cand;                 // [candCap] (a, b, world, live)
    int candCap;
    int* want;        // [partCap]
    int* fclaim;      // [partCap]
    int* absorbCell;  // [partCap]
    int* cclaim;      // [cellCap]
    const int2* cellBlocks;  // per block: (world, first global cell index)
    const int2* bondBlocks;  // the same over the worlds with bond work
    const int2* binBlocks;   // per block: (world, first global bin index) over cell bins
    const int2* pbinBlocks;  // per block: (world, first global bin index) over particle bins
    // per block of partBlockParticles live particles: (world, first global particle index), over
    // the active worlds' live ranges; written by the host when the tables are built and by the
    // particle bin sort for the ranges it publishes (*numPartBlocks entries)
    int2* partBlocks;
    int* numPartBlocks;
    int partBlockParticles;
    const int2* tileBlocks;  // per block: (world, tile index), -1 for padding blocks
    const int2* livingBlocks;  // per block: (world, first global cell index) over livingWork worlds
    int numCellBlocks, cellBlockCells;  // entries of cellBlocks, cells per entry
};

}  // namespace alien
