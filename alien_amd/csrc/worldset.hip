// worldset.hip — device memory, layout, transfer-object conversion and the step loop of a world
// set.  See worldset.hpp.
#include "worldset.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <numeric>

namespace alien {

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            throw Error(ALIEN_E_OOM, std::string(what) + ": " + hipGetErrorString(e));
        throw Error(ALIEN_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
    }
}

int bin_size_for(const alien_params& p) {
    float r = p.motionType == ALIEN_MOTION_FLUID ? p.smoothingLength * 2.0f : p.cellMaxCollisionDistance;
    int s = int(std::ceil(r));
    return s < 2 ? 2 : s;
}

void validate_params(const alien_params& p) {
    if (p.motionType != ALIEN_MOTION_FLUID && p.motionType != ALIEN_MOTION_COLLISION)
        throw Error(ALIEN_E_ARG, "motionType must be ALIEN_MOTION_FLUID or ALIEN_MOTION_COLLISION");
    if (!(p.innerFriction >= 0.0f && p.innerFriction <= 1.0f)) throw Error(ALIEN_E_ARG, "innerFriction must lie in [0, 1]");
    if (!(p.cellMaxForceDecayProb >= 0.0f && p.cellMaxForceDecayProb <= 1.0f))
        throw Error(ALIEN_E_ARG, "cellMaxForceDecayProb must lie in [0, 1]");
    for (int c = 0; c < kColors; ++c) {
        if (!(p.cellMaxForce[c] >= 0.0f)) throw Error(ALIEN_E_ARG, "cellMaxForce must not be negative");
        if (std::isnan(p.cellMinEnergy[c])) throw Error(ALIEN_E_ARG, "cellMinEnergy must be a number");
        if (!(p.highRadiationFactor[c] >= 0.0f) || !(p.radiationCellAgeStrength[c] >= 0.0f) ||
            std::isinf(p.highRadiationFactor[c]) || std::isinf(p.radiationCellAgeStrength[c]))
            throw Error(ALIEN_E_ARG, "radiation factors must be finite and not negative");
        if (std::isnan(p.highRadiationMinCellEnergy[c]))
            throw Error(ALIEN_E_ARG, "highRadiationMinCellEnergy must be a number");
        if (std::isnan(p.particleSplitEnergy[c])) throw Error(ALIEN_E_ARG, "particleSplitEnergy must be a number");
    }
    if (!(p.radiationProb >= 0.0f && p.radiationProb <= 1.0f)) throw Error(ALIEN_E_ARG, "radiationProb must lie in [0, 1]");
    if (!std::isfinite(p.radiationVelocityMultiplier) || !std::isfinite(p.radiationVelocityPerturbation))
        throw Error(ALIEN_E_ARG, "radiation velocity multiplier and perturbation must be finite");
    if (!(p.timestepSize > 0.0f)) throw Error(ALIEN_E_ARG, "timestepSize must be positive");
    if (p.motionType == ALIEN_MOTION_FLUID && !(p.smoothingLength > 0.0f))
        throw Error(ALIEN_E_ARG, "smoothingLength must be positive");
    if (p.motionType == ALIEN_MOTION_COLLISION && !(p.cellMaxCollisionDistance > 0.0f))
        throw Error(ALIEN_E_ARG, "cellMaxCollisionDistance must be positive");
    float r = p.motionType == ALIEN_MOTION_FLUID ? p.smoothingLength * 2.0f : p.cellMaxCollisionDistance;
    if (r > 64.0f) throw Error(ALIEN_E_ARG, "interaction radius above 64 is not supported");
}

void check_set_size(uint64_t cells, uint64_t particles) {
    if (cells > uint64_t(kMaxSetCells))
        throw Error(ALIEN_E_CAPACITY, "more than " + std::to_string(kMaxSetCells) +
                                          " cells in one world set (32-bit bond-plane offsets)");
    if (particles > uint64_t(kMaxSetParticles))
        throw Error(ALIEN_E_CAPACITY, "more than " + std::to_string(kMaxSetParticles) + " energy particles in one world set");
}

void validate_world(const alien_world_desc& d, const alien_params& p) {
    validate_params(p);
    int s = bin_size_for(p);
    if (d.worldSizeX < 3 * s || d.worldSizeY < 3 * s)
        throw Error(ALIEN_E_ARG, "world must span at least three interaction bins per axis");
    if (d.worldSizeX > (1 << 24) || d.worldSizeY > (1 << 24))
        throw Error(ALIEN_E_ARG, "world side above 2^24 is not supported");
    if (int64_t(d.worldSizeX) * int64_t(d.worldSizeY) > (int64_t(1) << 30))
        throw Error(ALIEN_E_ARG, "world area above 2^30 is not supported");
}

// largest float s with sqrtf(s) <= R (sqrtf is correctly rounded and monotone)
static float sqrt_threshold(float R) {
    float s = R * R;
    while (s > 0.0f && std::sqrt(s) > R) s = std::nextafter(s, 0.0f);
    while (std::sqrt(std::nextafter(s, std::numeric_limits<float>::infinity())) <= R)
        s = std::nextafter(s, std::numeric_limits<float>::infinity());
    return s;
}

// largest float s with sqrtf(s) < R (-1 if there is none)
static float sqrt_below(float R) {
    if (!(R > 0.0f)) return -1.0f;
    if (std::isinf(R)) return std::numeric_limits<float>::max();
    float s = R * R;
    while (s > 0.0f && !(std::sqrt(s) < R)) s = std::nextafter(s, 0.0f);
    while (std::sqrt(std::nextafter(s, std::numeric_limits<float>::infinity())) < R)
        s = std::nextafter(s, std::numeric_limits<float>::infinity());
    return std::sqrt(s) < R ? s : -1.0f;
}

DevParams to_dev_params(const alien_params& p) {
    DevParams d{};
    d.dt = p.timestepSize;
    d.motion = p.motionType == ALIEN_MOTION_FLUID ? 0 : 1;
    d.smoothing = p.smoothingLength;
    d.viscosity = p.viscosityStrength;
    d.pressure = p.pressureStrength;
    d.collisionDistance = p.cellMaxCollisionDistance;
    d.repulsion = p.cellRepulsionStrength;
    d.friction = p.friction;
    d.maxVelocity = p.cellMaxVelocity;
    d.minDistance = p.cellMinDistance;
    d.fusionVelocity = p.cellFusionVelocity;
    d.maxBindingEnergy = p.cellMaxBindingEnergy;
    for (int c = 0; c < kColors; ++c) {
        d.maxBindingDistance[c] = p.cellMaxBindingDistance[c];
        d.absorption[c] = p.radiationAbsorption[c];
    }
    d.reachSq = sqrt_threshold(d.motion == 0 ? d.smoothing * 2.0f : d.collisionDistance);
    d.maxVelSq = sqrt_threshold(d.maxVelocity);
    d.fastSq = sqrt_threshold(0.5f);
    d.fusionBelowSq = sqrt_below(d.fusionVelocity);
    d.overlapBelowSq = sqrt_below(d.minDistance);
    d.innerFriction = p.innerFriction;
    d.decayProb = p.cellMaxForceDecayProb;
    for (int c = 0; c < kColors; ++c) {
        d.maxForceSq[c] = std::isinf(p.cellMaxForce[c]) ? std::numeric_limits<float>::infinity()
                                                        : sqrt_threshold(p.cellMaxForce[c]);
        d.minEnergy[c] = p.cellMinEnergy[c];
    }
    d.radiates = radiates(p) ? 1 : 0;
    d.radProb = p.radiationProb;
    d.radVelMul = p.radiationVelocityMultiplier;
    d.radVelPert = p.radiationVelocityPerturbation;
    d.splits = splits(p) ? 1 : 0;
    for (int c = 0; c < kColors; ++c) {
        d.splitEnergy[c] = p.particleSplitEnergy[c];
        d.radMinAge[c] = p.radiationMinCellAge[c];
        d.highRadFactor[c] = p.highRadiationFactor[c];
        d.highRadMinEnergy[c] = p.highRadiationMinCellEnergy[c];
        d.radAgeStrength[c] = p.radiationCellAgeStrength[c];
    }
    d.invH = 1.0f / d.smoothing;
    d.invH2 = 1.0f / (d.smoothing * d.smoothing);
    d.invH3 = 1.0f / (d.smoothing * d.smoothing * d.smoothing);
    return d;
}

// BaseMap::correctPosition(float2&) (Map.cuh:20-26) on the host, same operations as the device
static inline int host_floor_int(float v) {
    int r = static_cast<int>(v);
    if (static_cast<float>(r) > v) --r;
    return r;
}
static inline void host_wrap(float& x, float& y, int W, int H) {
    int ix = host_floor_int(x), iy = host_floor_int(y);
    float fx = x - static_cast<float>(ix), fy = y - static_cast<float>(iy);
    ix = ((ix % W) + W) % W;
    iy = ((iy % H) + H) % H;
    x = static_cast<float>(ix) + fx;
    y = static_cast<float>(iy) + fy;
}

// ---------------------------------------------------------------------------------------------
// WorldImage <-> transfer objects
// ---------------------------------------------------------------------------------------------
WorldImage WorldImage::fromTO(const alien_world_desc& desc, const alien_params& params, const alien_cell_to* cells,
                              uint64_t numCells, const alien_particle_to* particles, uint64_t numParticles) {
    validate_world(desc, params);
    check_set_size(numCells, numParticles);
    if ((numCells && !cells) || (numParticles && !particles)) throw Error(ALIEN_E_ARG, "null transfer array");
    WorldImage img;
    const int n = int(numCells), np = int(numParticles);
    const int W = desc.worldSizeX, H = desc.worldSizeY;
    HostWorld& h = img.host;
    h.desc = desc;
    h.params = params;
    h.numCells = n;
    h.particleCap = np;
    h.cellId.resize(n);
    h.age.resize(n);
    h.creatureId.resize(n);
    h.mutationId.resize(n);
    h.activationTime.resize(n);
    h.livingState.resize(n);
    h.cellFunction.resize(n);
    h.ancestorMutationId.resize(n);
    h.selected.resize(n);
    h.genomeComplexity.resize(n);
    img.pos.resize(n);
    img.vel.resize(n);
    img.dens.assign(n, 1.0f);
    img.flags.resize(n);
    img.slot.resize(n);
    img.btgt.assign(size_t(kMaxBonds) * n, 0);
    img.brest.assign(size_t(kMaxBonds) * n, 0.0f);
    img.bang.assign(size_t(kMaxBonds) * n, 0.0f);
    img.energy.resize(n);
    img.stiff.resize(n);
    // Slots (the cells' stable identities, which index their cold state) are given in spatial
    // order, bin by bin of the cell map: the map keeps cells in bin order, so the cold state of
    // neighbouring map entries then lies close in memory.  Only locality depends on it.
    {
        const int bs = bin_size_for(params);
        const int nbx = std::max(1, W / bs), nby = std::max(1, H / bs);
        std::vector<int> key(n), cnt(size_t(nbx) * nby + 1, 0);
        for (int i = 0; i < n; ++i) {
            float x = cells[i].pos[0], y = cells[i].pos[1];
            int k = 0;
            if (std::isfinite(x) && std::isfinite(y)) {
                host_wrap(x, y, W, H);
                const int bx = std::min(int(x) / bs, nbx - 1), by = std::min(int(y) / bs, nby - 1);
                k = std::max(0, by) * nbx + std::max(0, bx);
            }
            key[i] = k;
            ++cnt[size_t(k) + 1];
        }
        for (size_t b = 1; b < cnt.size(); ++b) cnt[b] += cnt[b - 1];
        for (int i = 0; i < n; ++i) img.slot[i] = cnt[size_t(key[i])]++;
    }
    for (int i = 0; i < n; ++i) {
        const alien_cell_to& c = cells[i];
        if (c.numConnections > kMaxBonds) throw Error(ALIEN_E_ARG, "numConnections above 6");
        if (c.color >= kColors) throw Error(ALIEN_E_ARG, "cell color must be below 7 (MAX_COLORS)");
        float x = c.pos[0], y = c.pos[1];
        if (!std::isfinite(x) || !std::isfinite(y)) throw Error(ALIEN_E_ARG, "non-finite cell position");
        host_wrap(x, y, W, H);
        img.pos[i] = make_float2(x, y);
        img.vel[i] = make_float2(c.vel[0], c.vel[1]);
        uint32_t maxConn = c.maxConnections > 15 ? 15u : c.maxConnections;
        img.flags[i] = uint32_t(c.numConnections) | (maxConn << kMaxConnShift) | (uint32_t(c.color) << kColorShift) |
                       (c.barrier ? kBarrierBit : 0u);
        const int sl = img.slot[i];
        for (int b = 0; b < c.numConnections; ++b) {
            int t = c.connections[b].cellIndex;
            if (t < 0 || t >= n) throw Error(ALIEN_E_ARG, "bond cellIndex out of range");
            img.btgt[size_t(b) * n + i] = t;
            img.brest[size_t(b) * n + i] = c.connections[b].distance;
            img.bang[size_t(b) * n + i] = c.connections[b].angleFromPrevious;
        }
        img.energy[sl] = c.energy;
        img.stiff[sl] = c.stiffness;
        for (int b = 0; b < c.numConnections; ++b) {  // bonds come in pairs a->b, b->a (TOs.cuh:30-35)
            const int t = c.connections[b].cellIndex;
            const alien_cell_to& o = cells[t];
            int back = 0, same = 0;
            for (int k = 0; k < o.numConnections && k < kMaxBonds; ++k) back += o.connections[k].cellIndex == i;
            for (int k = 0; k < c.numConnections; ++k) same += c.connections[k].cellIndex == t;
            if (t == i || back != 1 || same != 1) throw Error(ALIEN_E_ARG, "bonds must be symmetric and unique");
        }
        h.cellId[sl] = c.id;
        h.age[sl] = c.age;
        h.creatureId[sl] = c.creatureId;
        h.mutationId[sl] = c.mutationId;
        h.activationTime[sl] = c.activationTime;
        h.livingState[sl] = c.livingState;
        h.cellFunction[sl] = c.cellFunction;
        h.ancestorMutationId[sl] = c.ancestorMutationId;
        h.selected[sl] = c.selected;
        h.genomeComplexity[sl] = c.genomeComplexity;
    }
    img.ppos.resize(np);
    img.pvel.resize(np);
    img.pen.resize(np);
    img.plast.assign(np, kNone);
    img.pslot.resize(np);
    img.pflags.resize(np);
    img.pid.resize(np);
    h.particleId.resize(np);
    h.particleSelected.resize(np);
    for (int i = 0; i < np; ++i) {
        const alien_particle_to& p = particles[i];
        if (p.color >= kColors) throw Error(ALIEN_E_ARG, "particle color must be below 7 (MAX_COLORS)");
        float x = p.pos[0], y = p.pos[1];
        if (!std::isfinite(x) || !std::isfinite(y)) throw Error(ALIEN_E_ARG, "non-finite particle position");
        host_wrap(x, y, W, H);
        img.ppos[i] = make_float2(x, y);
        img.pvel[i] = make_float2(p.vel[0], p.vel[1]);
        img.pen[i] = p.energy;
        img.pslot[i] = i;
        img.pflags[i] = uint8_t(p.color & kParticleColorMask);
        img.pid[i] = p.id;
        h.particleId[i] = p.id;
        h.particleSelected[i] = p.selected;
    }
    img.numParticles = np;
    // Living states need tracking on the device only if some cell is not trivially Ready: a state
    // other than Ready, an activation count, (non-barrier) energy below cellMinEnergy, or cell
    // radiation, the one process of this path that takes energy from cells.  A world without any
    // of these keeps every cell Ready with no count for good (its energies only grow).
    img.livingWork = radiates(params);
    for (int i = 0; i < n && !img.livingWork; ++i) {
        const alien_cell_to& c = cells[i];
        img.livingWork = c.livingState != 0 || c.activationTime != 0 ||
                         (!c.barrier && c.energy < params.cellMinEnergy[c.color]);
    }
    return img;
}

void WorldImage::updateLivingWork() {
    const HostWorld& h = host;
    if (radiates(h.params)) livingWork = true;
    for (int i = 0; i < h.numCells && !livingWork; ++i) {
        const int sl = slot[size_t(i)];
        const uint32_t f = flags[size_t(i)];
        livingWork = h.livingState[sl] != 0 || h.activationTime[sl] != 0 ||
                     (!flags_barrier(f) && energy[sl] < h.params.cellMinEnergy[flags_color(f)]);
    }
}

void WorldImage::toTO(std::vector<alien_cell_to>& cells, std::vector<alien_particle_to>& particles) const {
    const HostWorld& h = host;
    const int n = h.numCells;
    // ascending id (ties by index), sorted as compact (id, index) pairs
    std::vector<std::pair<uint64_t, int>> keys(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) keys[size_t(i)] = {h.cellId[slot[i]], i};
    std::sort(keys.begin(), keys.end());
    std::vector<int> order(n);
    for (int k = 0; k < n; ++k) order[k] = keys[size_t(k)].second;
    std::vector<int> posOf(n);
    for (int k = 0; k < n; ++k) posOf[order[k]] = k;
    cells.assign(n, alien_cell_to{});
    for (int k = 0; k < n; ++k) {
        const int i = order[k];
        const int sl = slot[i];
        alien_cell_to& c = cells[k];
        const uint32_t f = flags[i];
        c.id = h.cellId[sl];
        c.pos[0] = pos[i].x;
        c.pos[1] = pos[i].y;
        c.vel[0] = vel[i].x;
        c.vel[1] = vel[i].y;
        c.energy = energy[sl];
        c.stiffness = stiff[sl];
        c.color = uint8_t(flags_color(f));
        c.maxConnections = uint8_t(flags_max_conn(f));
        c.numConnections = uint8_t(flags_num_conn(f));
        c.barrier = flags_barrier(f) ? 1 : 0;
        for (int b = 0; b < c.numConnections; ++b) {
            c.connections[b].cellIndex = posOf[btgt[size_t(b) * n + i]];
            c.connections[b].distance = brest[size_t(b) * n + i];
            c.connections[b].angleFromPrevious = bang[size_t(b) * n + i];
        }
        c.age = h.age[sl];
        c.livingState = h.livingState[sl];
        c.activationTime = h.activationTime[sl];
        c.cellFunction = h.cellFunction[sl];
        c.ancestorMutationId = h.ancestorMutationId[sl];
        c.selected = h.selected[sl];
        c.creatureId = h.creatureId[sl];
        c.mutationId = h.mutationId[sl];
        c.genomeComplexity = h.genomeComplexity[sl];
    }
    const int np = numParticles;
    std::vector<std::pair<uint64_t, int>> pkeys(static_cast<size_t>(np));
    for (int i = 0; i < np; ++i) pkeys[size_t(i)] = {pid[i], i};
    std::sort(pkeys.begin(), pkeys.end());
    std::vector<int> po(np);
    for (int k = 0; k < np; ++k) po[k] = pkeys[size_t(k)].second;
    particles.assign(np, alien_particle_to{});
    for (int k = 0; k < np; ++k) {
        const int i = po[k];
        alien_particle_to& p = particles[k];
        p.id = pid[i];
        p.energy = pen[i];
        p.pos[0] = ppos[i].x;
        p.pos[1] = ppos[i].y;
        p.vel[0] = pvel[i].x;
        p.vel[1] = pvel[i].y;
        p.color = uint8_t(pflags[i] & kParticleColorMask);
        p.selected = pslot[i] >= 0 ? h.particleSelected[pslot[i]] : 0;
    }
}

// ---------------------------------------------------------------------------------------------
// WorldSet
// ---------------------------------------------------------------------------------------------
class WorldSet::EventRecorder : public Recorder {
public:
    explicit EventRecorder(hipStream_t s) : stream_(s) {}
    ~EventRecorder() override {
        for (auto e : pool_) (void)hipEventDestroy(e);
    }
    void mark(const char* name) override {
        if (used_ == pool_.size()) {
            hipEvent_t e;
            check_hip(hipEventCreate(&e), "hipEventCreate");
            pool_.push_back(e);
        }
        check_hip(hipEventRecord(pool_[used_], stream_), "hipEventRecord");
        names_.push_back(name);
        ++used_;
    }
    // elapsed time between consecutive marks, attributed to the kernel launched after the first
    void collect(std::map<std::string, KernelTime>& acc) {
        check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        for (size_t i = 0; i + 1 < used_; ++i) {
            if (!names_[i]) continue;
            float ms = 0.0f;
            check_hip(hipEventElapsedTime(&ms, pool_[i], pool_[i + 1]), "hipEventElapsedTime");
            KernelTime& k = acc[names_[i]];
            k.launches += 1;
            k.totalMs += double(ms);
        }
        used_ = 0;
        names_.clear();
    }

private:
    hipStream_t stream_;
    std::vector<hipEvent_t> pool_;
    std::vector<const char*> names_;
    size_t used_ = 0;
};

// both streams at the default priority
WorldSet::WorldSet() {
    check_hip(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipStreamCreateWithFlags(&side_.stream, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipEventCreateWithFlags(&side_.fork, hipEventDisableTiming), "hipEventCreate");
    check_hip(hipEventCreateWithFlags(&side_.mapped, hipEventDisableTiming), "hipEventCreate");
    check_hip(hipEventCreateWithFlags(&side_.join, hipEventDisableTiming), "hipEventCreate");
    check_hip(hipEventCreateWithFlags(&liveReady_, hipEventDisableTiming), "hipEventCreate");
    check_hip(hipHostMalloc(reinterpret_cast<void**>(&liveHost_), sizeof(int), hipHostMallocDefault), "hipHostMalloc");
}

WorldSet::~WorldSet() {
    invalidateGraphs();
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (side_.stream) (void)hipStreamSynchronize(side_.stream);
    freeAll();
    for (auto& a : tableAllocs_) (void)hipFree(a.p);
    tableAllocs_.clear();
    if (liveReady_) (void)hipEventDestroy(liveReady_);
    if (liveHost_) (void)hipHostFree(liveHost_);
    if (side_.fork) (void)hipEventDestroy(side_.fork);
    if (side_.mapped) (void)hipEventDestroy(side_.mapped);
    if (side_.join) (void)hipEventDestroy(side_.join);
    if (side_.stream) (void)hipStreamDestroy(side_.stream);
    if (stream_) (void)hipStreamDestroy(stream_);
}

void* WorldSet::alloc(size_t bytes) {
    if (bytes == 0) bytes = 64;
    bytes = (bytes + 255) & ~size_t(255);
    void* p = nullptr;
    check_hip(hipMalloc(&p, bytes), "hipMalloc");
    allocs_.push_back({p, bytes});
    deviceBytes_ += bytes;
    return p;
}

void WorldSet::freeAll() {
    for (auto& a : allocs_) (void)hipFree(a.p);
    allocs_.clear();
    deviceBytes_ = 0;
}

void WorldSet::invalidateGraphs() {
    for (auto* set : {graph_, graphMulti_})
        for (int p = 0; p < 2; ++p)
            if (set[p]) {
                (void)hipGraphExecDestroy(set[p]);
                set[p] = nullptr;
            }
}

void WorldSet::synchronize() { check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize"); }

uint64_t WorldSet::estimateBytes(uint64_t cells, uint64_t particles, int W, int H, int binSize) {
    const uint64_t perCell = 2 * (8 + 8 + 4 + 4 + 4 + kMaxBonds * 12) + 8 + 4 * 4 + 8 + 3 * 8 + 2 * kMaxBonds * 8 + 4 +
                             2 + 8 + 4 * 16 + 4 + (8 + 8 + 4 + 4) + 4 + 4;
    const uint64_t perParticle = 2 * (8 + 8 + 4 + 4 + 4 + 1 + 8) + 6 * 4;
    const uint64_t bins = uint64_t(std::max(1, W / binSize)) * uint64_t(std::max(1, H / binSize)) +
                          uint64_t(std::max(1, W / kParticleBin)) * uint64_t(std::max(1, H / kParticleBin));
    // per bin: two histograms, two offset arrays, two mover-list heads and deltas (cell bins), scan
    // partials
    return cells * perCell + particles * perParticle + bins * 36 + (1u << 20);
}

void WorldSet::clearAll() {
    invalidateGraphs();
    freeAll();
    hosts_.clear();
    worlds_.clear();
    params_.clear();
    densityPending_.clear();
    cellCap_ = partCap_ = numBins_ = candCap_ = 0;
    parity_ = 0;
}

int stageLayout(int64_t caps, int cells, bool radiation, bool splitting, int& splitStage, int& radStage);

void WorldSet::build(const std::vector<Source>& sources) {
    // collect the new layout first (sources may live in this very set's old buffers only via a
    // different WorldSet object)
    std::vector<HostWorld> hosts;
    std::vector<DevWorld> worlds;
    std::vector<DevParams> params;
    std::vector<uint64_t> timesteps;
    std::vector<int> liveParticles;
    std::vector<uint8_t> pending;
    std::vector<uint8_t> bondWork, livingWork;
    std::vector<int> stripChoice, stageChoice;
    // A set is built once, into fresh allocations; a failed build then leaves nothing behind (the
    // destructor frees what was allocated).  Sizes are checked before anything is read or allocated.
    if (!allocs_.empty() || !worlds_.empty())
        throw Error(ALIEN_E_STATE, "a world set is built once; build a fresh set to change its worlds");
    {
        uint64_t cells = 0, parts = 0;
        for (const auto& src : sources) {
            if (src.from == this) throw Error(ALIEN_E_ARG, "a world set cannot be rebuilt from itself");
            if (src.image) {
                cells += uint64_t(src.image->host.numCells);
                parts += uint64_t(src.image->numParticles);
            } else {
                cells += uint64_t(src.from->worlds_.at(src.fromIndex).numCells);
                parts += uint64_t(src.from->worlds_.at(src.fromIndex).partCap);
            }
        }
        check_set_size(cells, parts);
    }
    int64_t cellOff = 0, partOff = 0, bins = 0;
    for (const auto& src : sources) {
        HostWorld h;
        int live = 0;
        uint64_t ts = 0;
        bool pend = false;
        int strip = kStripMax, stage = 1;
        bool work = false, living = false;
        if (src.image) {
            h = src.image->host;
            live = src.image->numParticles;
            ts = src.image->timestep;
            pend = src.image->densityResetPending;
            living = src.image->livingWork;
            const WorldImage& im = *src.image;
            for (int i = 0; i < h.numCells && !work; ++i) {
                // bonds can appear only by fusion, which needs a free bond slot
                if (flags_num_conn(im.flags[i]) > 0 || flags_barrier(im.flags[i]) || flags_max_conn(im.flags[i]) > 0)
                    work = true;
            }
            const int bs = bin_size_for(h.params);
            strip = choose_strip(im.pos, h.desc.worldSizeX, h.desc.worldSizeY, bs,
                                 h.params.motionType == ALIEN_MOTION_COLLISION, stage);
        } else {
            h = src.from->hosts_.at(src.fromIndex);
            live = src.from->numParticles(src.fromIndex);
            ts = src.from->timestep(src.fromIndex);
            pend = src.from->densityPending_.at(src.fromIndex) != 0;
            strip = src.from->stripChoice_.at(src.fromIndex);
            stage = src.from->stageChoice_.at(src.fromIndex);
            work = src.from->bondWork_.at(src.fromIndex) != 0;
            living = src.from->livingWork_.at(src.fromIndex) != 0;
        }
        validate_world(h.desc, h.params);
        h.particleCap = live;
        DevWorld w{};
        w.W = h.desc.worldSizeX;
        w.H = h.desc.worldSizeY;
        w.binSize = bin_size_for(h.params);
        w.binShift = -1;
        for (int sh = 0; sh < 24; ++sh)
            if ((1 << sh) == w.binSize) w.binShift = sh;
        w.nbx = w.W / w.binSize;
        w.nby = w.H / w.binSize;
        w.pnbx = std::max(1, w.W / kParticleBin);
        w.pnby = std::max(1, w.H / kParticleBin);
        w.cellOff = int(cellOff);
        w.numCells = h.numCells;
        w.binOff = int(bins);
        w.partOff = int(partOff);
        const int64_t cap = particle_room(h.numCells, live, h.params);
        if (partOff + cap > kMaxSetParticles)
            throw Error(ALIEN_E_CAPACITY, "the particle room of the radiating worlds exceeds the particles of one world set");
        w.partCap = int(cap);
        w.ageTs = ts - h.stepsSinceSet;
        w.active = 1;
        w.bondWork = work ? 1 : 0;
        w.livingWork = living ? 1 : 0;
        w.seed = h.desc.randomSeed;
        strip_shape(w, std::min(strip, stripCap_), stage);
        stripChoice.push_back(strip);
        stageChoice.push_back(stage);
        cellOff += h.numCells;
        partOff += w.partCap;
        bins += int64_t(w.nbx) * w.nby;
        if (bins + int64_t(w.pnbx) * w.pnby >= (int64_t(1) << 31) - 2)
            throw Error(ALIEN_E_CAPACITY, "too many spatial bins in one world set");
        hosts.push_back(std::move(h));
        worlds.push_back(w);
        params.push_back(to_dev_params(hosts.back().params));
        timesteps.push_back(ts);
        liveParticles.push_back(live);
        pending.push_back(pend ? 1 : 0);
        bondWork.push_back(work ? 1 : 0);
        livingWork.push_back(living ? 1 : 0);
    }
    const int cellBins = int(bins);  // cell bins of all worlds come first, then the particle bins
    for (auto& w : worlds) {
        w.pbinOff = int(bins);
        bins += int64_t(w.pnbx) * w.pnby;
    }
    if (bins >= (int64_t(1) << 31) - 2) throw Error(ALIEN_E_CAPACITY, "too many spatial bins in one world set");

    WorldSet& me = *this;
    invalidateGraphs();

    bool anyRadiates = false, anySplits = false;
    for (const auto& h : hosts) {
        anyRadiates = anyRadiates || radiates(h.params);
        anySplits = anySplits || splits(h.params);
    }
    const int NC = int(std::max<int64_t>(cellOff, 1));
    const int NP = stageLayout(partOff, NC, anyRadiates, anySplits, splitStage_, partStage_);
    const int NB = int(std::max<int64_t>(bins, 1));
    cellCap_ = NC;
    partCap_ = NP;
    numBins_ = NB;
    candCap_ = std::max(4 * NC, 4096);
    auto A = [&](size_t bytes) { return me.alloc(bytes); };
    dWorlds_ = static_cast<DevWorld*>(A(sizeof(DevWorld) * std::max<size_t>(worlds.size(), 1)));
    dParams_ = static_cast<DevParams*>(A(sizeof(DevParams) * std::max<size_t>(worlds.size(), 1)));
    dCtr_ = static_cast<DevCounters*>(A(sizeof(DevCounters) * std::max<size_t>(worlds.size(), 1)));
    dSc_ = static_cast<StepCounters*>(A(sizeof(StepCounters)));
    for (int b = 0; b < 2; ++b) {
        HotBuffers& h = buf_[b];
        h.pos = static_cast<float2*>(A(sizeof(float2) * NC));
        h.vel = static_cast<float2*>(A(sizeof(float2) * NC));
        h.dens = static_cast<float*>(A(sizeof(float) * NC));
        h.flags = static_cast<uint32_t*>(A(sizeof(uint32_t) * NC));
        h.slot = static_cast<int*>(A(sizeof(int) * NC));
        h.btgt = static_cast<int*>(A(sizeof(int) * kMaxBonds * size_t(NC)));
        h.brest = static_cast<float*>(A(sizeof(float) * kMaxBonds * size_t(NC)));
        h.bang = static_cast<float*>(A(sizeof(float) * kMaxBonds * size_t(NC)));
        h.ppos = static_cast<float2*>(A(sizeof(float2) * NP));
        h.pvel = static_cast<float2*>(A(sizeof(float2) * NP));
        h.pen = static_cast<float*>(A(sizeof(float) * NP));
        h.plast = static_cast<int*>(A(sizeof(int) * NP));
        h.pslot = static_cast<int*>(A(sizeof(int) * NP));
        h.pflags = static_cast<uint8_t*>(A(NP));
        h.pid = static_cast<unsigned long long*>(A(sizeof(unsigned long long) * NP));
    }
    energy_ = static_cast<float*>(A(sizeof(float) * NC));
    stiff_ = static_cast<float*>(A(sizeof(float) * NC));
    cid_ = static_cast<unsigned long long*>(A(sizeof(unsigned long long) * NC));
    ageBase_ = static_cast<uint32_t*>(A(sizeof(uint32_t) * NC));
    staged_ = partStage_ > 0 ? static_cast<StagedParticle*>(A(sizeof(StagedParticle) * NC)) : nullptr;
    splitStaged_ = splitStage_ > 0 ? static_cast<StagedParticle*>(A(sizeof(StagedParticle) * size_t(splitStage_))) : nullptr;
    living_[0] = static_cast<uint8_t*>(A(NC));
    living_[1] = static_cast<uint8_t*>(A(NC));
    activation_ = static_cast<uint32_t*>(A(sizeof(uint32_t) * NC));
    creature_ = static_cast<uint32_t*>(A(sizeof(uint32_t) * NC));
    stretched_ = static_cast<uint8_t*>(A(NC));
    binOf_ = static_cast<int*>(A(sizeof(int) * NC));
    pbinOf_ = static_cast<int*>(A(sizeof(int) * NP));
    count_[0] = static_cast<uint32_t*>(A(sizeof(uint32_t) * NB));
    count_[1] = static_cast<uint32_t*>(A(sizeof(uint32_t) * NB));
    numCellBins_ = cellBins;
    const size_t NCB = size_t(std::max(cellBins, 1));
    for (int b = 0; b < 2; ++b) {
        start_[b] = static_cast<uint32_t*>(A(sizeof(uint32_t) * (size_t(NB) + 1)));
        moverHead_[b] = static_cast<int*>(A(sizeof(int) * NCB));
        delta_[b] = static_cast<int*>(A(sizeof(int) * NCB));
    }
    moverNext_ = static_cast<int*>(A(sizeof(int) * NC));
    mapStatus_ = static_cast<unsigned long long*>(A(sizeof(unsigned long long) * size_t(map_scan_tiles(NB))));
    mapEpochTicket_ = static_cast<uint32_t*>(A(sizeof(uint32_t) * 2));
    sorted_ = static_cast<int*>(A(sizeof(int) * NC));
    mpos_ = static_cast<float2*>(A(sizeof(float2) * NC));
    mvel_ = static_cast<float2*>(A(sizeof(float2) * NC));
    mdens_ = static_cast<float*>(A(sizeof(float) * NC));
    mflags_ = static_cast<uint32_t*>(A(sizeof(uint32_t) * NC));
    newidx_ = static_cast<int*>(A(sizeof(int) * NC));
    psorted_ = static_cast<int*>(A(sizeof(int) * NP));
    stagedMark_ = static_cast<uint8_t*>(A(NC));
    own_ = static_cast<float2*>(A(sizeof(float2) * NC));
    shift_ = static_cast<float2*>(A(sizeof(float2) * NC));
    bstiff_ = static_cast<float*>(A(sizeof(float) * NC));
    s2_ = static_cast<float2*>(A(sizeof(float2) * NC));
    out1_ = static_cast<float2*>(A(sizeof(float2) * kMaxBonds * size_t(NC)));
    out2_ = static_cast<float2*>(A(sizeof(float2) * kMaxBonds * size_t(NC)));
    valid_ = static_cast<uint32_t*>(A(sizeof(uint32_t) * NC));
    delFlag_ = static_cast<uint8_t*>(A(NC));
    mark_ = static_cast<uint8_t*>(A(NC));
    claim_ = static_cast<unsigned long long*>(A(sizeof(unsigned long long) * NC));
    cand_ = static_cast<int4*>(A(sizeof(int4) * size_t(candCap_)));
    want_ = static_cast<int*>(A(sizeof(int) * NP));
    fclaim_ = static_cast<int*>(A(sizeof(int) * NP));
    absorbCell_ = static_cast<int*>(A(sizeof(int) * NP));
    cclaim_ = static_cast<int*>(A(sizeof(int) * NC));

    // copy world states into buffer 0 (and, from other sets, the split halves still waiting)
    struct PendingCopy {
        size_t world;
        int fromOff;
        const StagedParticle* from;
        int count;
    };
    std::vector<PendingCopy> pendingCopies;
    HotBuffers& D = buf_[0];
    for (size_t wi = 0; wi < sources.size(); ++wi) {
        const auto& src = sources[wi];
        const DevWorld& w = worlds[wi];
        const int n = w.numCells, np = liveParticles[wi];
        auto h2d = [&](void* dst, const void* s, size_t bytes) {
            if (bytes) check_hip(hipMemcpyAsync(dst, s, bytes, hipMemcpyHostToDevice, stream_), "hipMemcpy H2D");
        };
        auto d2d = [&](void* dst, const void* s, size_t bytes) {
            if (bytes) check_hip(hipMemcpyAsync(dst, s, bytes, hipMemcpyDeviceToDevice, stream_), "hipMemcpy D2D");
        };
        if (src.image) {
            const WorldImage& im = *src.image;
            h2d(D.pos + w.cellOff, im.pos.data(), sizeof(float2) * n);
            h2d(D.vel + w.cellOff, im.vel.data(), sizeof(float2) * n);
            h2d(D.dens + w.cellOff, im.dens.data(), sizeof(float) * n);
            h2d(D.flags + w.cellOff, im.flags.data(), sizeof(uint32_t) * n);
            h2d(D.slot + w.cellOff, im.slot.data(), sizeof(int) * n);
            for (int b = 0; b < kMaxBonds; ++b) {
                h2d(D.btgt + size_t(b) * NC + w.cellOff, im.btgt.data() + size_t(b) * n, sizeof(int) * n);
                h2d(D.brest + size_t(b) * NC + w.cellOff, im.brest.data() + size_t(b) * n, sizeof(float) * n);
                h2d(D.bang + size_t(b) * NC + w.cellOff, im.bang.data() + size_t(b) * n, sizeof(float) * n);
            }
            h2d(energy_ + w.cellOff, im.energy.data(), sizeof(float) * n);
            h2d(stiff_ + w.cellOff, im.stiff.data(), sizeof(float) * n);
            static_assert(sizeof(unsigned long long) == sizeof(uint64_t), "cell ids");
            h2d(cid_ + w.cellOff, im.host.cellId.data(), sizeof(uint64_t) * n);
            h2d(ageBase_ + w.cellOff, im.host.age.data(), sizeof(uint32_t) * n);
            h2d(living_[0] + w.cellOff, im.host.livingState.data(), n);
            h2d(activation_ + w.cellOff, im.host.activationTime.data(), sizeof(uint32_t) * n);
            h2d(creature_ + w.cellOff, im.host.creatureId.data(), sizeof(uint32_t) * n);
            h2d(D.ppos + w.partOff, im.ppos.data(), sizeof(float2) * np);
            h2d(D.pvel + w.partOff, im.pvel.data(), sizeof(float2) * np);
            h2d(D.pen + w.partOff, im.pen.data(), sizeof(float) * np);
            h2d(D.plast + w.partOff, im.plast.data(), sizeof(int) * np);
            h2d(D.pslot + w.partOff, im.pslot.data(), sizeof(int) * np);
            h2d(D.pflags + w.partOff, im.pflags.data(), np);
            static_assert(sizeof(unsigned long long) == sizeof(uint64_t), "particle ids");
            h2d(D.pid + w.partOff, im.pid.data(), sizeof(uint64_t) * np);
            // the staging vectors must outlive the async copies
            check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
        } else {
            WorldSet& o = *src.from;
            o.synchronize();
            const DevWorld& ow = o.worlds_.at(src.fromIndex);
            const HotBuffers& S = o.buf_[o.parity_];
            const int oNC = o.cellCap_;
            d2d(D.pos + w.cellOff, S.pos + ow.cellOff, sizeof(float2) * n);
            d2d(D.vel + w.cellOff, S.vel + ow.cellOff, sizeof(float2) * n);
            d2d(D.dens + w.cellOff, S.dens + ow.cellOff, sizeof(float) * n);
            d2d(D.flags + w.cellOff, S.flags + ow.cellOff, sizeof(uint32_t) * n);
            d2d(D.slot + w.cellOff, S.slot + ow.cellOff, sizeof(int) * n);
            for (int b = 0; b < kMaxBonds; ++b) {
                d2d(D.btgt + size_t(b) * NC + w.cellOff, S.btgt + size_t(b) * oNC + ow.cellOff, sizeof(int) * n);
                d2d(D.brest + size_t(b) * NC + w.cellOff, S.brest + size_t(b) * oNC + ow.cellOff, sizeof(float) * n);
                d2d(D.bang + size_t(b) * NC + w.cellOff, S.bang + size_t(b) * oNC + ow.cellOff, sizeof(float) * n);
            }
            d2d(energy_ + w.cellOff, o.energy_ + ow.cellOff, sizeof(float) * n);
            d2d(stiff_ + w.cellOff, o.stiff_ + ow.cellOff, sizeof(float) * n);
            d2d(cid_ + w.cellOff, o.cid_ + ow.cellOff, sizeof(unsigned long long) * n);
            d2d(ageBase_ + w.cellOff, o.ageBase_ + ow.cellOff, sizeof(uint32_t) * n);
            d2d(living_[0] + w.cellOff, o.living_[o.parity_] + ow.cellOff, n);
            d2d(activation_ + w.cellOff, o.activation_ + ow.cellOff, sizeof(uint32_t) * n);
            d2d(creature_ + w.cellOff, o.creature_ + ow.cellOff, sizeof(uint32_t) * n);
            d2d(D.ppos + w.partOff, S.ppos + ow.partOff, sizeof(float2) * np);
            d2d(D.pvel + w.partOff, S.pvel + ow.partOff, sizeof(float2) * np);
            d2d(D.pen + w.partOff, S.pen + ow.partOff, sizeof(float) * np);
            d2d(D.plast + w.partOff, S.plast + ow.partOff, sizeof(int) * np);
            d2d(D.pslot + w.partOff, S.pslot + ow.partOff, sizeof(int) * np);
            d2d(D.pflags + w.partOff, S.pflags + ow.partOff, np);
            d2d(D.pid + w.partOff, S.pid + ow.partOff, sizeof(unsigned long long) * np);
            pendingCopies.push_back({wi, int(ow.partOff), o.splitStaged_, np});
        }
    }
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");

    hosts_ = std::move(hosts);
    worlds_ = std::move(worlds);
    params_ = std::move(params);
    densityPending_ = std::move(pending);
    bondWork_ = std::move(bondWork);
    livingWork_ = std::move(livingWork);
    active_.assign(worlds_.size(), 1);
    stripChoice_ = std::move(stripChoice);
    stageChoice_ = std::move(stageChoice);
    chunk_shape(worlds_, params_, chunkRows_);
    shape_.hasParticles = anyRadiates;
    for (int lp : liveParticles) shape_.hasParticles = shape_.hasParticles || lp > 0;
    parity_ = 0;
    shape_.candGrid = 16;
    // one particle block per workgroup (blocks past a world's live particles exit at once)

    std::vector<DevCounters> ctr(worlds_.size());
    for (size_t i = 0; i < worlds_.size(); ++i) {
        ctr[i].timestep = timesteps[i];
        ctr[i].numParticles = liveParticles[i];
        ctr[i].pad = 0;
    }
    if (!worlds_.empty()) {
        check_hip(hipMemcpy(dWorlds_, worlds_.data(), sizeof(DevWorld) * worlds_.size(), hipMemcpyHostToDevice), "H2D");
        check_hip(hipMemcpy(dParams_, params_.data(), sizeof(DevParams) * params_.size(), hipMemcpyHostToDevice), "H2D");
        check_hip(hipMemcpy(dCtr_, ctr.data(), sizeof(DevCounters) * ctr.size(), hipMemcpyHostToDevice), "H2D");
    }
    StepCounters sc{};
    check_hip(hipMemcpy(dSc_, &sc, sizeof(sc), hipMemcpyHostToDevice), "H2D");
    launch_fill_u32(count_[0], 0u, size_t(NB), stream_);
    launch_fill_u32(count_[1], 0u, size_t(NB), stream_);
    for (int b = 0; b < 2; ++b) {
        launch_fill_u32(start_[b], 0u, size_t(NB) + 1, stream_);
        launch_fill_u32(reinterpret_cast<uint32_t*>(moverHead_[b]), uint32_t(kNone), NCB, stream_);
        launch_fill_u32(reinterpret_cast<uint32_t*>(delta_[b]), 0u, NCB, stream_);
    }
    launch_fill_u32(reinterpret_cast<uint32_t*>(fclaim_), uint32_t(kClaimFree), size_t(NP), stream_);
    launch_fill_u32(reinterpret_cast<uint32_t*>(cclaim_), uint32_t(kClaimFree), size_t(NC), stream_);
    launch_fill_u32(reinterpret_cast<uint32_t*>(mapStatus_), 0u, 2 * size_t(map_scan_tiles(NB)), stream_);
    launch_fill_u32(mapEpochTicket_, 0u, 2, stream_);
    launch_fill_u32(reinterpret_cast<uint32_t*>(claim_), 0xffffffffu, size_t(NC) * 2, stream_);
    launch_fill_u32(reinterpret_cast<uint32_t*>(delFlag_), 0u, (size_t(NC) + 3) / 4, stream_);
    launch_fill_u32(reinterpret_cast<uint32_t*>(mark_), 0u, (size_t(NC) + 3) / 4, stream_);
    launch_fill_u32(reinterpret_cast<uint32_t*>(buf_[1].pflags), 0u, (size_t(NP) + 3) / 4, stream_);
    if (partStage_ > 0)  // no particle waits in a staging slot
        launch_fill_u32(reinterpret_cast<uint32_t*>(pbinOf_ + partStage_), 0xffffffffu, size_t(NC), stream_);
    if (splitStage_ > 0) {
        launch_fill_u32(reinterpret_cast<uint32_t*>(pbinOf_ + splitStage_), 0xffffffffu, size_t(splitStage_), stream_);
        // no split half is pending but those copied below
        launch_fill_u32(reinterpret_cast<uint32_t*>(splitStaged_), 0u, size_t(splitStage_) * sizeof(StagedParticle) / 4, stream_);
        for (const PendingCopy& pc : pendingCopies)
            if (pc.from && pc.count > 0)
                check_hip(hipMemcpyAsync(splitStaged_ + worlds_[pc.world].partOff, pc.from + pc.fromOff,
                                         sizeof(StagedParticle) * size_t(pc.count), hipMemcpyDeviceToDevice, stream_),
                          "hipMemcpy D2D");
    }
    check_hip(hipGetLastError(), "fill kernels");
    buildTables();
    launch_seed_bins(makeCtx(parity_), shape_, stream_);
    check_hip(hipGetLastError(), "seed bins");
    synchronize();
}

// Strip shape of the neighbour pass for one world, measured on the given positions.  Only
// performance depends on it.
int choose_strip(const std::vector<float2>& pos, int W, int H, int binSize, bool twoRowStages, int& rowsPerStage) {
    const int nbx = W / binSize, nby = H / binSize;
    std::vector<int> cnt(size_t(nbx) * nby, 0);
    for (const float2& p : pos) {
        int sx = host_floor_int(p.x), sy = host_floor_int(p.y);
        sx = ((sx % W) + W) % W;
        sy = ((sy % H) + H) % H;
        const int bx = std::min(sx / binSize, nbx - 1), by = std::min(sy / binSize, nby - 1);
        ++cnt[size_t(by) * nbx + bx];
    }
    // sparse rows (fewer than half a wave of cells per bin row of a strip on average) are staged
    // two at a time; a chunk where that overflows runs single rows
    auto stages = [&](int sw) {
        const double perRow = double(pos.size()) / (double(nby) * double((nbx + sw - 1) / sw));
        return twoRowStages && perRow < 32.0 && nby % 2 == 0 && nby >= 6 ? 2 : 1;
    };
    // entries of every staged row as the neighbour pass stages it (interior plus the two halo
    // columns, both wrapped around the torus), for each candidate width; a chunk whose rows exceed
    // the row capacity, or whose three consecutive rows exceed the LDS pool, takes the slower
    // global path, so the widest strip where no row does that at the start is taken
    for (int sw : {kStripMax, 32, 16, 8}) {
        bool fits = true;
        std::vector<int> rows(static_cast<size_t>(nby));
        for (int x0 = 0; x0 < nbx && fits; x0 += sw) {
            const int HX = std::min(sw, nbx - x0) + 2;
            for (int y = 0; y < nby; ++y) {
                int n = 0;
                for (int hx = 0; hx < HX; ++hx) n += cnt[size_t(y) * nbx + (x0 - 1 + hx + nbx) % nbx];
                rows[size_t(y)] = n;
            }
            for (int y = 0; y < nby && fits; ++y) {
                const int three = rows[size_t(y)] + rows[size_t((y + 1) % nby)] + rows[size_t((y + 2) % nby)];
                fits = rows[size_t(y)] <= kRowCap && three <= kRowPool;
            }
        }
        if (fits) {
            rowsPerStage = stages(sw);
            return sw;
        }
    }
    rowsPerStage = stages(4);
    return 4;
}

// Strip width (capped by the world's columns), strip count and staging; chunk_shape sets the rows.
void strip_shape(DevWorld& w, int strip, int rowsPerStage) {
    w.tileX = std::max(1, std::min(strip, w.nbx));
    w.ntx = (w.nbx + w.tileX - 1) / w.tileX;
    w.rowsPerStage = rowsPerStage;
    w.tileY = 1;
    w.nty = w.nby / rowsPerStage;
}

// Rows per chunk of the row-streaming (collision) pass, one value for all worlds of a set
// (forcedRows if nonzero).  Tall chunks stage fewer halo rows per computed row; short ones even
// out the work of the workgroups, whose cell counts vary.  Measured on MI355X, the pass ran fastest
// with the tallest chunks that still give at least 5 workgroups per resident slot on the
// 3.4M-cell soft-body world, whose bodies make the workgroups' loads unequal (2 % faster than at
// 3.5); sets too small for that take the shortest chunks.
void chunk_shape(std::vector<DevWorld>& worlds, const std::vector<DevParams>& params, int forcedRows) {
    static const int numCUs = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return 256;
        return std::max(1, n);
    }();
    // (only collision worlds run the row-streaming pass; the fluid pass is cell-parallel)
    bool collision = false;
    for (size_t i = 0; i < worlds.size(); ++i) collision = collision || params[i].motion != 0;
    const int perCU = collision ? neighbor_blocks_per_cu() : 1;
    const int64_t slots = int64_t(perCU) * numCUs;
    const double minRounds = 5.0;
    int best = 2;
    // staged rows per chunk of a world for `rows` bin rows (a chunk covers at most kChunkMax bin
    // rows: the LDS histogram of its region)
    auto stagedRows = [](const DevWorld& w, int rows) {
        return std::max(1, std::min(rows, w.nby) / w.rowsPerStage);
    };
    for (int rows = 2; rows <= kChunkMax; ++rows) {
        int64_t tiles = 0;
        for (size_t i = 0; i < worlds.size(); ++i) {
            const DevWorld& w = worlds[i];
            if (w.numCells == 0 || params[i].motion == 0) continue;
            const int ty = stagedRows(w, rows), nyS = w.nby / w.rowsPerStage;
            tiles += int64_t(w.ntx) * ((nyS + ty - 1) / ty);
        }
        if (double(tiles) >= minRounds * double(slots)) best = rows;
    }
    if (forcedRows > 0) best = forcedRows;
    for (DevWorld& w : worlds) {
        const int nyS = w.nby / w.rowsPerStage;
        w.tileY = stagedRows(w, best);
        w.nty = (nyS + w.tileY - 1) / w.tileY;
    }
}

void WorldSet::buildTables() {
    if (stream_) synchronize();  // (the tables and the live counts read below are in use until then)
    for (auto& a : tableAllocs_) (void)hipFree(a.p);
    tableAllocs_.clear();
    std::vector<int2> cells, bonds, bins, pbins, parts, living;
    const int ct = shape_.cellThreads, pt = shape_.partThreads, bt = shape_.binThreads;
    for (int wi = 0; wi < int(worlds_.size()); ++wi) {
        if (!active_[wi]) continue;
        const DevWorld& w = worlds_[wi];
        for (int s = 0; s < w.numCells; s += ct) cells.push_back(make_int2(wi, w.cellOff + s));
        if (bondWork_[wi])
            for (int s = 0; s < w.numCells; s += ct) bonds.push_back(make_int2(wi, w.cellOff + s));
        if (livingWork_[wi])
            for (int s = 0; s < w.numCells; s += ct) living.push_back(make_int2(wi, w.cellOff + s));
        for (int s = 0; s < w.nbx * w.nby; s += bt) bins.push_back(make_int2(wi, w.binOff + s));
        for (int s = 0; s < w.pnbx * w.pnby; s += bt) pbins.push_back(make_int2(wi, w.pbinOff + s));
    }
    auto up = [&](const std::vector<int2>& v) -> int2* {
        if (v.empty()) return nullptr;
        void* p = nullptr;
        check_hip(hipMalloc(&p, sizeof(int2) * v.size()), "hipMalloc");
        tableAllocs_.push_back({p, sizeof(int2) * v.size()});
        check_hip(hipMemcpy(p, v.data(), sizeof(int2) * v.size(), hipMemcpyHostToDevice), "H2D");
        return static_cast<int2*>(p);
    };
    // neighbour-pass tiles, row-major per world; workgroup b runs on XCD b % 8, so each XCD gets
    // one contiguous run of tiles and the halo rows its tiles share stay in that XCD's L2
    std::vector<int2> tiles;
    for (int wi = 0; wi < int(worlds_.size()); ++wi) {
        const DevWorld& w = worlds_[wi];
        if (w.numCells == 0 || !active_[wi] || params_[wi].motion == 0) continue;
        for (int t = 0; t < w.ntx * w.nty; ++t) tiles.push_back(make_int2(wi, t));
    }
    std::vector<int2> tileTable;
    if (!tiles.empty()) {
        const int T = int(tiles.size());
        const int chunk = (T + kXcds - 1) / kXcds;
        tileTable.assign(size_t(chunk) * kXcds, make_int2(-1, -1));
        for (int b = 0; b < int(tileTable.size()); ++b) {
            const int t = (b % kXcds) * chunk + b / kXcds;
            if (t < T) tileTable[b] = tiles[t];
        }
    }
    // particle blocks: a table sized for every active world's particle capacity, holding the
    // blocks of the live particles (the particle bin sort rewrites it every step); the grid of the
    // particle kernels follows the live particles (and the cells that can radiate)
    std::vector<DevCounters> ctr(worlds_.size());
    if (!worlds_.empty())
        check_hip(hipMemcpy(ctr.data(), dCtr_, sizeof(DevCounters) * ctr.size(), hipMemcpyDeviceToHost), "D2H");
    size_t partCapBlocks = 0;
    int64_t live = 0, radCells = 0;
    for (int wi = 0; wi < int(worlds_.size()); ++wi) {
        if (!active_[wi]) continue;
        const DevWorld& w = worlds_[wi];
        partCapBlocks += size_t((w.partCap + pt - 1) / pt);
        for (int s = 0; s < ctr[wi].numParticles; s += pt) parts.push_back(make_int2(wi, w.partOff + s));
        live += ctr[wi].numParticles;
        if (params_[wi].radiates) radCells += w.numCells;
    }
    {
        std::vector<int2> table(std::max<size_t>(partCapBlocks, 1), make_int2(0, 0));
        std::copy(parts.begin(), parts.end(), table.begin());
        partBlocks_ = up(table);
        const std::vector<int2> count(1, make_int2(int(parts.size()), 0));
        numPartBlocks_ = reinterpret_cast<int*>(up(count));
    }
    shape_.partGrid = int(std::max<int64_t>(64, std::min<int64_t>(65536, (live + pt - 1) / pt + (radCells / 8 + pt - 1) / pt)));
    cellBlocks_ = up(cells);
    bondBlocks_ = up(bonds);
    binBlocks_ = up(bins);
    pbinBlocks_ = up(pbins);
    tileBlocks_ = up(tileTable);
    livingBlocks_ = up(living);
    shape_.livingBlocks = int(living.size());
    shape_.cellBlocks = int(cells.size());
    shape_.bondBlocks = int(bonds.size());
    // the bond kernels run when an active world can have a bonded or barrier cell
    shape_.hasBondWork = !bonds.empty();
    shape_.partBlocks = int(partCapBlocks);
    shape_.binBlocks = int(bins.size());
    shape_.pbinBlocks = int(pbins.size());
    shape_.tileBlocks = int(tileTable.size());
    numCellBlocks_ = int(cells.size());
    shape_.hasFluid = shape_.hasCollision = false;
    for (const DevParams& p : params_) (p.motion == 0 ? shape_.hasFluid : shape_.hasCollision) = true;
    updateInnerFriction();
    updateRadiation();
    invalidateGraphs();
}

void WorldSet::updateRadiation() {
    bool on = false, split = false;
    for (size_t w = 0; w < params_.size(); ++w) {
        on = on || (active_[w] && params_[w].radiates);
        split = split || (active_[w] && params_[w].splits);
    }
    if (on != shape_.hasRadiation || split != shape_.hasSplitting) invalidateGraphs();
    shape_.hasRadiation = on;
    shape_.hasSplitting = split;
    if (on) shape_.hasParticles = true;  // (particles can appear)
}

// Index space of the particles: the worlds' ranges [0, caps), then the split staging slots (one per
// particle slot) with a world that can split, then the radiation staging slots (one per cell) with
// a radiating world.  Returns its size.
int stageLayout(int64_t caps, int cells, bool radiation, bool splitting, int& splitStage, int& radStage) {
    const int64_t c = std::max<int64_t>(caps, 1);
    int64_t end = c;
    splitStage = splitting ? int(end) : 0;
    if (splitting) end += c;
    radStage = radiation && end < (int64_t(1) << 31) ? int(end) : 0;
    if (radiation) end += cells;
    // (particle indices, staging slots included, are 32-bit)
    if (end >= (int64_t(1) << 31) - 1)
        throw Error(ALIEN_E_CAPACITY, "the particle slots and staging slots of the world set exceed 32-bit indices");
    return int(end);
}

// the inner-friction kernel runs when an active world with bond work has innerFriction > 0
void WorldSet::updateInnerFriction() {
    bool any = false;
    for (size_t i = 0; i < worlds_.size(); ++i)
        any = any || (active_[i] && bondWork_[i] && params_[i].innerFriction > 0.0f);
    if (any != shape_.hasInnerFriction) invalidateGraphs();
    shape_.hasInnerFriction = any;
}

void WorldSet::setBlockSize(int threads) {
    if (threads < 64 || threads > 256 || threads % 64 != 0)
        throw Error(ALIEN_E_ARG, "block size must be 64, 128, 192 or 256");
    shape_.cellThreads = threads;
    shape_.partThreads = threads;
    shape_.binThreads = threads;
    if (!worlds_.empty()) buildTables();
}

// The kernels reach a world only through the launch tables, and every offset they derive from the
// shared scan is relative to the world's own first bin, so the frozen counters of an inactive
// world change nothing for the others.
void WorldSet::setActive(int w, bool on) {
    if (active_.at(w) == uint8_t(on)) return;
    active_[w] = on;
    worlds_[w].active = on ? 1 : 0;
    check_hip(hipMemcpy(dWorlds_ + w, &worlds_[w], sizeof(DevWorld), hipMemcpyHostToDevice), "H2D");
    if (on) {
        synchronize();
        anchorAges(w);
    }
    buildTables();
    invalidateGraphs();
}

void WorldSet::setStripCap(int cap) {
    if (cap != 8 && cap != 16 && cap != 32 && cap != kStripMax)
        throw Error(ALIEN_E_ARG, "strip cap must be 8, 16, 32 or 64");
    stripCap_ = cap;
    if (worlds_.empty()) return;
    for (size_t i = 0; i < worlds_.size(); ++i)
        strip_shape(worlds_[i], std::min(stripChoice_[i], cap), stageChoice_[i]);
    chunk_shape(worlds_, params_, chunkRows_);
    check_hip(hipMemcpy(dWorlds_, worlds_.data(), sizeof(DevWorld) * worlds_.size(), hipMemcpyHostToDevice), "H2D");
    buildTables();
}

void WorldSet::setChunkRows(int rows) {
    if (rows != 0 && (rows < 2 || rows > kChunkMax)) throw Error(ALIEN_E_ARG, "chunk rows must be 0 or 2..8");
    chunkRows_ = rows;
    if (worlds_.empty()) return;
    chunk_shape(worlds_, params_, chunkRows_);
    check_hip(hipMemcpy(dWorlds_, worlds_.data(), sizeof(DevWorld) * worlds_.size(), hipMemcpyHostToDevice), "H2D");
    buildTables();
}

void WorldSet::setUseGraph(bool on) {
    useGraph_ = on;
    invalidateGraphs();
}

void WorldSet::pathCounts(uint64_t& fallbackChunks, uint64_t& directCells) {
    fallbackChunks = directCells = 0;
    if (!dSc_) return;
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    StepCounters sc{};
    check_hip(hipMemcpy(&sc, dSc_, sizeof(sc), hipMemcpyDeviceToHost), "D2H");
    fallbackChunks = sc.fallbackChunks;
    directCells = sc.directCells;
}

void WorldSet::setKernelTiming(bool on) {
    timing_ = on;
    if (on && !recorder_) recorder_ = std::make_unique<EventRecorder>(stream_);
}

std::map<std::string, WorldSet::KernelTime> WorldSet::takeKernelTimes() {
    std::map<std::string, KernelTime> r;
    r.swap(kernelTimes_);
    return r;
}

Ctx WorldSet::makeCtx(int parity) const {
    Ctx c{};
    c.numWorlds = int(worlds_.size());
    c.cellCap = cellCap_;
    c.partCap = partCap_;
    c.numBins = numBins_;
    c.numCellBins = numCellBins_;
    c.start = start_[parity];
    c.startPrev = start_[parity ^ 1];
    c.moverHead = moverHead_[parity];
    c.moverHeadNext = moverHead_[parity ^ 1];
    c.moverNext = moverNext_;
    c.delta = delta_[parity];
    c.deltaNext = delta_[parity ^ 1];
    c.worlds = dWorlds_;
    c.params = dParams_;
    c.ctr = dCtr_;
    c.sc = dSc_;
    c.A = buf_[parity];
    c.B = buf_[parity ^ 1];
    c.energy = energy_;
    c.stiff = stiff_;
    c.cid = cid_;
    c.ageBase = ageBase_;
    c.partStage = partStage_;
    c.staged = staged_;
    c.splitStage = splitStage_;
    c.splitStaged = splitStaged_;
    c.timestepAdvanced = shape_.hasBondWork ? 0 : 1;
    c.living = living_[parity];
    c.livingNext = living_[parity ^ 1];
    c.activation = activation_;
    c.creature = creature_;
    c.stretched = stretched_;
    c.livingBlocks = livingBlocks_;
    c.binOf = binOf_;
    c.pbinOf = pbinOf_;
    c.count = count_[parity];
    c.countNext = count_[parity ^ 1];
    c.mapStatus = mapStatus_;
    c.mapEpoch = mapEpochTicket_;
    c.mapTicket = mapEpochTicket_ + 1;
    c.sorted = sorted_;
    c.mpos = mpos_;
    c.mvel = mvel_;
    c.mdens = mdens_;
    c.mflags = mflags_;
    c.newidx = newidx_;
    c.psorted = psorted_;
    c.stagedMark = stagedMark_;
    c.own = own_;
    c.shift = shift_;
    c.bstiff = bstiff_;
    c.s2 = s2_;
    c.out1 = out1_;
    c.out2 = out2_;
    c.valid = valid_;
    c.delFlag = delFlag_;
    c.mark = mark_;
    c.claim = claim_;
    c.cand = cand_;
    c.candCap = candCap_;
    c.want = want_;
    c.fclaim = fclaim_;
    c.absorbCell = absorbCell_;
    c.cclaim = cclaim_;
    c.cellBlocks = cellBlocks_;
    c.bondBlocks = bondBlocks_;
    c.binBlocks = binBlocks_;
    c.pbinBlocks = pbinBlocks_;
    c.partBlocks = partBlocks_;
    c.tileBlocks = tileBlocks_;
    c.numPartBlocks = numPartBlocks_;
    c.partBlockParticles = shape_.partThreads;
    c.numCellBlocks = numCellBlocks_;
    c.cellBlockCells = shape_.cellThreads;
    return c;
}

int WorldSet::numParticles(int w) {
    DevCounters c{};
    check_hip(hipMemcpy(&c, dCtr_ + w, sizeof(c), hipMemcpyDeviceToHost), "D2H");
    return c.numParticles;
}

int WorldSet::liveParticles(int w) {
    synchronize();
    const int np = numParticles(w);
    std::vector<uint8_t> f(static_cast<size_t>(std::max(np, 0)));
    if (np > 0)
        check_hip(hipMemcpy(f.data(), buf_[parity_].pflags + worlds_.at(w).partOff, size_t(np), hipMemcpyDeviceToHost),
                  "D2H");
    int live = 0;
    for (uint8_t x : f) live += (x & kParticleDead) ? 0 : 1;
    if (splitStage_ > 0 && np > 0) {  // and the split halves waiting for the next map
        std::vector<StagedParticle> rec(static_cast<size_t>(np));
        check_hip(hipMemcpy(rec.data(), splitStaged_ + worlds_.at(w).partOff, sizeof(StagedParticle) * size_t(np),
                            hipMemcpyDeviceToHost),
                  "D2H");
        for (const StagedParticle& r : rec) live += (r.color & kStagedPending) ? 1 : 0;
    }
    return live;
}

uint64_t WorldSet::timestep(int w) {
    DevCounters c{};
    check_hip(hipMemcpy(&c, dCtr_ + w, sizeof(c), hipMemcpyDeviceToHost), "D2H");
    return c.timestep;
}

void WorldSet::setTimestep(int w, uint64_t t) {
    synchronize();
    DevCounters c{};
    check_hip(hipMemcpy(&c, dCtr_ + w, sizeof(c), hipMemcpyDeviceToHost), "D2H");
    c.timestep = t;
    check_hip(hipMemcpy(dCtr_ + w, &c, sizeof(c), hipMemcpyHostToDevice), "H2D");
    anchorAges(w);
}

// The device derives a cell's age from the world's timestep (ageBase + timestep - ageTs); after
// the timestep is set, or while the world was inactive (frozen, its ages too, though the step
// advances every world's timestep), the anchor moves so that the ages stay base + steps simulated.
void WorldSet::anchorAges(int w) {
    DevCounters c{};
    check_hip(hipMemcpy(&c, dCtr_ + w, sizeof(c), hipMemcpyDeviceToHost), "D2H");
    worlds_.at(w).ageTs = c.timestep - hosts_.at(w).stepsSinceSet;
    check_hip(hipMemcpy(dWorlds_ + w, &worlds_[w], sizeof(DevWorld), hipMemcpyHostToDevice), "H2D");
}

void WorldSet::setParams(int w, const alien_params& p) {
    validate_world(hosts_.at(w).desc, p);
    if (bin_size_for(p) != bin_size_for(hosts_.at(w).params))
        throw Error(ALIEN_E_STATE, "the bin grid changes with these parameters: build a fresh set around them");
    if (!livingWork_.at(w) && needs_fresh_set(hosts_.at(w).params, p))
        throw Error(ALIEN_E_STATE,
                    "cellMinEnergy changes or cell radiation starts: build a fresh set around the new parameters");
    const bool wasFluid = params_.at(w).motion == 0;
    hosts_.at(w).params = p;
    params_.at(w) = to_dev_params(p);
    check_hip(hipMemcpy(dParams_ + w, &params_[w], sizeof(DevParams), hipMemcpyHostToDevice), "H2D");
    densityPending_.at(w) = 1;  // prepareForSimulationParametersChanges, SimulationCudaFacade.cu:667-669
    // the set of motion modes selects which neighbour-pass variants the step launches
    const bool fluid = shape_.hasFluid, collision = shape_.hasCollision;
    shape_.hasFluid = shape_.hasCollision = false;
    for (const DevParams& q : params_) (q.motion == 0 ? shape_.hasFluid : shape_.hasCollision) = true;
    if (wasFluid != (params_[w].motion == 0)) {
        // only collision worlds run the row-streaming pass: its chunks and tiles follow the switch
        chunk_shape(worlds_, params_, chunkRows_);
        check_hip(hipMemcpy(dWorlds_, worlds_.data(), sizeof(DevWorld) * worlds_.size(), hipMemcpyHostToDevice), "H2D");
        buildTables();
        invalidateGraphs();
    } else if (fluid != shape_.hasFluid || collision != shape_.hasCollision) {
        invalidateGraphs();
    }
    updateInnerFriction();
    updateRadiation();  // (the room for emitted particles is made before the next step)
}

// A world whose living states are not tracked on the device stays trivially Ready only while no
// cell can fall below cellMinEnergy: a change of it needs a fresh set (built from the world's
// current state, which decides anew whether to track).
bool needs_fresh_set(const alien_params& before, const alien_params& after) {
    for (int c = 0; c < kColors; ++c)
        if (before.cellMinEnergy[c] != after.cellMinEnergy[c]) return true;
    return radiates(after) && !radiates(before);
}

int64_t particle_room(int64_t cells, int64_t particles, const alien_params& p) {
    constexpr int64_t kRoomSteps = 8;  // twice the steps between the room checks of a radiating set
    return (particles + (radiates(p) ? kRoomSteps * cells : 0)) * (splits(p) ? 2 : 1);
}

// RadiationProcessor::splitting acts on a particle above a finite particleSplitEnergy of its colour
bool splits(const alien_params& p) {
    for (int c = 0; c < kColors; ++c)
        if (!std::isinf(p.particleSplitEnergy[c]) || p.particleSplitEnergy[c] < 0.0f) return true;
    return false;
}

// CellProcessor::radiation takes energy from cells only when a draw below radiationProb can meet
// a positive factor
bool radiates(const alien_params& p) {
    if (!(p.radiationProb > 0.0f)) return false;
    for (int c = 0; c < kColors; ++c)
        if (p.highRadiationFactor[c] > 0.0f || p.radiationCellAgeStrength[c] > 0.0f) return true;
    return false;
}

// Reads the last step's candidate-list size: reports overflow and grows the grid of the fusion
// kernels when the list outgrew it (the graphs are captured again on the next step).
void WorldSet::checkErrors() {
    StepCounters sc{};
    check_hip(hipMemcpy(&sc, dSc_, sizeof(sc), hipMemcpyDeviceToHost), "D2H");
    if (sc.errors & 1) {
        sc.errors = 0;
        check_hip(hipMemcpy(dSc_, &sc, sizeof(sc), hipMemcpyHostToDevice), "H2D");
        throw Error(ALIEN_E_CAPACITY, "fusion candidate list overflow");
    }
    if (sc.errors & 2) {
        sc.errors = 0;
        check_hip(hipMemcpy(dSc_, &sc, sizeof(sc), hipMemcpyHostToDevice), "H2D");
        throw Error(ALIEN_E_STATE, "map build: a bin tile's offset was not published in time");
    }
    if (sc.errors & 4) {
        sc.errors = 0;
        check_hip(hipMemcpy(dSc_, &sc, sizeof(sc), hipMemcpyHostToDevice), "H2D");
        throw Error(ALIEN_E_CAPACITY, "a world's energy particles exceeded its particle capacity");
    }
    const int wantCand = std::min(2048, (std::min(sc.candidates, candCap_) + 255) / 256);
    if (wantCand > 2 * shape_.candGrid) {
        shape_.candGrid = std::max(shape_.candGrid, wantCand);
        invalidateGraphs();
    }
}

void WorldSet::step(uint64_t n) {
    if (worlds_.empty() || n == 0) return;
    // density reset requested by a parameter change, applied before the next step
    for (int w = 0; w < numWorlds(); ++w) {
        if (!densityPending_[w]) continue;
        const DevWorld& dw = worlds_[w];
        launch_fill_u32(reinterpret_cast<uint32_t*>(buf_[parity_].dens + dw.cellOff), 0x3f800000u, size_t(dw.numCells),
                        stream_);
        densityPending_[w] = 0;
    }
    // the captured graph of `count` steps from parity p; all four are built together on first
    // use, so that no later call pays for a capture
    auto build = [&](hipGraphExec_t& g, int p, uint64_t count) {
        if (g) return;
        hipGraph_t graph = nullptr;
        check_hip(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        for (uint64_t k = 0; k < count; ++k) launch_step(makeCtx(p ^ int(k & 1)), shape_, stream_, nullptr, &side_);
        check_hip(hipStreamEndCapture(stream_, &graph), "hipStreamEndCapture");
        check_hip(hipGraphInstantiate(&g, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
        (void)hipGraphDestroy(graph);
    };
    auto graphOf = [&](uint64_t count) -> hipGraphExec_t {
        for (int p = 0; p < 2; ++p) {
            build(graph_[p], p, 1);
            build(graphMulti_[p], p, kStepsPerGraph);
        }
        return (count == 1 ? graph_ : graphMulti_)[parity_];
    };
    uint64_t roomUntil = 0;  // with cell radiation: steps of this call covered by the last room check
    for (uint64_t s = 0; s < n;) {
        if ((shape_.hasRadiation || shape_.hasSplitting) && s >= roomUntil) {
            // (with splitting the live particles can double every step: room is made per step)
            const uint64_t k = std::min<uint64_t>(n - s, shape_.hasSplitting ? 1 : kRadiationSteps);
            ensureParticleRoom(k);  // (a growth rebuilds the tables; the graphs follow)
            roomUntil = s + k;
        }
        // while particles live, every 32 steps their count is read without waiting for the device:
        // once none is left (seen a check later), the particle kernels go
        if (shape_.hasParticles && !shape_.hasRadiation && s > 0 && s % 32 == 0) checkParticlesAsync();
        uint64_t done = 1;
        if (timing_) {
            launch_step(makeCtx(parity_), shape_, stream_, recorder_.get());
            check_hip(hipGetLastError(), "kernel launch");
        } else if (useGraph_) {
            // several steps per graph while they fit before the end and the next particle check
            const uint64_t toCheck =
                (shape_.hasRadiation || shape_.hasSplitting) ? roomUntil - s
                                                             : (shape_.hasParticles ? 32 - s % 32 : n - s);
            if (n - s >= kStepsPerGraph && toCheck >= kStepsPerGraph) done = kStepsPerGraph;
            check_hip(hipGraphLaunch(graphOf(done), stream_), "hipGraphLaunch");
        } else {
            launch_step(makeCtx(parity_), shape_, stream_, nullptr, &side_);
            check_hip(hipGetLastError(), "kernel launch");
        }
        if (done & 1) parity_ ^= 1;
        s += done;
    }
    for (size_t i = 0; i < hosts_.size(); ++i)
        if (active_[i]) hosts_[i].stepsSinceSet += n;  // (an inactive world is frozen)
    synchronize();
    if (timing_) recorder_->collect(kernelTimes_);
    checkErrors();
    if (shape_.hasParticles) dropParticlesIfNone();
}

void WorldSet::checkParticlesAsync() {
    if (livePending_) {
        const hipError_t q = hipEventQuery(liveReady_);
        if (q == hipSuccess) {
            livePending_ = false;
            if (*liveHost_ == 0 && !shape_.hasRadiation) {
                shape_.hasParticles = false;
                invalidateGraphs();
                return;
            }
        } else if (q == hipErrorNotReady) {
            (void)hipGetLastError();  // (not ready yet: not an error of the launches)
        } else {
            check_hip(q, "hipEventQuery");
        }
    }
    if (!livePending_) {
        check_hip(hipMemcpyAsync(liveHost_, &dSc_->liveParticles, sizeof(int), hipMemcpyDeviceToHost, stream_),
                  "hipMemcpyAsync");
        check_hip(hipEventRecord(liveReady_, stream_), "hipEventRecord");
        livePending_ = true;
    }
}

void WorldSet::dropParticlesIfNone() {
    synchronize();
    livePending_ = false;  // (a read in flight has landed; the count below is newer)
    StepCounters sc{};
    check_hip(hipMemcpy(&sc, dSc_, sizeof(sc), hipMemcpyDeviceToHost), "D2H");
    if (sc.liveParticles == 0 && !shape_.hasRadiation) {
        shape_.hasParticles = false;
        invalidateGraphs();
    }
}

void WorldSet::ensureParticleRoom(uint64_t steps) {
    static_assert(2 * kRadiationSteps == 8, "particle_room builds radiating worlds with room for 8 steps");
    if (!shape_.hasRadiation && !shape_.hasSplitting) return;
    synchronize();
    std::vector<DevCounters> ctr(worlds_.size());
    check_hip(hipMemcpy(ctr.data(), dCtr_, sizeof(DevCounters) * ctr.size(), hipMemcpyDeviceToHost), "D2H");
    bool grow = (shape_.hasRadiation && partStage_ == 0) || (shape_.hasSplitting && splitStage_ == 0);
    std::vector<int64_t> caps(worlds_.size());
    for (size_t w = 0; w < worlds_.size(); ++w) {
        caps[w] = worlds_[w].partCap;
        const bool rad = params_[w].radiates != 0, spl = params_[w].splits != 0;
        if (!active_[w] || !(rad || spl)) continue;
        // (steps is 1 with splitting)
        int64_t need = int64_t(ctr[w].numParticles) + (rad ? int64_t(steps) * worlds_[w].numCells : 0);
        if (spl) need *= 2;
        if (need > caps[w]) {
            caps[w] = 2 * need;
            grow = true;
        }
    }
    if (grow) growParticles(caps);
}

void WorldSet::release(void* p) {
    for (size_t k = 0; k < allocs_.size(); ++k)
        if (allocs_[k].p == p) {
            (void)hipFree(p);
            deviceBytes_ -= allocs_[k].bytes;
            allocs_.erase(allocs_.begin() + std::ptrdiff_t(k));
            return;
        }
}

void WorldSet::growParticles(const std::vector<int64_t>& caps) {
    synchronize();
    int64_t total = 0;
    for (int64_t c : caps) total += c;
    if (total > kMaxSetParticles)
        throw Error(ALIEN_E_CAPACITY, "the particle room of the radiating worlds exceeds the particles of one world set");
    std::vector<DevCounters> ctr(worlds_.size());
    check_hip(hipMemcpy(ctr.data(), dCtr_, sizeof(DevCounters) * ctr.size(), hipMemcpyDeviceToHost), "D2H");
    bool anyRadiates = false, anySplits = false;
    for (const DevParams& q : params_) {
        anyRadiates = anyRadiates || q.radiates;
        anySplits = anySplits || q.splits;
    }
    int splitStage = 0, radStage = 0;
    const int NP = stageLayout(total, cellCap_, anyRadiates, anySplits, splitStage, radStage);
    // the particle arrays of both buffers and the per-particle step arrays, reallocated; the
    // current state (and its next-step bins) moves to the new world offsets
    HotBuffers nb[2];
    for (int b = 0; b < 2; ++b) {
        nb[b] = buf_[b];
        nb[b].ppos = static_cast<float2*>(alloc(sizeof(float2) * NP));
        nb[b].pvel = static_cast<float2*>(alloc(sizeof(float2) * NP));
        nb[b].pen = static_cast<float*>(alloc(sizeof(float) * NP));
        nb[b].plast = static_cast<int*>(alloc(sizeof(int) * NP));
        nb[b].pslot = static_cast<int*>(alloc(sizeof(int) * NP));
        nb[b].pflags = static_cast<uint8_t*>(alloc(NP));
        nb[b].pid = static_cast<unsigned long long*>(alloc(sizeof(unsigned long long) * NP));
    }
    int* pbinOf = static_cast<int*>(alloc(sizeof(int) * NP));
    // no particle waits in a new staging slot, but for the split particles of the last step, which
    // move with the particles they split from
    launch_fill_u32(reinterpret_cast<uint32_t*>(pbinOf), 0xffffffffu, size_t(NP), stream_);
    StagedParticle* splitStaged =
        splitStage > 0 ? static_cast<StagedParticle*>(alloc(sizeof(StagedParticle) * size_t(splitStage))) : nullptr;
    if (splitStaged)
        launch_fill_u32(reinterpret_cast<uint32_t*>(splitStaged), 0u, size_t(splitStage) * sizeof(StagedParticle) / 4, stream_);
    const HotBuffers& S = buf_[parity_];
    HotBuffers& D = nb[parity_];
    int64_t off = 0;
    auto d2d = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) check_hip(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_), "hipMemcpy D2D");
    };
    for (size_t w = 0; w < worlds_.size(); ++w) {
        DevWorld& dw = worlds_[w];
        const size_t np = size_t(ctr[w].numParticles);
        const int from = dw.partOff, to = int(off);
        d2d(D.ppos + to, S.ppos + from, sizeof(float2) * np);
        d2d(D.pvel + to, S.pvel + from, sizeof(float2) * np);
        d2d(D.pen + to, S.pen + from, sizeof(float) * np);
        d2d(D.plast + to, S.plast + from, sizeof(int) * np);
        d2d(D.pslot + to, S.pslot + from, sizeof(int) * np);
        d2d(D.pflags + to, S.pflags + from, np);
        d2d(D.pid + to, S.pid + from, sizeof(unsigned long long) * np);
        d2d(pbinOf + to, pbinOf_ + from, sizeof(int) * np);
        if (splitStage_ > 0 && splitStage > 0) {
            d2d(pbinOf + splitStage + to, pbinOf_ + splitStage_ + from, sizeof(int) * np);
            d2d(splitStaged + to, splitStaged_ + from, sizeof(StagedParticle) * np);
        }
        dw.partOff = to;
        dw.partCap = int(caps[w]);
        off += caps[w];
    }
    check_hip(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    for (int b = 0; b < 2; ++b) {
        for (void* p : {static_cast<void*>(buf_[b].ppos), static_cast<void*>(buf_[b].pvel), static_cast<void*>(buf_[b].pen),
                        static_cast<void*>(buf_[b].plast), static_cast<void*>(buf_[b].pslot),
                        static_cast<void*>(buf_[b].pflags), static_cast<void*>(buf_[b].pid)})
            release(p);
        buf_[b] = nb[b];
    }
    for (void* p : {static_cast<void*>(pbinOf_), static_cast<void*>(psorted_), static_cast<void*>(want_),
                    static_cast<void*>(fclaim_), static_cast<void*>(absorbCell_), static_cast<void*>(splitStaged_)})
        if (p) release(p);
    splitStaged_ = splitStaged;
    pbinOf_ = pbinOf;
    psorted_ = static_cast<int*>(alloc(sizeof(int) * NP));
    want_ = static_cast<int*>(alloc(sizeof(int) * NP));
    fclaim_ = static_cast<int*>(alloc(sizeof(int) * NP));
    absorbCell_ = static_cast<int*>(alloc(sizeof(int) * NP));
    partCap_ = NP;
    partStage_ = radStage;
    splitStage_ = splitStage;
    if (partStage_ > 0 && !staged_) staged_ = static_cast<StagedParticle*>(alloc(sizeof(StagedParticle) * size_t(cellCap_)));
    launch_fill_u32(reinterpret_cast<uint32_t*>(fclaim_), uint32_t(kClaimFree), size_t(NP), stream_);
    launch_fill_u32(reinterpret_cast<uint32_t*>(buf_[parity_ ^ 1].pflags), 0u, (size_t(NP) + 3) / 4, stream_);
    check_hip(hipGetLastError(), "fill kernels");
    check_hip(hipMemcpy(dWorlds_, worlds_.data(), sizeof(DevWorld) * worlds_.size(), hipMemcpyHostToDevice), "H2D");
    buildTables();  // (particle blocks of the new capacities; the graphs are captured again)
    synchronize();
}

void WorldSet::extract(int wi, WorldImage& img) {
    synchronize();
    const DevWorld& w = worlds_.at(wi);
    DevCounters ctr{};
    check_hip(hipMemcpy(&ctr, dCtr_ + wi, sizeof(ctr), hipMemcpyDeviceToHost), "D2H");
    const int n = w.numCells, np = ctr.numParticles;
    img.host = hosts_.at(wi);
    img.timestep = ctr.timestep;
    img.numParticles = np;
    img.densityResetPending = densityPending_.at(wi) != 0;
    const HotBuffers& S = buf_[parity_];
    auto d2h = [&](void* dst, const void* src, size_t bytes) {
        if (bytes) check_hip(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost), "D2H");
    };
    img.pos.resize(n);
    img.vel.resize(n);
    img.dens.resize(n);
    img.flags.resize(n);
    img.slot.resize(n);
    img.btgt.resize(size_t(kMaxBonds) * n);
    img.brest.resize(size_t(kMaxBonds) * n);
    img.bang.resize(size_t(kMaxBonds) * n);
    img.energy.resize(n);
    img.stiff.resize(n);
    d2h(img.pos.data(), S.pos + w.cellOff, sizeof(float2) * n);
    d2h(img.vel.data(), S.vel + w.cellOff, sizeof(float2) * n);
    d2h(img.dens.data(), S.dens + w.cellOff, sizeof(float) * n);
    d2h(img.flags.data(), S.flags + w.cellOff, sizeof(uint32_t) * n);
    d2h(img.slot.data(), S.slot + w.cellOff, sizeof(int) * n);
    for (int b = 0; b < kMaxBonds; ++b) {
        d2h(img.btgt.data() + size_t(b) * n, S.btgt + size_t(b) * cellCap_ + w.cellOff, sizeof(int) * n);
        d2h(img.brest.data() + size_t(b) * n, S.brest + size_t(b) * cellCap_ + w.cellOff, sizeof(float) * n);
        d2h(img.bang.data() + size_t(b) * n, S.bang + size_t(b) * cellCap_ + w.cellOff, sizeof(float) * n);
    }
    d2h(img.energy.data(), energy_ + w.cellOff, sizeof(float) * n);
    d2h(img.stiff.data(), stiff_ + w.cellOff, sizeof(float) * n);
    // the current age, living state and activation count (CellProcessor.cuh:561-676): aging adds
    // one per step to every non-barrier cell; a world with livingWork has its states and counts on
    // the device, any other world is Ready without counts throughout (its barrier cells too)
    HostWorld& h = img.host;
    img.livingWork = livingWork_.at(wi) != 0;
    if (img.livingWork) {
        d2h(h.livingState.data(), living_[parity_] + w.cellOff, size_t(n));
        d2h(h.activationTime.data(), activation_ + w.cellOff, sizeof(uint32_t) * n);
    }
    if (h.stepsSinceSet > 0)
        for (int i = 0; i < n; ++i) {
            const int sl = img.slot[i];
            if (flags_barrier(img.flags[i])) {
                h.livingState[sl] = 0;
                continue;
            }
            h.age[sl] = uint32_t(uint64_t(h.age[sl]) + h.stepsSinceSet);
        }
    h.stepsSinceSet = 0;
    img.ppos.resize(np);
    img.pvel.resize(np);
    img.pen.resize(np);
    img.plast.resize(np);
    img.pslot.resize(np);
    img.pflags.resize(np);
    img.pid.resize(np);
    d2h(img.ppos.data(), S.ppos + w.partOff, sizeof(float2) * np);
    d2h(img.pvel.data(), S.pvel + w.partOff, sizeof(float2) * np);
    d2h(img.pen.data(), S.pen + w.partOff, sizeof(float) * np);
    d2h(img.plast.data(), S.plast + w.partOff, sizeof(int) * np);
    d2h(img.pslot.data(), S.pslot + w.partOff, sizeof(int) * np);
    d2h(img.pflags.data(), S.pflags + w.partOff, np);
    d2h(img.pid.data(), S.pid + w.partOff, sizeof(uint64_t) * np);
    // the halves split off at the end of the last step, which enter the map at the next step,
    // follow the particles that lived (the order the next map gives them)
    std::vector<StagedParticle> halves;
    if (splitStage_ > 0 && np > 0) {
        std::vector<StagedParticle> rec(static_cast<size_t>(np));
        d2h(rec.data(), splitStaged_ + w.partOff, sizeof(StagedParticle) * size_t(np));
        for (const StagedParticle& r : rec)
            if (r.color & kStagedPending) halves.push_back(r);
    }
    // particles killed in the last step are still in the live range until the next sort
    int keep = 0;
    for (int i = 0; i < np; ++i) {
        if (img.pflags[i] & kParticleDead) continue;
        img.ppos[keep] = img.ppos[i];
        img.pvel[keep] = img.pvel[i];
        img.pen[keep] = img.pen[i];
        img.plast[keep] = img.plast[i];
        img.pslot[keep] = img.pslot[i];
        img.pflags[keep] = img.pflags[i];
        img.pid[keep] = img.pid[i];
        ++keep;
    }
    img.ppos.resize(keep);
    img.pvel.resize(keep);
    img.pen.resize(keep);
    img.plast.resize(keep);
    img.pslot.resize(keep);
    img.pflags.resize(keep);
    img.pid.resize(keep);
    for (const StagedParticle& r : halves) {
        img.ppos.push_back(r.pos);
        img.pvel.push_back(r.vel);
        img.pen.push_back(r.energy);
        img.plast.push_back(kNone);
        img.pslot.push_back(kNone);
        img.pflags.push_back(uint8_t(r.color & kParticleColorMask));
        img.pid.push_back(r.id);
    }
    img.numParticles = int(img.ppos.size());
}

alien_totals WorldSet::totals(int wi) {
    WorldImage img;
    extract(wi, img);
    alien_totals t{};
    for (int i = 0; i < img.host.numCells; ++i) {
        t.cellEnergy += double(img.energy[img.slot[i]]);
        t.momentum[0] += double(img.vel[i].x);
        t.momentum[1] += double(img.vel[i].y);
        t.kinetic += 0.5 * (double(img.vel[i].x) * double(img.vel[i].x) + double(img.vel[i].y) * double(img.vel[i].y));
    }
    for (int i = 0; i < img.numParticles; ++i) t.particleEnergy += double(img.pen[i]);
    t.numCells = uint64_t(img.host.numCells);
    t.numParticles = uint64_t(img.numParticles);
    return t;
}

alien_statistics WorldSet::statistics(int wi) {
    const DevWorld& w = worlds_.at(wi);
    alien_statistics st{};
    st.timestep = timestep(wi);
    int cellBlocks = 0;
    const int blocks = statistics_blocks(w.numCells, w.partCap, cellBlocks);
    if (blocks == 0) return st;
    synchronize();
    int* cnt = nullptr;
    double* en = nullptr;
    check_hip(hipMalloc(&cnt, sizeof(int) * size_t(blocks) * kColors), "hipMalloc");
    if (hipMalloc(&en, sizeof(double) * size_t(blocks) * kColors) != hipSuccess) {
        (void)hipFree(cnt);
        throw Error(ALIEN_E_OOM, "hipMalloc failed (statistics)");
    }
    std::vector<int> hc(size_t(blocks) * kColors);
    std::vector<double> he(size_t(blocks) * kColors);
    try {
        launch_statistics(makeCtx(parity_), wi, cellBlocks, blocks, cnt, en, stream_);
        check_hip(hipGetLastError(), "statistics");
        synchronize();
        check_hip(hipMemcpy(hc.data(), cnt, sizeof(int) * hc.size(), hipMemcpyDeviceToHost), "D2H");
        check_hip(hipMemcpy(he.data(), en, sizeof(double) * he.size(), hipMemcpyDeviceToHost), "D2H");
    } catch (...) {
        (void)hipFree(cnt);
        (void)hipFree(en);
        throw;
    }
    (void)hipFree(cnt);
    (void)hipFree(en);
    for (int b = 0; b < blocks; ++b)
        for (int k = 0; k < kColors; ++k) {
            (b < cellBlocks ? st.numCells : st.numParticles)[k] += hc[size_t(b) * kColors + k];
            st.totalEnergy[k] += he[size_t(b) * kColors + k];
        }
    return st;
}

void WorldSet::statisticsNoAlloc(int wi, alien_statistics* out) {
    const DevWorld& w = worlds_.at(wi);
    int cellBlocks = 0;
    const int blocks = statistics_blocks(w.numCells, w.partCap, cellBlocks);
    if (blocks > kStatScratchBlocks)
        throw Error(ALIEN_E_CAPACITY, "world too large for the allocation-free statistics: " + std::to_string(blocks) +
                                          " blocks of 256 cells or particle slots, at most " +
                                          std::to_string(kStatScratchBlocks));
    if (blocks == 0) {
        synchronize();
        *out = alien_statistics{};
    } else {
        void* base = nullptr;
        check_hip(statistics_scratch(&base), "hipGetSymbolAddress (statistics scratch)");
        check_hip(launch_statistics_noalloc(makeCtx(parity_), wi, cellBlocks, blocks, base, out, stream_),
                  "statistics");
    }
    out->timestep = timestep(wi);  // (the launch returned after the stream drained)
}

void WorldSet::cellMap(int wi, std::vector<int>& bins, std::vector<int>& slots) {
    synchronize();
    const Ctx c = makeCtx(parity_);
    launch_debug_map(c, shape_, stream_);  // (which seeds the bins again)
    check_hip(hipGetLastError(), "sort");
    synchronize();
    const DevWorld& w = worlds_.at(wi);
    const int n = w.numCells;
    std::vector<int> sorted(n), binOfAll(cellCap_), slotAll(cellCap_);
    if (n) {
        check_hip(hipMemcpy(sorted.data(), sorted_ + w.cellOff, sizeof(int) * n, hipMemcpyDeviceToHost), "D2H");
        check_hip(hipMemcpy(binOfAll.data(), binOf_, sizeof(int) * cellCap_, hipMemcpyDeviceToHost), "D2H");
        check_hip(hipMemcpy(slotAll.data(), buf_[parity_].slot, sizeof(int) * cellCap_, hipMemcpyDeviceToHost), "D2H");
    }
    bins.resize(n);
    slots.resize(n);
    for (int k = 0; k < n; ++k) {
        bins[k] = int(uint32_t(binOfAll[sorted[k]]) & ~kMoverBit) - w.binOff;
        slots[k] = slotAll[sorted[k]];
    }
}

void WorldSet::neighbors(int wi, std::vector<std::vector<int>>& bySlot) {
    synchronize();
    const Ctx c = makeCtx(parity_);
    launch_debug_map(c, shape_, stream_);  // (which seeds the bins again)
    int* counts = nullptr;
    int* out = nullptr;
    uint64_t* offs = nullptr;
    check_hip(hipMalloc(&counts, sizeof(int) * cellCap_), "hipMalloc");
    launch_debug_neighbors(c, shape_, nullptr, nullptr, counts, stream_);
    check_hip(hipGetLastError(), "debug neighbours");
    synchronize();
    std::vector<int> cnt(cellCap_);
    check_hip(hipMemcpy(cnt.data(), counts, sizeof(int) * cellCap_, hipMemcpyDeviceToHost), "D2H");
    std::vector<uint64_t> off(size_t(cellCap_) + 1, 0);
    for (int i = 0; i < cellCap_; ++i) off[i + 1] = off[i] + uint64_t(std::max(cnt[i], 0));
    check_hip(hipMalloc(&offs, sizeof(uint64_t) * off.size()), "hipMalloc");
    check_hip(hipMalloc(&out, sizeof(int) * std::max<uint64_t>(off.back(), 1)), "hipMalloc");
    check_hip(hipMemcpy(offs, off.data(), sizeof(uint64_t) * off.size(), hipMemcpyHostToDevice), "H2D");
    launch_debug_neighbors(c, shape_, offs, out, nullptr, stream_);
    check_hip(hipGetLastError(), "debug neighbours");
    synchronize();
    std::vector<int> ids(off.back());
    if (!ids.empty())
        check_hip(hipMemcpy(ids.data(), out, sizeof(int) * ids.size(), hipMemcpyDeviceToHost), "D2H");
    std::vector<int> slotAll(cellCap_);
    check_hip(hipMemcpy(slotAll.data(), buf_[parity_].slot, sizeof(int) * cellCap_, hipMemcpyDeviceToHost), "D2H");
    (void)hipFree(counts);
    (void)hipFree(out);
    (void)hipFree(offs);
    const DevWorld& w = worlds_.at(wi);
    bySlot.assign(w.numCells, {});
    for (int i = w.cellOff; i < w.cellOff + w.numCells; ++i) {
        auto& v = bySlot[slotAll[i]];
        for (uint64_t q = off[i]; q < off[i + 1]; ++q) v.push_back(slotAll[ids[q]]);
    }
}

}  // namespace alien
