// capi.cpp — the C ABI declared in include/alien_hip.h.
//
// Nothing throws across this boundary: every entry point converts alien::Error (and any other
// exception) into a negative code plus a thread-local message, the way ALIEN's
// CHECK_FOR_CUDA_ERROR / checkAndThrowError turn CUDA errors into exceptions
// (source/EngineGpuKernels/Macros.cuh:15-59) one level up.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/alien_hip.h"
#include "capi_guard.hpp"
#include "worldset.hpp"

using alien::Error;
using alien::WorldImage;
using alien::WorldSet;

std::string& alien::last_error() {
    thread_local std::string message;
    return message;
}

namespace {
using alien::guarded;

void require(bool cond, const char* msg) {
    if (!cond) throw Error(ALIEN_E_ARG, msg);
}

void copyOut(const std::vector<alien_cell_to>& c, const std::vector<alien_particle_to>& p, alien_cell_to* cells,
             uint64_t cellCap, uint64_t* numCells, alien_particle_to* parts, uint64_t partCap, uint64_t* numParts) {
    if (numCells) *numCells = c.size();
    if (numParts) *numParts = p.size();
    if (c.size() > cellCap || p.size() > partCap)
        throw Error(ALIEN_E_CAPACITY, "output buffer too small (counts were written)");
    if (!c.empty()) std::memcpy(cells, c.data(), c.size() * sizeof(alien_cell_to));
    if (!p.empty()) std::memcpy(parts, p.data(), p.size() * sizeof(alien_particle_to));
}

constexpr int kStripCapMax = alien::kStripMax;

double nowSeconds() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}
}  // namespace

// ---- one world --------------------------------------------------------------------------------
struct alien_world {
    alien_world_desc desc{};
    alien_params params{};
    std::unique_ptr<WorldSet> set;
    bool poisoned = false;
    std::string poisonReason;
    uint64_t pendingTimestep = 0;
};

namespace {
void requireHealthy(const alien_world* w) {
    require(w != nullptr, "null world");
    if (w->poisoned)
        throw Error(ALIEN_E_STATE, "world is in an invalid state after a failed step: " + w->poisonReason);
}
}  // namespace

extern "C" {

const char* alien_last_error(void) { return alien::last_error().c_str(); }
int alien_abi_version(void) { return ALIEN_ABI_VERSION; }

int alien_get_device_name(char* buf, size_t cap) {
    return guarded([&] {
        require(buf && cap > 0, "null buffer");
        int dev = 0;
        alien::check_hip(hipGetDevice(&dev), "hipGetDevice");
        hipDeviceProp_t prop{};
        alien::check_hip(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
        std::string name = prop.name[0] ? std::string(prop.name) : std::string("AMD GPU");
        name += " (" + std::string(prop.gcnArchName) + ", " + std::to_string(prop.multiProcessorCount) + " CUs, " +
                std::to_string(prop.totalGlobalMem >> 30) + " GiB)";
        std::strncpy(buf, name.c_str(), cap - 1);
        buf[cap - 1] = 0;
    });
}

void alien_params_default(alien_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->timestepSize = 1.0f;
    p->motionType = ALIEN_MOTION_FLUID;
    p->smoothingLength = 0.8f;
    p->viscosityStrength = 0.1f;
    p->pressureStrength = 0.1f;
    p->cellMaxCollisionDistance = 1.3f;
    p->cellRepulsionStrength = 0.08f;
    p->friction = 0.001f;
    p->innerFriction = 0.3f;
    p->cellMaxVelocity = 2.0f;
    p->cellMinDistance = 0.3f;
    p->cellFusionVelocity = 0.6f;
    p->cellMaxBindingEnergy = std::numeric_limits<float>::infinity();
    p->cellMaxForceDecayProb = 0.2f;
    for (int c = 0; c < ALIEN_MAX_COLORS; ++c) {
        p->cellMaxBindingDistance[c] = 3.6f;
        p->radiationAbsorption[c] = 1.0f;
        p->cellMaxForce[c] = 0.8f;
        p->cellMinEnergy[c] = 50.0f;
        p->radiationMinCellAge[c] = 0;
        p->highRadiationFactor[c] = 0.0f;
        p->highRadiationMinCellEnergy[c] = 500.0f;
        p->radiationCellAgeStrength[c] = 0.00002f;
        p->particleSplitEnergy[c] = std::numeric_limits<float>::infinity();
    }
    p->radiationProb = 0.03f;
    p->radiationVelocityMultiplier = 1.0f;
    p->radiationVelocityPerturbation = 0.5f;
}

int alien_world_create(const alien_world_desc* desc, const alien_params* params, alien_world** out) {
    return guarded([&] {
        require(desc && params && out, "null argument");
        alien::validate_world(*desc, *params);
        auto w = std::make_unique<alien_world>();
        w->desc = *desc;
        w->params = *params;
        w->set = std::make_unique<WorldSet>();
        WorldImage empty = WorldImage::fromTO(*desc, *params, nullptr, 0, nullptr, 0);
        std::vector<WorldSet::Source> src(1);
        src[0].image = &empty;
        w->set->build(src);
        *out = w.release();
    });
}

int alien_world_destroy(alien_world* w) {
    return guarded([&] { delete w; });
}

int alien_set_params(alien_world* w, const alien_params* params) {
    return guarded([&] {
        require(w && params, "null argument");
        alien::validate_world(w->desc, *params);
        if (!w->poisoned) {
            if (alien::bin_size_for(*params) != alien::bin_size_for(w->params) ||
                alien::needs_fresh_set(w->params, *params)) {
                // the bin grid or cellMinEnergy changes: a fresh set is built around the new
                // parameters (density reset pending), and replaces the old one only once complete
                WorldImage img;
                w->set->extract(0, img);
                img.host.params = *params;
                img.densityResetPending = true;
                img.updateLivingWork();
                std::vector<WorldSet::Source> src(1);
                src[0].image = &img;
                auto fresh = std::make_unique<WorldSet>();
                fresh->setBlockSize(w->set->blockSize());
                fresh->setKernelTiming(w->set->kernelTiming());
                fresh->build(src);
                w->set = std::move(fresh);
            } else {
                try {
                    w->set->setParams(0, *params);
                } catch (const Error& e) {
                    w->poisoned = true;
                    w->poisonReason = e.what();
                    throw;
                }
            }
        }
        w->params = *params;
    });
}

int alien_get_params(const alien_world* w, alien_params* params) {
    return guarded([&] {
        require(w && params, "null argument");
        *params = w->params;
    });
}

int alien_set_data(alien_world* w, const alien_cell_to* cells, uint64_t numCells, const alien_particle_to* particles,
                   uint64_t numParticles) {
    return guarded([&] {
        require(w != nullptr, "null world");
        uint64_t ts = w->poisoned ? w->pendingTimestep : w->set->timestep(0);
        WorldImage img = WorldImage::fromTO(w->desc, w->params, cells, numCells, particles, numParticles);
        img.timestep = ts;
        std::vector<WorldSet::Source> src(1);
        src[0].image = &img;
        auto fresh = std::make_unique<WorldSet>();
        fresh->setBlockSize(w->set->blockSize());
        fresh->setKernelTiming(w->set->kernelTiming());
        fresh->build(src);
        w->set = std::move(fresh);
        w->poisoned = false;
    });
}

int alien_get_data(alien_world* w, alien_cell_to* cells, uint64_t cellCap, uint64_t* numCells,
                   alien_particle_to* particles, uint64_t particleCap, uint64_t* numParticles) {
    return guarded([&] {
        requireHealthy(w);
        WorldImage img;
        w->set->extract(0, img);
        std::vector<alien_cell_to> c;
        std::vector<alien_particle_to> p;
        img.toTO(c, p);
        copyOut(c, p, cells, cellCap, numCells, particles, particleCap, numParticles);
    });
}

int alien_get_counts(alien_world* w, uint64_t* numCells, uint64_t* numParticles) {
    return guarded([&] {
        requireHealthy(w);
        if (numCells) *numCells = uint64_t(w->set->numCells(0));
        if (numParticles) *numParticles = uint64_t(w->set->liveParticles(0));
    });
}

int alien_step(alien_world* w, uint64_t steps) {
    return guarded([&] {
        requireHealthy(w);
        try {
            w->set->step(steps);
        } catch (const Error& e) {
            w->poisoned = true;
            w->poisonReason = e.what();
            throw;
        }
    });
}

uint64_t alien_get_timestep(const alien_world* w) {
    if (!w || w->poisoned) return 0;
    try {
        return const_cast<alien_world*>(w)->set->timestep(0);
    } catch (...) {
        return 0;
    }
}

int alien_set_timestep(alien_world* w, uint64_t timestep) {
    return guarded([&] {
        require(w != nullptr, "null world");
        if (w->poisoned) {
            w->pendingTimestep = timestep;
            return;
        }
        w->set->setTimestep(0, timestep);
    });
}

int alien_get_totals(alien_world* w, alien_totals* out) {
    return guarded([&] {
        requireHealthy(w);
        require(out != nullptr, "null argument");
        *out = w->set->totals(0);
    });
}

int alien_get_statistics(alien_world* w, alien_statistics* out) {
    return guarded([&] {
        requireHealthy(w);
        require(out != nullptr, "null argument");
        *out = w->set->statistics(0);
    });
}

int alien_get_statistics_noalloc(alien_world* w, alien_statistics* out) {
    return guarded([&] {
        requireHealthy(w);
        require(out != nullptr, "null argument");
        w->set->statisticsNoAlloc(0, out);
    });
}

int alien_clear(alien_world* w) {
    return alien_set_data(w, nullptr, 0, nullptr, 0);
}

int alien_debug_get_map(alien_world* w, int32_t* cellBin, uint64_t* cellId, uint64_t cap, uint64_t* numCells,
                        int32_t* binSize, int32_t* binsX, int32_t* binsY) {
    return guarded([&] {
        requireHealthy(w);
        std::vector<int> bins, slots;
        w->set->cellMap(0, bins, slots);
        if (numCells) *numCells = bins.size();
        if (binSize) *binSize = alien::bin_size_for(w->params);
        if (binsX) *binsX = w->desc.worldSizeX / alien::bin_size_for(w->params);
        if (binsY) *binsY = w->desc.worldSizeY / alien::bin_size_for(w->params);
        if (bins.size() > cap) throw Error(ALIEN_E_CAPACITY, "output buffer too small (count was written)");
        const auto& ids = w->set->host(0).cellId;
        for (size_t k = 0; k < bins.size(); ++k) {
            cellBin[k] = bins[k];
            cellId[k] = ids[slots[k]];
        }
    });
}

int alien_debug_neighbors(alien_world* w, uint64_t* offsets, uint64_t offsetsCap, uint64_t* neighborIds,
                          uint64_t idsCap, uint64_t* numIds) {
    return guarded([&] {
        requireHealthy(w);
        std::vector<std::vector<int>> bySlot;
        w->set->neighbors(0, bySlot);
        const auto& ids = w->set->host(0).cellId;
        const int n = int(bySlot.size());
        std::vector<int> order(n);
        for (int i = 0; i < n; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return ids[a] < ids[b]; });
        uint64_t total = 0;
        for (auto& v : bySlot) total += v.size();
        if (numIds) *numIds = total;
        if (offsetsCap < uint64_t(n) + 1 || idsCap < total)
            throw Error(ALIEN_E_CAPACITY, "output buffer too small (count was written)");
        uint64_t o = 0;
        offsets[0] = 0;
        for (int k = 0; k < n; ++k) {
            for (int s : bySlot[order[k]]) neighborIds[o++] = ids[s];
            offsets[k + 1] = o;
        }
    });
}

static void copyTimes(WorldSet& set, alien_kernel_time* out, int32_t cap, int32_t* count) {
    auto times = set.takeKernelTimes();
    int32_t n = 0;
    for (auto& kv : times) {
        if (n < cap && out) {
            std::memset(&out[n], 0, sizeof(alien_kernel_time));
            std::strncpy(out[n].name, kv.first.c_str(), sizeof(out[n].name) - 1);
            out[n].launches = kv.second.launches;
            out[n].totalMs = kv.second.totalMs;
        }
        ++n;
    }
    if (count) *count = n;
    if (n > cap) throw Error(ALIEN_E_CAPACITY, "timing buffer too small (count was written)");
}

int alien_set_kernel_timing(alien_world* w, int32_t enable) {
    return guarded([&] {
        requireHealthy(w);
        w->set->setKernelTiming(enable != 0);
    });
}

int alien_get_kernel_timing(alien_world* w, alien_kernel_time* out, int32_t cap, int32_t* count) {
    return guarded([&] {
        requireHealthy(w);
        copyTimes(*w->set, out, cap, count);
    });
}

int alien_get_path_counts(alien_world* w, uint64_t* fallbackChunks, uint64_t* directCells) {
    return guarded([&] {
        requireHealthy(w);
        if (!fallbackChunks || !directCells) throw Error(ALIEN_E_ARG, "null output pointer");
        w->set->pathCounts(*fallbackChunks, *directCells);
    });
}

// ---- batched worlds ---------------------------------------------------------------------------
}  // extern "C"

struct alien_batch {
    std::vector<WorldImage> pending;  // added but not yet resident
    std::unique_ptr<WorldSet> set;
    int resident = 0;
    uint64_t cells = 0, particles = 0;  // totals of the resident and pending worlds
    int blockSize = 256;
    int stripCap = kStripCapMax;  // launch shape chosen by the autotuner, kept when worlds are added
    int chunkRows = 0;
    bool useGraph = true;
    bool timing = false;
    void flush() {
        if (pending.empty()) return;
        auto fresh = std::make_unique<WorldSet>();
        fresh->setBlockSize(blockSize);
        fresh->setStripCap(stripCap);
        fresh->setChunkRows(chunkRows);
        fresh->setUseGraph(useGraph);
        fresh->setKernelTiming(timing);
        std::vector<WorldSet::Source> src;
        for (int i = 0; i < resident; ++i) {
            WorldSet::Source s;
            s.from = set.get();
            s.fromIndex = i;
            src.push_back(s);
        }
        for (auto& img : pending) {
            WorldSet::Source s;
            s.image = &img;
            src.push_back(s);
        }
        fresh->build(src);
        set = std::move(fresh);
        resident += int(pending.size());
        pending.clear();
    }
};

extern "C" {

int alien_batch_create(alien_batch** out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        *out = new alien_batch();
    });
}

int alien_batch_destroy(alien_batch* b) {
    return guarded([&] { delete b; });
}

int alien_batch_add_world(alien_batch* b, const alien_world_desc* desc, const alien_params* params,
                          const alien_cell_to* cells, uint64_t numCells, const alien_particle_to* particles,
                          uint64_t numParticles, int32_t* worldIndex) {
    return guarded([&] {
        require(b && desc && params, "null argument");
        // the batch is one world set: its totals must stay within the set limits
        alien::check_set_size(b->cells + numCells, b->particles + numParticles);
        b->pending.push_back(WorldImage::fromTO(*desc, *params, cells, numCells, particles, numParticles));
        b->cells += numCells;
        b->particles += numParticles;
        if (worldIndex) *worldIndex = b->resident + int(b->pending.size()) - 1;
    });
}

int alien_batch_num_worlds(const alien_batch* b) { return b ? b->resident + int(b->pending.size()) : 0; }

int alien_batch_step(alien_batch* b, uint64_t steps) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        b->flush();
        if (b->set) b->set->step(steps);
    });
}

int alien_batch_get_data(alien_batch* b, int32_t world, alien_cell_to* cells, uint64_t cellCap, uint64_t* numCells,
                         alien_particle_to* particles, uint64_t particleCap, uint64_t* numParticles) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        b->flush();
        require(world >= 0 && world < b->resident, "world index out of range");
        WorldImage img;
        b->set->extract(world, img);
        std::vector<alien_cell_to> c;
        std::vector<alien_particle_to> p;
        img.toTO(c, p);
        copyOut(c, p, cells, cellCap, numCells, particles, particleCap, numParticles);
    });
}

int alien_batch_get_counts(alien_batch* b, int32_t world, uint64_t* numCells, uint64_t* numParticles) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        b->flush();
        require(world >= 0 && world < b->resident, "world index out of range");
        if (numCells) *numCells = uint64_t(b->set->numCells(world));
        if (numParticles) *numParticles = uint64_t(b->set->liveParticles(world));
    });
}

uint64_t alien_batch_get_timestep(const alien_batch* b, int32_t world) {
    if (!b || !b->set || world < 0 || world >= b->resident) return 0;
    try {
        return b->set->timestep(world);
    } catch (...) {
        return 0;
    }
}

int alien_batch_set_block_size(alien_batch* b, int32_t threads) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        if (threads < 64 || threads > 256 || threads % 64 != 0)
            throw Error(ALIEN_E_ARG, "block size must be 64, 128, 192 or 256");
        b->blockSize = threads;
        if (b->set) b->set->setBlockSize(threads);
    });
}

int alien_batch_set_kernel_timing(alien_batch* b, int32_t enable) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        b->timing = enable != 0;
        if (b->set) b->set->setKernelTiming(b->timing);
    });
}

int alien_batch_get_path_counts(alien_batch* b, uint64_t* fallbackChunks, uint64_t* directCells) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        if (!fallbackChunks || !directCells) throw Error(ALIEN_E_ARG, "null output pointer");
        *fallbackChunks = *directCells = 0;
        if (b->set) b->set->pathCounts(*fallbackChunks, *directCells);
    });
}

int alien_batch_get_kernel_timing(alien_batch* b, alien_kernel_time* out, int32_t cap, int32_t* count) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        if (!b->set) {
            if (count) *count = 0;
            return;
        }
        copyTimes(*b->set, out, cap, count);
    });
}

int alien_batch_autotune(alien_batch* b, uint32_t trialSteps, int32_t* blockSize, int32_t* stripCap,
                         int32_t* chunkRows, double* msPerStep) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        require(trialSteps > 0, "trialSteps must be positive");
        b->flush();
        if (!b->set) return;
        // time every launch shape on a copy of the batch; results never depend on the shape
        std::vector<WorldSet::Source> src(b->set->numWorlds());
        for (int i = 0; i < int(src.size()); ++i) {
            src[i].from = b->set.get();
            src[i].fromIndex = i;
        }
        WorldSet trial;
        trial.setUseGraph(b->useGraph);
        trial.build(src);
        double best = 1e300;
        int bestBlock = b->blockSize, bestStrip = b->set->stripCap(), bestRows = 0;
        auto time = [&]() {
            trial.step(1);  // captures the graphs of this shape
            const double t0 = nowSeconds();
            trial.step(trialSteps);
            return (nowSeconds() - t0) / double(trialSteps);
        };
        // block size x strip cap with the default chunk rows, then the chunk rows for the best
        trial.setChunkRows(0);
        for (int blk : {64, 128, 256})
            for (int sc : {kStripCapMax, 32, 16}) {
                trial.setBlockSize(blk);
                trial.setStripCap(sc);
                const double dt = time();
                if (dt < best) {
                    best = dt;
                    bestBlock = blk;
                    bestStrip = sc;
                }
            }
        trial.setBlockSize(bestBlock);
        trial.setStripCap(bestStrip);
        for (int rows : {2, 4, alien::kChunkMax}) {
            trial.setChunkRows(rows);
            const double dt = time();
            if (dt < best) {
                best = dt;
                bestRows = rows;
            }
        }
        b->blockSize = bestBlock;
        b->stripCap = bestStrip;
        b->chunkRows = bestRows;
        b->set->setBlockSize(bestBlock);
        b->set->setStripCap(bestStrip);
        b->set->setChunkRows(bestRows);
        if (blockSize) *blockSize = bestBlock;
        if (stripCap) *stripCap = bestStrip;
        if (chunkRows) *chunkRows = bestRows;
        if (msPerStep) *msPerStep = best * 1000.0;
    });
}

int alien_batch_set_use_graph(alien_batch* b, int32_t enable) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        b->useGraph = enable != 0;
        if (b->set) b->set->setUseGraph(b->useGraph);
    });
}

}  // extern "C"

// ---- serving ----------------------------------------------------------------------------------
namespace {
struct Job {
    uint64_t id = 0;
    int state = ALIEN_JOB_QUEUED;
    uint64_t stepsRequested = 0, stepsDone = 0;
    uint64_t numCells = 0, numParticles = 0;
    uint64_t bytes = 0;
    double submitTime = 0, startTime = 0, finishTime = 0;
    std::unique_ptr<WorldImage> image;
    std::vector<alien_cell_to> snapCells;
    std::vector<alien_particle_to> snapParticles;
    std::string error;
    bool cancel = false;
};
}  // namespace

struct alien_server {
    uint64_t budget = 0;
    uint32_t chunk = 50;
    std::mutex mu;
    std::condition_variable cv;      // wakes the controller
    std::condition_variable doneCv;  // wakes waiters
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<uint64_t> queue;
    std::vector<uint64_t> running;  // order = world index in the resident set
    std::unique_ptr<WorldSet> set;
    uint64_t nextId = 1;
    bool stop = false;
    std::thread controller;

    uint64_t usedBytes() const {
        uint64_t s = 0;
        for (auto id : running)
            if (id) s += jobs.at(id)->bytes;
        return s;
    }

    // `running` lists the jobs in world order of the resident set; 0 marks a finished world that
    // stays in the set, inactive, until the set is next rebuilt (when jobs are admitted or
    // cancelled), so a completion costs no copy of the other worlds.
    void loop() {
        int device = 0;
        (void)hipGetDevice(&device);
        std::unique_lock<std::mutex> lk(mu);
        while (true) {
            cv.wait(lk, [&] { return stop || !queue.empty() || !running.empty(); });
            if (stop) break;
            // drop cancelled running jobs, admit queued jobs that fit the budget (FIFO)
            std::vector<uint64_t> keep;
            bool changed = false;
            for (auto id : running) {
                if (!id) continue;
                if (jobs.at(id)->cancel) {
                    jobs.erase(id);
                    changed = true;
                } else {
                    keep.push_back(id);
                }
            }
            std::vector<uint64_t> admitted;
            uint64_t kept = 0, cells = 0, parts = 0;
            for (auto id : keep) {
                kept += jobs.at(id)->bytes;
                cells += jobs.at(id)->numCells;
                parts += jobs.at(id)->numParticles;
            }
            // the old set keeps its memory until the new one is built (the budget bounds the peak)
            const uint64_t oldSet = set ? uint64_t(set->deviceBytes()) : 0;
            uint64_t used = 0;
            while (!queue.empty()) {
                uint64_t id = queue.front();
                auto it = jobs.find(id);
                if (it == jobs.end() || it->second->cancel) {
                    queue.pop_front();
                    if (it != jobs.end()) jobs.erase(it);
                    continue;
                }
                Job& j = *it->second;
                // the budget, and the size limits of one world set (alien::kMaxSetCells)
                if (alien_server_plan_admission(budget, oldSet, kept + used, &j.bytes, 1) == 0 ||
                    cells + j.numCells > uint64_t(alien::kMaxSetCells) ||
                    parts + j.numParticles > uint64_t(alien::kMaxSetParticles))
                    break;
                used += j.bytes;
                cells += j.numCells;
                parts += j.numParticles;
                admitted.push_back(id);
                queue.pop_front();
            }
            if (!admitted.empty()) changed = true;
            std::vector<WorldImage*> newImages;
            for (auto id : admitted) newImages.push_back(jobs.at(id)->image.get());
            std::vector<uint64_t> members = keep;
            members.insert(members.end(), admitted.begin(), admitted.end());
            if (members.empty()) {
                running.clear();
                set.reset();
                continue;
            }
            uint64_t chunkSteps = chunk;
            for (auto id : members) {
                const Job& j = *jobs.at(id);
                chunkSteps = std::min<uint64_t>(chunkSteps, j.stepsRequested - j.stepsDone);
            }
            // device work happens without the lock; the job objects stay alive (only this thread
            // erases running jobs)
            std::vector<uint64_t> oldRunning = running;
            lk.unlock();
            std::string err, buildErr;
            try {
                if (changed) {
                    // rebuild: the live worlds (device copies) and the admitted ones, finished
                    // worlds left out
                    auto fresh = std::make_unique<WorldSet>();
                    std::vector<WorldSet::Source> src;
                    for (size_t i = 0; i < members.size(); ++i) {
                        WorldSet::Source s;
                        auto pos = std::find(oldRunning.begin(), oldRunning.end(), members[i]);
                        if (pos != oldRunning.end()) {
                            s.from = set.get();
                            s.fromIndex = int(pos - oldRunning.begin());
                        } else {
                            s.image = newImages[i - keep.size()];
                        }
                        src.push_back(s);
                    }
                    try {
                        fresh->build(src);
                    } catch (const std::exception& e) {
                        buildErr = e.what();
                    }
                    if (buildErr.empty()) set = std::move(fresh);
                }
                if (buildErr.empty() && chunkSteps > 0) set->step(chunkSteps);
            } catch (const std::exception& e) {
                err = e.what();
            }
            lk.lock();
            const double t = nowSeconds();
            if (!buildErr.empty()) {
                // the old set is untouched: only the jobs being admitted fail; the resident ones go
                // on with the old set (cancelled ones were dropped from it above)
                for (auto id : admitted) {
                    Job& j = *jobs.at(id);
                    j.state = ALIEN_JOB_FAILED;
                    j.error = buildErr;
                    j.image.reset();
                    j.finishTime = t;
                }
                for (auto& id : oldRunning)
                    if (id && !jobs.count(id)) id = 0;  // cancelled and erased: inactive in place
                running = oldRunning;
                for (size_t i = 0; i < running.size(); ++i)
                    if (!running[i] && set && i < size_t(set->numWorlds())) set->setActive(int(i), false);
                bool anyLive = false;
                for (auto id : running) anyLive = anyLive || id != 0;
                if (!anyLive) {
                    running.clear();
                    set.reset();
                }
                doneCv.notify_all();
                continue;
            }
            for (auto id : admitted) {
                Job& j = *jobs.at(id);
                j.state = ALIEN_JOB_RUNNING;
                j.startTime = t;
                j.image.reset();
            }
            running = changed ? members : oldRunning;
            if (!err.empty()) {
                for (auto id : running) {
                    if (!id) continue;
                    Job& j = *jobs.at(id);
                    j.state = ALIEN_JOB_FAILED;
                    j.error = err;
                    j.finishTime = nowSeconds();
                }
                running.clear();
                set.reset();
                doneCv.notify_all();
                continue;
            }
            // finished jobs: snapshot, then the world goes inactive in place
            std::vector<std::pair<uint64_t, int>> finished;
            for (size_t i = 0; i < running.size(); ++i) {
                if (!running[i]) continue;
                Job& j = *jobs.at(running[i]);
                j.stepsDone += chunkSteps;
                if (j.stepsDone >= j.stepsRequested) finished.push_back({running[i], int(i)});
            }
            if (finished.empty()) continue;
            lk.unlock();
            std::vector<std::pair<uint64_t, std::pair<std::vector<alien_cell_to>, std::vector<alien_particle_to>>>>
                snaps;
            std::string snapErr;
            try {
                for (auto& f : finished) {
                    WorldImage img;
                    set->extract(f.second, img);
                    std::vector<alien_cell_to> c;
                    std::vector<alien_particle_to> p;
                    img.toTO(c, p);
                    snaps.push_back({f.first, {std::move(c), std::move(p)}});
                    set->setActive(f.second, false);
                }
            } catch (const std::exception& e) {
                snapErr = e.what();
            }
            lk.lock();
            for (auto& sn : snaps) {
                Job& j = *jobs.at(sn.first);
                j.snapCells = std::move(sn.second.first);
                j.snapParticles = std::move(sn.second.second);
                j.numCells = j.snapCells.size();
                j.numParticles = j.snapParticles.size();
                j.state = ALIEN_JOB_DONE;
                j.finishTime = nowSeconds();
            }
            for (auto& f : finished) {
                Job& j = *jobs.at(f.first);
                if (j.state != ALIEN_JOB_DONE) {
                    j.state = ALIEN_JOB_FAILED;
                    j.error = snapErr;
                    j.finishTime = nowSeconds();
                }
                running[size_t(f.second)] = 0;
            }
            bool anyLive = false;
            for (auto id : running) anyLive = anyLive || id != 0;
            if (!anyLive) {
                running.clear();
                lk.unlock();
                set.reset();
                lk.lock();
            }
            doneCv.notify_all();
        }
    }
};

extern "C" {

uint64_t alien_server_estimate_bytes(uint64_t numCells, uint64_t numParticles, int32_t worldSizeX, int32_t worldSizeY) {
    return WorldSet::estimateBytes(numCells, numParticles, worldSizeX, worldSizeY, 2);
}

uint64_t alien_server_plan_admission(uint64_t budgetBytes, uint64_t oldSetBytes, uint64_t keptBytes,
                                     const uint64_t* queuedBytes, uint64_t numQueued) {
    uint64_t need = oldSetBytes + keptBytes, n = 0;
    for (; n < numQueued && queuedBytes; ++n) {
        if (need + queuedBytes[n] > budgetBytes || need + queuedBytes[n] < need) break;
        need += queuedBytes[n];
    }
    return n;
}

int alien_server_create(uint64_t budgetBytes, uint32_t chunkSteps, alien_server** out) {
    return guarded([&] {
        require(out != nullptr, "null argument");
        require(chunkSteps > 0, "chunkSteps must be positive");
        auto s = new alien_server();
        s->budget = budgetBytes;
        s->chunk = chunkSteps;
        s->controller = std::thread([s] { s->loop(); });
        *out = s;
    });
}

int alien_server_destroy(alien_server* s) {
    return guarded([&] {
        if (!s) return;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv.notify_all();
        if (s->controller.joinable()) s->controller.join();
        delete s;
    });
}

int alien_server_submit(alien_server* s, const alien_world_desc* desc, const alien_params* params,
                        const alien_cell_to* cells, uint64_t numCells, const alien_particle_to* particles,
                        uint64_t numParticles, uint64_t steps, uint64_t* jobId) {
    return guarded([&] {
        require(s && desc && params, "null argument");
        auto job = std::make_unique<Job>();
        job->image = std::make_unique<WorldImage>(
            WorldImage::fromTO(*desc, *params, cells, numCells, particles, numParticles));
        job->stepsRequested = steps;
        job->numCells = numCells;
        job->numParticles = numParticles;
        // (a radiating or splitting world is built with room for the particles it will create)
        job->bytes = WorldSet::estimateBytes(
            numCells, uint64_t(alien::particle_room(int64_t(numCells), int64_t(numParticles), *params)),
            desc->worldSizeX, desc->worldSizeY, alien::bin_size_for(*params));
        if (job->bytes > s->budget) throw Error(ALIEN_E_BUDGET, "job does not fit the device memory budget");
        job->submitTime = nowSeconds();
        std::lock_guard<std::mutex> lk(s->mu);
        job->id = s->nextId++;
        if (jobId) *jobId = job->id;
        if (steps == 0) {
            job->image->toTO(job->snapCells, job->snapParticles);
            job->state = ALIEN_JOB_DONE;
            job->startTime = job->finishTime = job->submitTime;
            job->image.reset();
            s->jobs[job->id] = std::move(job);
            s->doneCv.notify_all();
            return;
        }
        s->queue.push_back(job->id);
        s->jobs[job->id] = std::move(job);
        s->cv.notify_all();
    });
}

int alien_server_poll(alien_server* s, uint64_t jobId, alien_job_info* info) {
    return guarded([&] {
        require(s && info, "null argument");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(jobId);
        if (it == s->jobs.end()) throw Error(ALIEN_E_NOTFOUND, "unknown job id");
        const Job& j = *it->second;
        info->state = j.state;
        info->stepsDone = j.stepsDone;
        info->stepsRequested = j.stepsRequested;
        info->numCells = j.numCells;
        info->numParticles = j.numParticles;
        info->submitTime = j.submitTime;
        info->startTime = j.startTime;
        info->finishTime = j.finishTime;
        info->deviceBytes = j.bytes;
        if (j.state == ALIEN_JOB_FAILED) alien::last_error() = j.error;
    });
}

int alien_server_wait(alien_server* s, uint64_t jobId, double timeoutSeconds) {
    return guarded([&] {
        require(s != nullptr, "null server");
        std::unique_lock<std::mutex> lk(s->mu);
        auto pred = [&] {
            auto it = s->jobs.find(jobId);
            return it == s->jobs.end() || it->second->state == ALIEN_JOB_DONE || it->second->state == ALIEN_JOB_FAILED;
        };
        bool ok = s->doneCv.wait_for(lk, std::chrono::duration<double>(timeoutSeconds), pred);
        auto it = s->jobs.find(jobId);
        if (it == s->jobs.end()) throw Error(ALIEN_E_NOTFOUND, "unknown job id");
        if (!ok) throw Error(ALIEN_E_STATE, "timeout waiting for job");
        if (it->second->state == ALIEN_JOB_FAILED) throw Error(ALIEN_E_STATE, "job failed: " + it->second->error);
    });
}

int alien_server_get_snapshot(alien_server* s, uint64_t jobId, alien_cell_to* cells, uint64_t cellCap,
                              uint64_t* numCells, alien_particle_to* particles, uint64_t particleCap,
                              uint64_t* numParticles) {
    return guarded([&] {
        require(s != nullptr, "null server");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(jobId);
        if (it == s->jobs.end()) throw Error(ALIEN_E_NOTFOUND, "unknown job id");
        const Job& j = *it->second;
        if (j.state != ALIEN_JOB_DONE) throw Error(ALIEN_E_STATE, "job not finished");
        copyOut(j.snapCells, j.snapParticles, cells, cellCap, numCells, particles, particleCap, numParticles);
    });
}

int alien_server_release(alien_server* s, uint64_t jobId) {
    return guarded([&] {
        require(s != nullptr, "null server");
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(jobId);
        if (it == s->jobs.end()) throw Error(ALIEN_E_NOTFOUND, "unknown job id");
        Job& j = *it->second;
        if (j.state == ALIEN_JOB_DONE || j.state == ALIEN_JOB_FAILED) {
            s->jobs.erase(it);
        } else {
            j.cancel = true;
            s->cv.notify_all();
        }
    });
}

}  // extern "C"
