// worldset.hpp — host engine of a set of independent worlds resident in device memory.
//
// Plays the role of ALIEN's device seam _SimulationCudaFacade (SimulationCudaFacade.cuh:27-139)
// and its SimulationData (SimulationData.cuh:17-54): owns device state, converts transfer objects
// to the device layout and back (DataAccessKernels.cu, DescriptionConverter.cpp), and runs the
// timestep loop (SimulationCudaFacade.cu:128-160).  Many worlds share one set and one launch
// sequence; a single simulation is a set of one world.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/alien_hip.h"
#include "engine_types.hpp"
#include "kernels.hpp"

namespace alien {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void check_hip(hipError_t e, const char* what);

int bin_size_for(const alien_params& p);
void validate_params(const alien_params& p);
void validate_world(const alien_world_desc& d, const alien_params& p);
// throws ALIEN_E_CAPACITY when a world set of this many cells / particles exceeds kMaxSetCells /
// kMaxSetParticles (checked before any transfer array is read or any device memory is allocated)
void check_set_size(uint64_t cells, uint64_t particles);
// whether new parameters for a world without device-tracked living states need a fresh set
bool needs_fresh_set(const alien_params& before, const alien_params& after);
// whether the parameters let cell radiation take energy from cells (radiationProb > 0 and some
// colour's high-radiation factor or age strength positive)
bool radiates(const alien_params& p);
// whether some colour has a finite particleSplitEnergy (particles can split)
bool splits(const alien_params& p);
// particle slots a world is built with: its particles, plus room for 8 steps of cell radiation
// (one particle per cell and step at most), doubled when its particles can split
int64_t particle_room(int64_t cells, int64_t particles, const alien_params& p);
DevParams to_dev_params(const alien_params& p);
int choose_strip(const std::vector<float2>& pos, int W, int H, int binSize, bool twoRowStages, int& rowsPerStage);
void strip_shape(DevWorld& w, int strip, int rowsPerStage);
void chunk_shape(std::vector<DevWorld>& worlds, const std::vector<DevParams>& params, int forcedRows = 0);
constexpr int kXcds = kXcdCount;  // accelerator complex dies of an MI355X

// host-side identity and pass-through data of one world (fields the physics path never touches).
// age, activationTime and livingState hold the values at the last set (stepsSinceSet steps ago):
// aging adds one per step to every non-barrier cell, and a world whose living states are all
// trivially Ready (no livingWork) keeps them so; otherwise the device arrays are authoritative.
struct HostWorld {
    alien_world_desc desc{};
    alien_params params{};
    int numCells = 0;
    int particleCap = 0;
    uint64_t stepsSinceSet = 0;
    std::vector<uint64_t> cellId;
    std::vector<uint32_t> age, creatureId, mutationId, activationTime;
    std::vector<uint8_t> livingState, cellFunction, ancestorMutationId, selected;
    std::vector<float> genomeComplexity;
    std::vector<uint64_t> particleId;
    std::vector<uint8_t> particleSelected;
};

// complete state of one world in the device layout, staged in host memory
struct WorldImage {
    HostWorld host;
    uint64_t timestep = 0;
    int numParticles = 0;
    bool densityResetPending = false;
    std::vector<float2> pos, vel;
    std::vector<float> dens;
    std::vector<uint32_t> flags;
    std::vector<int> slot;
    std::vector<int> btgt;  // [kMaxBonds][numCells]
    std::vector<float> brest, bang;
    std::vector<float> energy, stiff;  // by slot
    bool livingWork = false;           // living states tracked on the device (see WorldSet::build)
    std::vector<float2> ppos, pvel;
    std::vector<float> pen;
    std::vector<int> plast, pslot;
    std::vector<uint8_t> pflags;
    std::vector<uint64_t> pid;  // particle ids

    // from transfer objects (ObjectFactory.cuh:46-86: positions wrapped, density 1)
    static WorldImage fromTO(const alien_world_desc& desc, const alien_params& params, const alien_cell_to* cells,
                             uint64_t numCells, const alien_particle_to* particles, uint64_t numParticles);
    // to transfer objects, cells and particles in ascending id order
    void toTO(std::vector<alien_cell_to>& cells, std::vector<alien_particle_to>& particles) const;
    // livingWork for the image's current states, energies and parameters (once set, it stays set)
    void updateLivingWork();
};

class WorldSet {
public:
    WorldSet();
    ~WorldSet();
    WorldSet(const WorldSet&) = delete;
    WorldSet& operator=(const WorldSet&) = delete;

    // A world of the new layout comes either from a host image or from a world of another set
    // (copied device to device).
    struct Source {
        const WorldImage* image = nullptr;
        WorldSet* from = nullptr;
        int fromIndex = -1;
    };
    void build(const std::vector<Source>& sources);
    void clearAll();

    int numWorlds() const { return int(hosts_.size()); }
    HostWorld& host(int w) { return hosts_.at(w); }
    const HostWorld& host(int w) const { return hosts_.at(w); }
    int numCells(int w) const { return worlds_.at(w).numCells; }
    int numParticles(int w);   // live range, may still hold particles killed in the last step
    int liveParticles(int w);  // the particles get_data returns
    uint64_t timestep(int w);
    void setTimestep(int w, uint64_t t);
    // new parameters of world w with the same bin grid (bin_size_for unchanged); a change of the
    // grid needs a fresh set built around the new parameters
    void setParams(int w, const alien_params& p);

    // advance every world by n timesteps
    void step(uint64_t n);
    void extract(int w, WorldImage& out);
    alien_totals totals(int w);
    alien_statistics statistics(int w);  // per-colour counts and energies, reduced on the device
    // the same result without allocating device memory (the library's per-device scratch array);
    // ALIEN_E_CAPACITY beyond kStatScratchBlocks blocks of 256 cells or particle slots
    void statisticsNoAlloc(int w, alien_statistics* out);
    // cell map of the current state: bin of each cell in map order, and its slot
    void cellMap(int w, std::vector<int>& bins, std::vector<int>& slots);
    // neighbour lists under the step predicate: per slot, neighbour slots in visiting order
    void neighbors(int w, std::vector<std::vector<int>>& bySlot);

    void setBlockSize(int threads);
    int blockSize() const { return shape_.cellThreads; }
    // neighbour-pass strips no wider than cap bins (8, 16, 32 or 64); launch shape only
    void setStripCap(int cap);
    // An inactive world stays in place but is left out of every launch (its state is frozen);
    // all worlds are active after build.
    void setActive(int w, bool on);
    int stripCap() const { return stripCap_; }
    // bin rows per neighbour-pass chunk (2..kChunkMax), or 0 for the occupancy rule; launch shape only
    void setChunkRows(int rows);
    int chunkRows() const { return chunkRows_; }
    void setUseGraph(bool on);
    // kernel timing mode: no graphs, every launch bracketed by events (see alien_set_kernel_timing)
    void setKernelTiming(bool on);
    bool kernelTiming() const { return timing_; }
    // neighbour-pass fallback counts since the data was set: chunks, cells
    void pathCounts(uint64_t& fallbackChunks, uint64_t& directCells);
    struct KernelTime {
        uint64_t launches = 0;
        double totalMs = 0.0;
    };
    std::map<std::string, KernelTime> takeKernelTimes();
    size_t deviceBytes() const { return deviceBytes_; }
    static uint64_t estimateBytes(uint64_t cells, uint64_t particles, int W, int H, int binSize);
    void synchronize();
    hipStream_t stream() const { return stream_; }

private:
    struct Arr {
        void* p = nullptr;
        size_t bytes = 0;
    };
    void* alloc(size_t bytes);
    void freeAll();
    void buildTables();
    void invalidateGraphs();
    Ctx makeCtx(int parity) const;
    void ensureSorted();
    void checkErrors();
    void dropParticlesIfNone();  // stops launching the particle kernels once no particle is left
    // the same without a host synchronisation: acts on the count read at the previous call (if it
    // has arrived) and starts reading the current one
    void checkParticlesAsync();
    // Cell radiation adds at most one particle per cell and step: before `steps` more steps, every
    // radiating world gets room for its live particles plus steps times its cells (twice that when
    // it grows), and the set gets its staging area (a host synchronisation).
    void ensureParticleRoom(uint64_t steps);
    // moves the particles to arrays with the given capacities per world (and a staging area)
    void growParticles(const std::vector<int64_t>& caps);
    void release(void* p);  // frees one allocation of alloc()
    void updateRadiation();  // shape_.hasRadiation from the active worlds' parameters
    void anchorAges(int w);  // worlds_[w].ageTs from its timestep and steps simulated

    hipStream_t stream_ = nullptr;
    SideStream side_;  // particle kernels beside the neighbour pass
    int* liveHost_ = nullptr;        // pinned: live particles as read by checkParticlesAsync
    hipEvent_t liveReady_ = nullptr;  // recorded after that read
    bool livePending_ = false;
    std::vector<HostWorld> hosts_;
    std::vector<DevWorld> worlds_;
    std::vector<DevParams> params_;
    std::vector<uint8_t> densityPending_;
    std::vector<Arr> allocs_;
    size_t deviceBytes_ = 0;
    int cellCap_ = 0, partCap_ = 0, numBins_ = 0, candCap_ = 0;
    int parity_ = 0;  // index of the buffer holding the current state
    LaunchShape shape_;
    bool useGraph_ = true;

    // device memory
    DevWorld* dWorlds_ = nullptr;
    DevParams* dParams_ = nullptr;
    DevCounters* dCtr_ = nullptr;
    StepCounters* dSc_ = nullptr;
    HotBuffers buf_[2]{};
    float* energy_ = nullptr;
    float* stiff_ = nullptr;
    unsigned long long* cid_ = nullptr;
    uint32_t* ageBase_ = nullptr;  // cell ages by slot at each world's ageTs
    int partStage_ = 0;            // first staging slot of radiation-emitted particles (0: none)
    StagedParticle* staged_ = nullptr;  // [cellCap] with a staging area
    int splitStage_ = 0;                     // first split staging slot (0: none)
    StagedParticle* splitStaged_ = nullptr;  // [splitStage_] with a split staging area
    int numCellBlocks_ = 0;
    uint8_t* living_[2] = {nullptr, nullptr};  // living states by slot, alternating with the parity
    uint32_t* activation_ = nullptr;
    uint32_t* creature_ = nullptr;
    uint8_t* stretched_ = nullptr;
    int2* livingBlocks_ = nullptr;
    std::vector<uint8_t> livingWork_;  // per world
    int *binOf_ = nullptr, *pbinOf_ = nullptr;
    uint32_t* count_[2] = {nullptr, nullptr};  // bin histograms, alternating with the state buffers
    // map build: offsets of the map of the step of that parity and the mover lists (per cell bin)
    // read in it, alternating with the state buffers
    uint32_t* start_[2] = {nullptr, nullptr};
    int* moverHead_[2] = {nullptr, nullptr};
    int* delta_[2] = {nullptr, nullptr};
    int* moverNext_ = nullptr;
    int numCellBins_ = 0;
    unsigned long long* mapStatus_ = nullptr;
    uint32_t* mapEpochTicket_ = nullptr;  // [2]: epoch, ticket
    int *sorted_ = nullptr, *newidx_ = nullptr, *psorted_ = nullptr;
    float2 *mpos_ = nullptr, *mvel_ = nullptr;
    float* mdens_ = nullptr;
    uint32_t* mflags_ = nullptr;
    uint8_t* stagedMark_ = nullptr;
    float2 *own_ = nullptr, *shift_ = nullptr, *s2_ = nullptr, *out1_ = nullptr, *out2_ = nullptr;
    float* bstiff_ = nullptr;
    uint32_t* valid_ = nullptr;
    uint8_t *delFlag_ = nullptr, *mark_ = nullptr;
    unsigned long long* claim_ = nullptr;
    int4* cand_ = nullptr;
    int *want_ = nullptr, *fclaim_ = nullptr, *absorbCell_ = nullptr, *cclaim_ = nullptr;
    int2 *cellBlocks_ = nullptr, *bondBlocks_ = nullptr, *binBlocks_ = nullptr, *pbinBlocks_ = nullptr;
    int2* partBlocks_ = nullptr;
    int2* tileBlocks_ = nullptr;
    int* numPartBlocks_ = nullptr;  // (device) entries of partBlocks_
    std::vector<uint8_t> bondWork_;  // per world: a bonded or barrier cell exists or can appear by fusion
    std::vector<uint8_t> active_;    // per world: included in the launch tables
    std::vector<int> stripChoice_;   // per world: strip width chosen from its cell density
    std::vector<int> stageChoice_;   // per world: bin rows per staged row (1 or 2)
    int stripCap_ = kStripMax;
    int chunkRows_ = 0;
    void adaptGrids();
    void updateInnerFriction();
    std::vector<Arr> tableAllocs_;

    hipGraphExec_t graph_[2] = {nullptr, nullptr};  // one step from parity 0 / 1
    // kStepsPerGraph steps from parity 0 / 1 (an even count: the parity is the same after them);
    // consecutive graph launches leave the device idle for ~20 us between them
    static constexpr uint64_t kStepsPerGraph = 4;
    // steps between the particle-room checks of a set with cell radiation; a radiating world is
    // built with room for twice that many steps of emissions
    static constexpr uint64_t kRadiationSteps = 4;
    hipGraphExec_t graphMulti_[2] = {nullptr, nullptr};

    class EventRecorder;
    std::unique_ptr<EventRecorder> recorder_;
    bool timing_ = false;
    std::map<std::string, KernelTime> kernelTimes_;
};

}  // namespace alien
