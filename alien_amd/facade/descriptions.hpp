// descriptions.hpp — the data descriptions of ALIEN's facade and their conversion to and from the
// transfer objects of the C ABI.
//
// Mirrors DataDescription / ClusteredDataDescription (source/EngineInterface/Descriptions.h:363-669)
// and DescriptionConverter (source/EngineImpl/DescriptionConverter.cpp:132-711) for the physics
// path.  The transfer objects of include/alien_hip.h carry what the timestep reads and the identity
// fields; the rest of a cell's description — its cell-function state, genome bytes and metadata
// (name, description), which ALIEN keeps in DataTO::auxiliaryData and the cell-function union of
// CellTO — is kept on the host by cell id (AuxStore) and reattached on the way back, so it passes
// through a simulation unchanged.  Host-only code: nothing here touches the GPU.
#pragma once

#include <cstdint>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/alien_hip.h"

namespace alien {

// DataTO-shaped host data (TOs.cuh:195-236) in vectors
struct SimulationData {
    std::vector<alien_cell_to> cells;
    std::vector<alien_particle_to> particles;
};

struct RealVector2D {
    float x = 0, y = 0;
};

// ConnectionDescription (Descriptions.h:29-52): cellId 0 means "not present in the data"
struct ConnectionDescription {
    uint64_t cellId = 0;
    float distance = 0;
    float angleFromPrevious = 0;
};

// CellMetadataDescription (Descriptions.h:10-27)
struct CellMetadataDescription {
    std::string name;
    std::string description;
};

// The cell-function part of CellDescription (Descriptions.h:353-361, 385-391), which the physics
// path carries through: the function type, its state as opaque bytes, and its genome (constructor
// and injector cells)
struct CellFunctionDescription {
    uint8_t type = 0;  // CellFunction_* (CellFunctionConstants.h); 0 = None
    std::vector<uint8_t> state;
    std::vector<uint8_t> genome;
};

// CellDescription (Descriptions.h:363-525), physics subset plus the pass-through fields
struct CellDescription {
    uint64_t id = 0;
    std::vector<ConnectionDescription> connections;
    RealVector2D pos;
    RealVector2D vel;
    float energy = 100.0f;
    float stiffness = 1.0f;
    int color = 0;
    int maxConnections = 0;
    bool barrier = false;
    int age = 0;
    int livingState = 0;  // LivingState_* (CellFunctionConstants.h:26-35)
    int creatureId = 0;
    int mutationId = 0;
    uint8_t ancestorMutationId = 0;
    float genomeComplexity = 0;
    int activationTime = 0;
    CellFunctionDescription cellFunction;
    CellMetadataDescription metadata;
};

// ClusterDescription (Descriptions.h:527-546): cells connected by bonds
struct ClusterDescription {
    std::vector<CellDescription> cells;
};

// ParticleDescription (Descriptions.h:548-584)
struct ParticleDescription {
    uint64_t id = 0;
    RealVector2D pos;
    RealVector2D vel;
    float energy = 0;
    int color = 0;
};

// ClusteredDataDescription (Descriptions.h:586-635)
struct ClusteredDataDescription {
    std::vector<ClusterDescription> clusters;
    std::vector<ParticleDescription> particles;
};

// DataDescription (Descriptions.h:637-667); constructible from clustered data (cells of all
// clusters in order)
struct DataDescription {
    std::vector<CellDescription> cells;
    std::vector<ParticleDescription> particles;

    DataDescription() = default;
    explicit DataDescription(ClusteredDataDescription const& clustered);
};

// what the transfer objects do not carry, by cell id
struct CellAux {
    CellFunctionDescription cellFunction;
    CellMetadataDescription metadata;
};
using AuxStore = std::unordered_map<uint64_t, CellAux>;

// NumberGenerator::getId (Base/NumberGenerator.cpp:56-59): ids for entities described with id 0,
// 2^48 | a running number of this process
uint64_t newEntityId();

// DescriptionConverter (DescriptionConverter.cpp).  toTO appends to `out` (convertDescriptionToTO,
// :230-262): a cell or particle with id 0 gets a new id; an invalid energy (NaN, negative, above
// 1e12) becomes 0 (:524-529); bonds to cell id 0 are dropped and their angle goes to the next kept
// bond, a remainder to the first (setConnections, :692-712); a bond to an id that is not in the data
// throws std::invalid_argument, as does a cell with more than ALIEN_MAX_CELL_BONDS kept bonds.  The
// pass-through data of every cell goes into `aux` under its id.
class DescriptionConverter {
public:
    static void toTO(DataDescription const& description, SimulationData& out, AuxStore& aux);
    static void toTO(ClusteredDataDescription const& description, SimulationData& out, AuxStore& aux);
    // convertTOtoDataDescription (:176-201): cells and particles in transfer order
    static DataDescription toDataDescription(SimulationData const& data, AuxStore const& aux);
    // convertTOtoClusteredDataDescription (:132-174): one cluster per connected component of the
    // bond graph.  The reference takes components and their cells in hash-set order; here clusters
    // are ordered by their first cell in transfer order and a cluster lists its cells breadth
    // first from that cell, each cell's partners in bond order.
    static ClusteredDataDescription toClusteredDataDescription(SimulationData const& data, AuxStore const& aux);
};

}  // namespace alien
