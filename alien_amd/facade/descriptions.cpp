// descriptions.cpp — see descriptions.hpp.
#include "descriptions.hpp"

#include <atomic>
#include <cmath>
#include <stdexcept>
#include <string>

namespace alien {

namespace {

std::atomic<uint64_t> g_runningNumber{0};

// checkAndCorrectInvalidEnergy (DescriptionConverter.cpp:524-529)
float validEnergy(float e) { return (std::isnan(e) || e < 0.0f || e > 1e12f) ? 0.0f : e; }

uint8_t toU8(int v, const char* what) {
    if (v < 0 || v > 255) throw std::invalid_argument(std::string(what) + " out of range");
    return uint8_t(v);
}

// addCell (DescriptionConverter.cpp:545-690): the physics and identity fields; bonds follow
void addCell(SimulationData& out, CellDescription const& d, std::unordered_map<uint64_t, int>& indexById,
             AuxStore& aux) {
    alien_cell_to c{};
    c.id = d.id == 0 ? newEntityId() : d.id;
    c.pos[0] = d.pos.x;
    c.pos[1] = d.pos.y;
    c.vel[0] = d.vel.x;
    c.vel[1] = d.vel.y;
    c.energy = validEnergy(d.energy);
    c.stiffness = d.stiffness;
    c.color = toU8(d.color, "cell color");
    c.maxConnections = toU8(d.maxConnections, "maxConnections");
    c.numConnections = 0;
    c.barrier = d.barrier ? 1 : 0;
    c.age = uint32_t(d.age);
    c.livingState = toU8(d.livingState, "livingState");
    c.cellFunction = d.cellFunction.type;
    c.ancestorMutationId = d.ancestorMutationId;
    c.creatureId = uint32_t(d.creatureId);
    c.mutationId = uint32_t(d.mutationId);
    c.genomeComplexity = d.genomeComplexity;
    c.activationTime = uint32_t(d.activationTime);
    for (int b = 0; b < ALIEN_MAX_CELL_BONDS; ++b) c.connections[b].cellIndex = -1;
    indexById.insert_or_assign(c.id, int(out.cells.size()));
    out.cells.push_back(c);
    aux[c.id] = CellAux{d.cellFunction, d.metadata};
}

// setConnections (DescriptionConverter.cpp:692-712)
void setConnections(SimulationData& out, CellDescription const& d, std::unordered_map<uint64_t, int> const& indexById) {
    alien_cell_to& c = out.cells[size_t(indexById.at(d.id))];
    int index = 0;
    float angleOffset = 0;
    for (ConnectionDescription const& cd : d.connections) {
        if (cd.cellId != 0) {
            auto it = indexById.find(cd.cellId);
            if (it == indexById.end())
                throw std::invalid_argument("cell " + std::to_string(d.id) + " is bonded to cell " +
                                            std::to_string(cd.cellId) + ", which is not in the data");
            if (index >= ALIEN_MAX_CELL_BONDS)
                throw std::invalid_argument("cell " + std::to_string(d.id) + " has more than " +
                                            std::to_string(ALIEN_MAX_CELL_BONDS) + " bonds");
            c.connections[index].cellIndex = it->second;
            c.connections[index].distance = cd.distance;
            c.connections[index].angleFromPrevious = cd.angleFromPrevious + angleOffset;
            ++index;
            angleOffset = 0;
        } else {
            angleOffset += cd.angleFromPrevious;
        }
    }
    if (angleOffset != 0 && index > 0) c.connections[0].angleFromPrevious += angleOffset;
    c.numConnections = uint8_t(index);
}

// addParticle (DescriptionConverter.cpp:532-543)
void addParticle(SimulationData& out, ParticleDescription const& d) {
    alien_particle_to p{};
    p.id = d.id == 0 ? newEntityId() : d.id;
    p.pos[0] = d.pos.x;
    p.pos[1] = d.pos.y;
    p.vel[0] = d.vel.x;
    p.vel[1] = d.vel.y;
    p.energy = validEnergy(d.energy);
    p.color = toU8(d.color, "particle color");
    out.particles.push_back(p);
}

// createCellDescription (DescriptionConverter.cpp:363-520): bonds by id (0 for a bond without a
// partner in the data)
CellDescription cellDescription(SimulationData const& data, int i, AuxStore const& aux) {
    const alien_cell_to& c = data.cells[size_t(i)];
    CellDescription d;
    d.id = c.id;
    d.pos = {c.pos[0], c.pos[1]};
    d.vel = {c.vel[0], c.vel[1]};
    d.energy = c.energy;
    d.stiffness = c.stiffness;
    d.color = c.color;
    d.maxConnections = c.maxConnections;
    d.barrier = c.barrier != 0;
    d.age = int(c.age);
    d.livingState = c.livingState;
    d.creatureId = int(c.creatureId);
    d.mutationId = int(c.mutationId);
    d.ancestorMutationId = c.ancestorMutationId;
    d.genomeComplexity = c.genomeComplexity;
    d.activationTime = int(c.activationTime);
    for (int b = 0; b < c.numConnections && b < ALIEN_MAX_CELL_BONDS; ++b) {
        const alien_connection_to& ct = c.connections[b];
        ConnectionDescription cd;
        cd.cellId = (ct.cellIndex >= 0 && size_t(ct.cellIndex) < data.cells.size()) ? data.cells[size_t(ct.cellIndex)].id : 0;
        cd.distance = ct.distance;
        cd.angleFromPrevious = ct.angleFromPrevious;
        d.connections.push_back(cd);
    }
    auto it = aux.find(c.id);
    if (it != aux.end()) {
        d.cellFunction = it->second.cellFunction;
        d.metadata = it->second.metadata;
    }
    d.cellFunction.type = c.cellFunction;
    return d;
}

ParticleDescription particleDescription(alien_particle_to const& p) {
    ParticleDescription d;
    d.id = p.id;
    d.pos = {p.pos[0], p.pos[1]};
    d.vel = {p.vel[0], p.vel[1]};
    d.energy = p.energy;
    d.color = p.color;
    return d;
}

template <class Cells>
void cellsToTO(Cells const& forEachCell, SimulationData& out, AuxStore& aux) {
    std::unordered_map<uint64_t, int> indexById;
    forEachCell([&](CellDescription const& d) { addCell(out, d, indexById, aux); });
    forEachCell([&](CellDescription const& d) {
        if (d.id != 0) setConnections(out, d, indexById);
    });
}

}  // namespace

uint64_t newEntityId() { return (uint64_t(1) << 48) | ++g_runningNumber; }

DataDescription::DataDescription(ClusteredDataDescription const& clustered) {
    for (auto const& cl : clustered.clusters) cells.insert(cells.end(), cl.cells.begin(), cl.cells.end());
    particles = clustered.particles;
}

void DescriptionConverter::toTO(DataDescription const& description, SimulationData& out, AuxStore& aux) {
    cellsToTO(
        [&](auto&& f) {
            for (auto const& c : description.cells) f(c);
        },
        out, aux);
    for (auto const& p : description.particles) addParticle(out, p);
}

void DescriptionConverter::toTO(ClusteredDataDescription const& description, SimulationData& out, AuxStore& aux) {
    cellsToTO(
        [&](auto&& f) {
            for (auto const& cl : description.clusters)
                for (auto const& c : cl.cells) f(c);
        },
        out, aux);
    for (auto const& p : description.particles) addParticle(out, p);
}

DataDescription DescriptionConverter::toDataDescription(SimulationData const& data, AuxStore const& aux) {
    DataDescription d;
    d.cells.reserve(data.cells.size());
    for (int i = 0; i < int(data.cells.size()); ++i) d.cells.push_back(cellDescription(data, i, aux));
    for (auto const& p : data.particles) d.particles.push_back(particleDescription(p));
    return d;
}

// scanAndCreateClusterDescription (DescriptionConverter.cpp:321-360): breadth first over bonds
ClusteredDataDescription DescriptionConverter::toClusteredDataDescription(SimulationData const& data,
                                                                           AuxStore const& aux) {
    ClusteredDataDescription result;
    const int n = int(data.cells.size());
    std::vector<uint8_t> seen(size_t(n), 0);
    std::vector<int> queue;
    for (int start = 0; start < n; ++start) {
        if (seen[size_t(start)]) continue;
        ClusterDescription cluster;
        queue.assign(1, start);
        seen[size_t(start)] = 1;
        for (size_t q = 0; q < queue.size(); ++q) {
            const int i = queue[q];
            cluster.cells.push_back(cellDescription(data, i, aux));
            const alien_cell_to& c = data.cells[size_t(i)];
            for (int b = 0; b < c.numConnections && b < ALIEN_MAX_CELL_BONDS; ++b) {
                const int j = c.connections[b].cellIndex;
                if (j >= 0 && j < n && !seen[size_t(j)]) {
                    seen[size_t(j)] = 1;
                    queue.push_back(j);
                }
            }
        }
        result.clusters.push_back(std::move(cluster));
    }
    for (auto const& p : data.particles) result.particles.push_back(particleDescription(p));
    return result;
}

}  // namespace alien
