// pymodule.cpp — Python binding (pybind11) of the C++ facade, module alien_amd._facade.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <sstream>

#include "simulation_facade.hpp"

namespace py = pybind11;
using alien::SimulationFacade;

namespace {
template <class T> std::vector<T> fromArray(py::array a) {
    py::buffer_info bi = a.request();
    if (bi.ndim != 1 || size_t(bi.itemsize) != sizeof(T))
        throw std::invalid_argument("array does not have the transfer-object layout");
    std::vector<T> v(size_t(bi.shape[0]));
    if (!v.empty()) std::memcpy(v.data(), bi.ptr, v.size() * sizeof(T));
    return v;
}
template <class T> py::bytes toBytes(const std::vector<T>& v) {
    return py::bytes(reinterpret_cast<const char*>(v.data()), v.size() * sizeof(T));
}
alien_params paramsFromBytes(py::bytes b) {
    std::string s = b;
    if (s.size() != sizeof(alien_params)) throw std::invalid_argument("params must be an alien_params blob");
    alien_params p;
    std::memcpy(&p, s.data(), sizeof(p));
    return p;
}
}  // namespace

PYBIND11_MODULE(_facade, m) {
    m.doc() = "C++ SimulationFacade (physics subset of ALIEN's _SimulationFacade) over the HIP C ABI";
    using namespace alien;
    // data descriptions (Descriptions.h) and their conversion (host only)
    py::class_<RealVector2D>(m, "RealVector2D")
        .def(py::init([](float x, float y) { return RealVector2D{x, y}; }), py::arg("x") = 0.0f, py::arg("y") = 0.0f)
        .def_readwrite("x", &RealVector2D::x)
        .def_readwrite("y", &RealVector2D::y);
    py::class_<ConnectionDescription>(m, "ConnectionDescription")
        .def(py::init<>())
        .def_readwrite("cellId", &ConnectionDescription::cellId)
        .def_readwrite("distance", &ConnectionDescription::distance)
        .def_readwrite("angleFromPrevious", &ConnectionDescription::angleFromPrevious);
    py::class_<CellMetadataDescription>(m, "CellMetadataDescription")
        .def(py::init<>())
        .def_readwrite("name", &CellMetadataDescription::name)
        .def_readwrite("description", &CellMetadataDescription::description);
    py::class_<CellFunctionDescription>(m, "CellFunctionDescription")
        .def(py::init<>())
        .def_readwrite("type", &CellFunctionDescription::type)
        .def_readwrite("state", &CellFunctionDescription::state)
        .def_readwrite("genome", &CellFunctionDescription::genome);
    py::class_<CellDescription>(m, "CellDescription")
        .def(py::init<>())
        .def_readwrite("id", &CellDescription::id)
        .def_readwrite("connections", &CellDescription::connections)
        .def_readwrite("pos", &CellDescription::pos)
        .def_readwrite("vel", &CellDescription::vel)
        .def_readwrite("energy", &CellDescription::energy)
        .def_readwrite("stiffness", &CellDescription::stiffness)
        .def_readwrite("color", &CellDescription::color)
        .def_readwrite("maxConnections", &CellDescription::maxConnections)
        .def_readwrite("barrier", &CellDescription::barrier)
        .def_readwrite("age", &CellDescription::age)
        .def_readwrite("livingState", &CellDescription::livingState)
        .def_readwrite("creatureId", &CellDescription::creatureId)
        .def_readwrite("mutationId", &CellDescription::mutationId)
        .def_readwrite("ancestorMutationId", &CellDescription::ancestorMutationId)
        .def_readwrite("genomeComplexity", &CellDescription::genomeComplexity)
        .def_readwrite("activationTime", &CellDescription::activationTime)
        .def_readwrite("cellFunction", &CellDescription::cellFunction)
        .def_readwrite("metadata", &CellDescription::metadata);
    py::class_<ClusterDescription>(m, "ClusterDescription")
        .def(py::init<>())
        .def_readwrite("cells", &ClusterDescription::cells);
    py::class_<ParticleDescription>(m, "ParticleDescription")
        .def(py::init<>())
        .def_readwrite("id", &ParticleDescription::id)
        .def_readwrite("pos", &ParticleDescription::pos)
        .def_readwrite("vel", &ParticleDescription::vel)
        .def_readwrite("energy", &ParticleDescription::energy)
        .def_readwrite("color", &ParticleDescription::color);
    py::class_<ClusteredDataDescription>(m, "ClusteredDataDescription")
        .def(py::init<>())
        .def_readwrite("clusters", &ClusteredDataDescription::clusters)
        .def_readwrite("particles", &ClusteredDataDescription::particles);
    py::class_<DataDescription>(m, "DataDescription")
        .def(py::init<>())
        .def(py::init<ClusteredDataDescription const&>())
        .def_readwrite("cells", &DataDescription::cells)
        .def_readwrite("particles", &DataDescription::particles);
    // conversions with a fresh store of pass-through data (TO = transfer-object bytes)
    m.def("dataToTO", [](DataDescription const& d) {
        SimulationData to;
        AuxStore aux;
        DescriptionConverter::toTO(d, to, aux);
        return py::make_tuple(toBytes(to.cells), toBytes(to.particles));
    });
    m.def("clusteredToTO", [](ClusteredDataDescription const& d) {
        SimulationData to;
        AuxStore aux;
        DescriptionConverter::toTO(d, to, aux);
        return py::make_tuple(toBytes(to.cells), toBytes(to.particles));
    });
    m.def("toDataDescription", [](py::array cells, py::array particles) {
        SimulationData to;
        to.cells = fromArray<alien_cell_to>(cells);
        to.particles = fromArray<alien_particle_to>(particles);
        return DescriptionConverter::toDataDescription(to, AuxStore{});
    });
    m.def("toClusteredDataDescription", [](py::array cells, py::array particles) {
        SimulationData to;
        to.cells = fromArray<alien_cell_to>(cells);
        to.particles = fromArray<alien_particle_to>(particles);
        return DescriptionConverter::toClusteredDataDescription(to, AuxStore{});
    });
    // description -> transfer objects -> clustered description, pass-through data kept by id
    m.def("roundTripClustered", [](DataDescription const& d) {
        SimulationData to;
        AuxStore aux;
        DescriptionConverter::toTO(d, to, aux);
        return DescriptionConverter::toClusteredDataDescription(to, aux);
    });
    // statistics history (StatisticsHistory.h, DataPointCollection.h) and its .statistics.csv form
    py::class_<DataPoint>(m, "DataPoint")
        .def(py::init<>())
        .def_property(
            "values", [](DataPoint const& d) { return std::vector<double>(d.values, d.values + kStatColors); },
            [](DataPoint& d, std::vector<double> const& v) {
                if (v.size() != size_t(kStatColors)) throw std::invalid_argument("seven values, one per colour");
                for (int i = 0; i < kStatColors; ++i) d.values[i] = v[size_t(i)];
            })
        .def_readwrite("summedValues", &DataPoint::summedValues);
    py::class_<DataPointCollection>(m, "DataPointCollection")
        .def(py::init<>())
        .def_readwrite("time", &DataPointCollection::time)
        .def_readwrite("systemClock", &DataPointCollection::systemClock)
        .def_readwrite("numCells", &DataPointCollection::numCells)
        .def_readwrite("numSelfReplicators", &DataPointCollection::numSelfReplicators)
        .def_readwrite("numColonies", &DataPointCollection::numColonies)
        .def_readwrite("numViruses", &DataPointCollection::numViruses)
        .def_readwrite("numFreeCells", &DataPointCollection::numFreeCells)
        .def_readwrite("numParticles", &DataPointCollection::numParticles)
        .def_readwrite("averageGenomeCells", &DataPointCollection::averageGenomeCells)
        .def_readwrite("averageGenomeComplexity", &DataPointCollection::averageGenomeComplexity)
        .def_readwrite("varianceGenomeComplexity", &DataPointCollection::varianceGenomeComplexity)
        .def_readwrite("maxGenomeComplexityOfColonies", &DataPointCollection::maxGenomeComplexityOfColonies)
        .def_readwrite("totalEnergy", &DataPointCollection::totalEnergy)
        .def_readwrite("numCreatedCells", &DataPointCollection::numCreatedCells)
        .def_readwrite("numAttacks", &DataPointCollection::numAttacks)
        .def_readwrite("numMuscleActivities", &DataPointCollection::numMuscleActivities)
        .def_readwrite("numDefenderActivities", &DataPointCollection::numDefenderActivities)
        .def_readwrite("numTransmitterActivities", &DataPointCollection::numTransmitterActivities)
        .def_readwrite("numDetonations", &DataPointCollection::numDetonations);
    m.def("serializeStatistics", [](StatisticsHistoryData const& h) {
        std::ostringstream out;
        serializeStatistics(h, out);
        return out.str();
    });
    m.def("deserializeStatistics", [](std::string const& text) {
        std::istringstream in(text);
        return deserializeStatistics(in);
    });
    // the history sampling alone, fed with given per-colour statistics (test hook of the host logic)
    py::class_<StatisticsHistory>(m, "StatisticsHistory")
        .def(py::init<>())
        .def("addDataPoint",
             [](StatisticsHistory& h, std::vector<int> cells, std::vector<int> freeCells, std::vector<int> particles,
                std::vector<double> energy, uint64_t timestep, double clock) {
                 if (cells.size() != 7 || freeCells.size() != 7 || particles.size() != 7 || energy.size() != 7)
                     throw std::invalid_argument("seven values per colour");
                 TimestepStatistics raw;
                 for (int i = 0; i < 7; ++i) {
                     raw.device.numCells[i] = cells[size_t(i)];
                     raw.numFreeCells[i] = freeCells[size_t(i)];
                     raw.device.numParticles[i] = particles[size_t(i)];
                     raw.device.totalEnergy[i] = energy[size_t(i)];
                 }
                 raw.device.timestep = timestep;
                 h.addDataPoint(raw, timestep, clock);
             })
        .def("resetTime", &StatisticsHistory::resetTime)
        .def("rewrite", &StatisticsHistory::rewrite)
        .def("data", &StatisticsHistory::data)
        .def("longtermTimestepDelta", &StatisticsHistory::longtermTimestepDelta);
    py::register_exception<alien::MemoryAllocationException>(m, "MemoryAllocationException", PyExc_MemoryError);
    py::class_<SimulationFacade>(m, "SimulationFacade")
        .def(py::init<>())
        .def("newSimulation",
             [](SimulationFacade& f, uint64_t ts, int sx, int sy, py::bytes params) {
                 alien::GeneralSettings g;
                 g.worldSizeX = sx;
                 g.worldSizeY = sy;
                 const alien_params p = paramsFromBytes(params);  // reads a Python object: GIL held
                 py::gil_scoped_release rel;
                 f.newSimulation(ts, g, p);
             })
        .def("setSimulationData",
             [](SimulationFacade& f, py::array cells, py::array particles) {
                 alien::SimulationData d;
                 d.cells = fromArray<alien_cell_to>(cells);
                 d.particles = fromArray<alien_particle_to>(particles);
                 py::gil_scoped_release rel;
                 f.setSimulationTO(d);
             })
        .def("getSimulationData",
             [](SimulationFacade& f) {
                 alien::SimulationData d;
                 {
                     py::gil_scoped_release rel;
                     d = f.getSimulationTO();
                 }
                 return py::make_tuple(toBytes(d.cells), toBytes(d.particles));
             })
        .def("calcTimesteps",
             [](SimulationFacade& f, uint64_t n) {
                 py::gil_scoped_release rel;
                 f.calcTimesteps(n);
             })
        .def("runSimulation", &SimulationFacade::runSimulation)
        .def("pauseSimulation", &SimulationFacade::pauseSimulation, py::call_guard<py::gil_scoped_release>())
        .def("isSimulationRunning", &SimulationFacade::isSimulationRunning)
        .def("closeSimulation", &SimulationFacade::closeSimulation, py::call_guard<py::gil_scoped_release>())
        .def("clear", &SimulationFacade::clear, py::call_guard<py::gil_scoped_release>())
        .def("getCurrentTimestep", &SimulationFacade::getCurrentTimestep, py::call_guard<py::gil_scoped_release>())
        .def("setCurrentTimestep", &SimulationFacade::setCurrentTimestep, py::call_guard<py::gil_scoped_release>())
        .def("getSimulationParameters",
             [](SimulationFacade& f) {
                 alien_params p = f.getSimulationParameters();
                 return py::bytes(reinterpret_cast<const char*>(&p), sizeof(p));
             })
        .def("setSimulationParameters",
             [](SimulationFacade& f, py::bytes b) {
                 alien_params p = paramsFromBytes(b);
                 py::gil_scoped_release rel;
                 f.setSimulationParameters(p);
             })
        .def("getWorldSize",
             [](SimulationFacade& f) {
                 auto s = f.getWorldSize();
                 return py::make_tuple(s.x, s.y);
             })
        .def("getGpuName", &SimulationFacade::getGpuName)
        .def("getTps", &SimulationFacade::getTps)
        .def("getStatisticsHistory", &SimulationFacade::getStatisticsHistory)
        .def("setStatisticsHistory", &SimulationFacade::setStatisticsHistory, py::call_guard<py::gil_scoped_release>())
        .def("getStatistics",
             [](SimulationFacade& f) {
                 alien_statistics s;
                 {
                     py::gil_scoped_release rel;
                     s = f.getStatistics();
                 }
                 py::dict d;
                 py::list nc, np_, e;
                 for (int k = 0; k < 7; ++k) {
                     nc.append(s.numCells[k]);
                     np_.append(s.numParticles[k]);
                     e.append(s.totalEnergy[k]);
                 }
                 d["numCells"] = nc;
                 d["numParticles"] = np_;
                 d["totalEnergy"] = e;
                 d["timestep"] = s.timestep;
                 return d;
             })
        .def("getRawStatistics",
             [](SimulationFacade& f) {
                 alien_totals t;
                 {
                     py::gil_scoped_release rel;
                     t = f.getRawStatistics();
                 }
                 py::dict d;
                 d["cellEnergy"] = t.cellEnergy;
                 d["particleEnergy"] = t.particleEnergy;
                 d["momentum"] = py::make_tuple(t.momentum[0], t.momentum[1]);
                 d["kinetic"] = t.kinetic;
                 d["numCells"] = t.numCells;
                 d["numParticles"] = t.numParticles;
                 return d;
             })
        .def("setDataDescription",
             [](SimulationFacade& f, DataDescription const& d) {
                 py::gil_scoped_release rel;
                 f.setSimulationData(d);
             })
        .def("getDataDescription",
             [](SimulationFacade& f) {
                 py::gil_scoped_release rel;
                 return f.getSimulationData();
             })
        .def("setClusteredSimulationData",
             [](SimulationFacade& f, ClusteredDataDescription const& d) {
                 py::gil_scoped_release rel;
                 f.setClusteredSimulationData(d);
             })
        .def("getClusteredSimulationData",
             [](SimulationFacade& f) {
                 py::gil_scoped_release rel;
                 return f.getClusteredSimulationData();
             })
        .def("addAndSelectSimulationData",
             [](SimulationFacade& f, DataDescription const& d) {
                 py::gil_scoped_release rel;
                 f.addAndSelectSimulationData(d);
             })
        .def("testOnly_failNextWorkerStep", &SimulationFacade::testOnly_failNextWorkerStep)
        .def("testOnly_setGuardTimeoutMs",
             [](SimulationFacade& f, int ms) { f.testOnly_setGuardTimeout(std::chrono::milliseconds(ms)); });
}
