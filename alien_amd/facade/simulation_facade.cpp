// simulation_facade.cpp — see simulation_facade.hpp.
#include "simulation_facade.hpp"

#include <cstring>

#include "../../include/alien_hip.h"

namespace alien {

void SimulationFacade::check(int rc) {
    if (rc == ALIEN_OK) return;
    std::string msg = alien_last_error();
    if (rc == ALIEN_E_OOM) throw MemoryAllocationException(msg);
    throw std::runtime_error(msg);
}

// Parks the worker thread for the duration of an API call (EngineWorkerGuard, EngineWorker.cpp:
// 608-632).  Throws if the worker does not answer within the timeout.
class SimulationFacade::Guard {
public:
    explicit Guard(SimulationFacade& f) : _f(f), _lock(f._accessMutex) {
        std::unique_lock<std::mutex> lk(_f._stateMutex);
        _f._accessState = 1;
        _f._stateCv.notify_all();
        if (!_f._stateCv.wait_for(lk, _f._guardTimeout, [&] { return _f._accessState == 2 || !_f._worker.joinable(); })) {
            _f._accessState = 0;
            throw std::runtime_error("GPU worker thread is not reachable.");
        }
    }
    ~Guard() {
        std::lock_guard<std::mutex> lk(_f._stateMutex);
        _f._accessState = 0;
        _f._stateCv.notify_all();
    }

private:
    SimulationFacade& _f;
    std::unique_lock<std::mutex> _lock;
};

SimulationFacade::SimulationFacade() {}

SimulationFacade::~SimulationFacade() { closeSimulation(); }

void SimulationFacade::workerLoop() {
    auto last = std::chrono::steady_clock::now();
    uint64_t stepsSince = 0;
    while (true) {
        {
            std::unique_lock<std::mutex> lk(_stateMutex);
            _stateCv.wait(lk, [&] { return _stop || _accessState == 1 || (_running && !_workerError); });
            if (_stop) return;
            if (_accessState == 1) {
                _accessState = 2;  // parked: the caller has exclusive access
                _stateCv.notify_all();
                // leave the park when the caller released it (0) or the next caller already asked
                // again (1) before this thread woke up
                _stateCv.wait(lk, [&] { return _stop || _accessState != 2; });
                continue;
            }
        }
        // running mode: one timestep at a time (EngineWorker.cpp:450-476)
        try {
            if (auto m = takeInjectedFailure()) throw std::runtime_error(*m);
            check(alien_step(_world, 1));
            ++stepsSince;
            auto now = std::chrono::steady_clock::now();
            if (!_lastStatisticsUpdate || now - *_lastStatisticsUpdate > std::chrono::milliseconds(30)) updateStatistics();
            double dt = std::chrono::duration<double>(now - last).count();
            if (dt > 0.5) {
                _tps = float(double(stepsSince) / dt);
                stepsSince = 0;
                last = now;
            }
        } catch (...) {
            std::lock_guard<std::mutex> lk(_stateMutex);
            _workerError = std::current_exception();
            _running = false;
        }
    }
}

std::optional<std::string> SimulationFacade::takeInjectedFailure() {
    std::lock_guard<std::mutex> lk(_stateMutex);
    auto m = _injectedFailure;
    _injectedFailure.reset();
    return m;
}

void SimulationFacade::throwIfWorkerFailed() {
    std::lock_guard<std::mutex> lk(_stateMutex);
    if (_workerError) throw std::runtime_error("GPU worker thread is in an invalid state.");
}

void SimulationFacade::newSimulation(uint64_t timestep, GeneralSettings const& generalSettings,
                                     alien_params const& parameters) {
    closeSimulation();
    alien_world_desc desc{};
    desc.worldSizeX = generalSettings.worldSizeX;
    desc.worldSizeY = generalSettings.worldSizeY;
    alien_world* w = nullptr;
    check(alien_world_create(&desc, &parameters, &w));
    _world = w;
    check(alien_set_timestep(_world, timestep));
    _settings = generalSettings;
    _params = parameters;
    {
        std::lock_guard<std::mutex> lk(_historyMutex);
        _history = StatisticsHistory{};
    }
    _lastStatisticsUpdate.reset();
    for (int k = 0; k < kStatColors; ++k) _numFreeCells[k] = 0;
    _stop = false;
    _running = false;
    _workerError = nullptr;
    _accessState = 0;
    _worker = std::thread([this] { workerLoop(); });
}

void SimulationFacade::closeSimulation() {
    if (_worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(_stateMutex);
            _stop = true;
        }
        _stateCv.notify_all();
        _worker.join();
    }
    if (_world) {
        alien_world_destroy(_world);
        _world = nullptr;
    }
}

void SimulationFacade::clear() {
    Guard g(*this);
    throwIfWorkerFailed();
    check(alien_clear(_world));
}

std::string SimulationFacade::getGpuName() const {
    char buf[256];
    check(alien_get_device_name(buf, sizeof(buf)));
    return buf;
}

SimulationData SimulationFacade::readTO() {
    SimulationData d;
    uint64_t n = 0, p = 0;
    check(alien_get_counts(_world, &n, &p));
    d.cells.resize(n);
    d.particles.resize(p);
    check(alien_get_data(_world, d.cells.data(), n, &n, d.particles.data(), p, &p));
    d.cells.resize(n);
    d.particles.resize(p);
    return d;
}

void SimulationFacade::writeTO(SimulationData const& data) {
    check(alien_set_data(_world, data.cells.data(), data.cells.size(), data.particles.data(), data.particles.size()));
    {
        std::lock_guard<std::mutex> lk(_stateMutex);
        _workerError = nullptr;  // fresh data makes the engine usable again
    }
    // free cells (StatisticsKernels.cu:19-21: mutation id MutationIdForFreeCell = 1); no kernel of
    // this path changes a cell's colour or mutation id, so the counts hold until the next set
    for (int k = 0; k < kStatColors; ++k) _numFreeCells[k] = 0;
    for (auto const& c : data.cells)
        if (c.mutationId == 1 && c.color < kStatColors) ++_numFreeCells[c.color];
    updateStatistics();
}

void SimulationFacade::updateStatistics() {
    TimestepStatistics raw;
    check(alien_get_statistics(_world, &raw.device));
    for (int k = 0; k < kStatColors; ++k) raw.numFreeCells[k] = _numFreeCells[k];
    const double clock = double(std::chrono::duration_cast<std::chrono::seconds>(
                                    std::chrono::system_clock::now().time_since_epoch())
                                    .count());
    _lastStatisticsUpdate = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> lk(_historyMutex);
    _history.addDataPoint(raw, raw.device.timestep, clock);
}

StatisticsHistoryData SimulationFacade::getStatisticsHistory() const {
    std::lock_guard<std::mutex> lk(_historyMutex);
    return _history.data();
}

void SimulationFacade::setStatisticsHistory(StatisticsHistoryData const& data) {
    Guard g(*this);
    const uint64_t t = alien_get_timestep(_world);
    std::lock_guard<std::mutex> lk(_historyMutex);
    _history.rewrite(data, t);
}

SimulationData SimulationFacade::getSimulationTO() {
    Guard g(*this);
    throwIfWorkerFailed();
    return readTO();
}

void SimulationFacade::setSimulationTO(SimulationData const& data) {
    Guard g(*this);
    writeTO(data);
}

DataDescription SimulationFacade::getSimulationData() {
    Guard g(*this);
    throwIfWorkerFailed();
    return DescriptionConverter::toDataDescription(readTO(), _aux);
}

ClusteredDataDescription SimulationFacade::getClusteredSimulationData() {
    Guard g(*this);
    throwIfWorkerFailed();
    return DescriptionConverter::toClusteredDataDescription(readTO(), _aux);
}

void SimulationFacade::setSimulationData(DataDescription const& data) {
    SimulationData to;
    AuxStore aux;
    DescriptionConverter::toTO(data, to, aux);  // (checked before anything changes)
    Guard g(*this);
    writeTO(to);
    _aux = std::move(aux);
}

void SimulationFacade::setClusteredSimulationData(ClusteredDataDescription const& data) {
    SimulationData to;
    AuxStore aux;
    DescriptionConverter::toTO(data, to, aux);
    Guard g(*this);
    writeTO(to);
    _aux = std::move(aux);
}

// the current data plus the new (whose bonds may refer only to its own cells), set as a whole
void SimulationFacade::addAndSelectSimulationData(DataDescription const& data) {
    SimulationData added;
    AuxStore aux;
    DescriptionConverter::toTO(data, added, aux);
    Guard g(*this);
    throwIfWorkerFailed();
    SimulationData to = readTO();
    const int32_t offset = int32_t(to.cells.size());
    for (auto c : added.cells) {
        for (int b = 0; b < c.numConnections; ++b) c.connections[b].cellIndex += offset;
        to.cells.push_back(c);
    }
    to.particles.insert(to.particles.end(), added.particles.begin(), added.particles.end());
    writeTO(to);
    for (auto& kv : aux) _aux[kv.first] = std::move(kv.second);
}

void SimulationFacade::calcTimesteps(uint64_t timesteps) {
    Guard g(*this);
    throwIfWorkerFailed();
    try {
        if (auto m = takeInjectedFailure()) throw std::runtime_error(*m);
        check(alien_step(_world, timesteps));
        updateStatistics();  // (forceUpdateStatistics of calcTimesteps)
    } catch (...) {
        // a failed step leaves the engine unusable until new data is set (EngineWorker.cpp:472-475)
        std::lock_guard<std::mutex> lk(_stateMutex);
        _workerError = std::current_exception();
        throw;
    }
}

void SimulationFacade::runSimulation() {
    throwIfWorkerFailed();
    std::lock_guard<std::mutex> lk(_stateMutex);
    _running = true;
    _stateCv.notify_all();
}

void SimulationFacade::pauseSimulation() {
    Guard g(*this);
    std::lock_guard<std::mutex> lk(_stateMutex);
    _running = false;
}

bool SimulationFacade::isSimulationRunning() const {
    auto& self = const_cast<SimulationFacade&>(*this);
    std::lock_guard<std::mutex> lk(self._stateMutex);
    return _running;
}

uint64_t SimulationFacade::getCurrentTimestep() {
    Guard g(*this);
    throwIfWorkerFailed();
    return alien_get_timestep(_world);
}

void SimulationFacade::setCurrentTimestep(uint64_t value) {
    Guard g(*this);
    throwIfWorkerFailed();
    check(alien_set_timestep(_world, value));
}

alien_params SimulationFacade::getSimulationParameters() const { return _params; }

void SimulationFacade::setSimulationParameters(alien_params const& parameters) {
    Guard g(*this);
    throwIfWorkerFailed();
    check(alien_set_params(_world, &parameters));
    _params = parameters;
}

GeneralSettings SimulationFacade::getGeneralSettings() const { return _settings; }

IntVector2D SimulationFacade::getWorldSize() const { return {_settings.worldSizeX, _settings.worldSizeY}; }

alien_totals SimulationFacade::getRawStatistics() {
    Guard g(*this);
    throwIfWorkerFailed();
    alien_totals t{};
    check(alien_get_totals(_world, &t));
    return t;
}

alien_statistics SimulationFacade::getStatistics() {
    Guard g(*this);
    throwIfWorkerFailed();
    alien_statistics s{};
    check(alien_get_statistics(_world, &s));
    return s;
}

float SimulationFacade::getTps() const { return _tps.load(); }

void SimulationFacade::testOnly_failNextWorkerStep(std::string const& message) {
    std::lock_guard<std::mutex> lk(_stateMutex);
    _injectedFailure = message;
}

void SimulationFacade::testOnly_setGuardTimeout(std::chrono::milliseconds timeout) { _guardTimeout = timeout; }

}  // namespace alien
