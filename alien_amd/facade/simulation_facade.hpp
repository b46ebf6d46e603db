// simulation_facade.hpp — C++ host engine with ALIEN's facade surface for the physics path.
//
// Mirrors the physics subset of _SimulationFacade (source/EngineInterface/SimulationFacade.h:13-110)
// and the threading model of _SimulationFacadeImpl / EngineWorker
// (source/EngineImpl/SimulationFacadeImpl.cpp, EngineWorker.cpp): one worker thread owns the
// device world; API calls take access through a guard that parks the worker (timeout 7 s,
// EngineWorker.cpp:608-632); an exception thrown on the worker thread is rethrown on the next
// access as "GPU worker thread is in an invalid state" (EngineWorker.cpp:645-651).  Device work
// goes only through the C ABI of include/alien_hip.h.
#pragma once

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <exception>
#include <mutex>
#include <optional>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/alien_hip.h"
#include "descriptions.hpp"
#include "statistics_history.hpp"

namespace alien {

// GeneralSettings (source/EngineInterface/GeneralSettings.h)
struct GeneralSettings {
    int worldSizeX = 1000;
    int worldSizeY = 1000;
};

struct IntVector2D {
    int x = 0, y = 0;
};

// CudaMemoryAllocationException analog (Base/Exceptions.h:15-20)
struct MemoryAllocationException : std::runtime_error {
    using std::runtime_error::runtime_error;
};

class SimulationFacade {
public:
    SimulationFacade();
    ~SimulationFacade();
    SimulationFacade(const SimulationFacade&) = delete;
    SimulationFacade& operator=(const SimulationFacade&) = delete;

    void newSimulation(uint64_t timestep, GeneralSettings const& generalSettings, alien_params const& parameters);
    void clear();
    std::string getGpuName() const;

    // getSimulationData / getClusteredSimulationData / setSimulationData /
    // setClusteredSimulationData / addAndSelectSimulationData (SimulationFacade.h:39-47), the
    // whole world; adding selects nothing (selection is not part of the physics path).  The
    // pass-through data of the cells (cell-function state, genome, metadata) is kept here by id.
    DataDescription getSimulationData();
    ClusteredDataDescription getClusteredSimulationData();
    void setSimulationData(DataDescription const& data);
    void setClusteredSimulationData(ClusteredDataDescription const& data);
    void addAndSelectSimulationData(DataDescription const& data);
    // the same data as transfer objects (DataTO of the device seam)
    SimulationData getSimulationTO();
    void setSimulationTO(SimulationData const& data);

    void calcTimesteps(uint64_t timesteps);
    void runSimulation();
    void pauseSimulation();
    bool isSimulationRunning() const;
    void closeSimulation();

    uint64_t getCurrentTimestep();
    void setCurrentTimestep(uint64_t value);

    alien_params getSimulationParameters() const;
    void setSimulationParameters(alien_params const& parameters);

    GeneralSettings getGeneralSettings() const;
    IntVector2D getWorldSize() const;
    alien_totals getRawStatistics();
    // per-colour counts and energies (the TimestepStatistics colour vectors of getRawStatistics)
    alien_statistics getStatistics();
    float getTps() const;
    // getStatisticsHistory / setStatisticsHistory (SimulationFacade.h:99-100): the history is
    // sampled like _SimulationCudaFacade::updateStatistics (SimulationCudaFacade.cu:129-160,
    // 250-280): after data is set, after calcTimesteps, and every 30 ms while running
    StatisticsHistoryData getStatisticsHistory() const;
    void setStatisticsHistory(StatisticsHistoryData const& data);

    // test hooks: make the worker thread fail on its next step (failure detection path), and set
    // the guard timeout
    void testOnly_failNextWorkerStep(std::string const& message);
    void testOnly_setGuardTimeout(std::chrono::milliseconds timeout);

private:
    class Guard;
    void workerLoop();
    void throwIfWorkerFailed();
    std::optional<std::string> takeInjectedFailure();
    static void check(int rc);

    SimulationData readTO();
    void writeTO(SimulationData const& data);
    // one data point of the statistics history from the device's counts (caller owns the world)
    void updateStatistics();

    alien_world* _world = nullptr;
    AuxStore _aux;  // pass-through data of the cells, by id
    GeneralSettings _settings;
    alien_params _params{};

    std::thread _worker;
    mutable std::mutex _accessMutex;  // serializes API callers (EngineWorkerGuard)
    std::mutex _stateMutex;
    std::condition_variable _stateCv;
    int _accessState = 0;  // 0 worker free, 1 access requested, 2 worker parked
    bool _running = false;
    bool _stop = false;
    std::exception_ptr _workerError;
    std::optional<std::string> _injectedFailure;
    std::chrono::milliseconds _guardTimeout{7000};
    std::atomic<float> _tps{0.0f};

    mutable std::mutex _historyMutex;
    StatisticsHistory _history;
    int32_t _numFreeCells[kStatColors] = {0, 0, 0, 0, 0, 0, 0};  // per colour, constant on this path
    std::optional<std::chrono::steady_clock::time_point> _lastStatisticsUpdate;
};

}  // namespace alien
