// statistics_history.hpp — the statistics history of ALIEN's facade and its .statistics.csv form.
//
// Mirrors DataPoint / DataPointCollection (source/EngineInterface/DataPointCollection.h:5-48,
// DataPointCollection.cpp), the sampling of the history (StatisticsService::addDataPoint /
// resetTime / rewriteHistory, source/EngineGpuKernels/StatisticsService.cu:12-110), the conversion
// of one timestep's raw statistics into a data point (StatisticsConverterService::convert,
// source/EngineInterface/StatisticsConverterService.cpp:78-140) and the CSV reader and writer
// (SerializerService::serializeStatistics / deserializeStatistics,
// source/PersisterInterface/SerializerService.cpp:1212-1484).  The physics path fills the columns
// it has: cells, free cells, energy particles and total energy per colour; the cell-function
// columns (replicators, viruses, genome statistics, activities, colonies) stay 0.  Host-only code.
#pragma once

#include <cstdint>
#include <istream>
#include <optional>
#include <ostream>
#include <string>
#include <vector>

#include "../../include/alien_hip.h"

namespace alien {

constexpr int kStatColors = 7;

struct DataPoint {
    double values[kStatColors] = {0, 0, 0, 0, 0, 0, 0};
    double summedValues = 0;

    DataPoint operator+(DataPoint const& other) const;
    DataPoint operator/(double divisor) const;
};

struct DataPointCollection {
    double time = 0;  // a timestep
    double systemClock = 0;

    DataPoint numCells;
    DataPoint numSelfReplicators;
    DataPoint numColonies;
    DataPoint numViruses;
    DataPoint numFreeCells;
    DataPoint numParticles;
    DataPoint averageGenomeCells;
    DataPoint averageGenomeComplexity;
    DataPoint varianceGenomeComplexity;
    DataPoint maxGenomeComplexityOfColonies;
    DataPoint totalEnergy;

    DataPoint numCreatedCells;
    DataPoint numAttacks;
    DataPoint numMuscleActivities;
    DataPoint numDefenderActivities;
    DataPoint numTransmitterActivities;
    DataPoint numInjectionActivities;
    DataPoint numCompletedInjections;
    DataPoint numNervePulses;
    DataPoint numNeuronActivities;
    DataPoint numSensorActivities;
    DataPoint numSensorMatches;
    DataPoint numReconnectorCreated;
    DataPoint numReconnectorRemoved;
    DataPoint numDetonations;

    DataPointCollection operator+(DataPointCollection const& other) const;
    DataPointCollection operator/(double divisor) const;
};

using StatisticsHistoryData = std::vector<DataPointCollection>;

// One timestep's statistics of the physics path: the per-colour counts and energies of the device
// (alien_statistics) and the free cells per colour (cells with mutation id 1, which keep their
// colour and mutation id on this path).
struct TimestepStatistics {
    alien_statistics device{};
    int32_t numFreeCells[kStatColors] = {0, 0, 0, 0, 0, 0, 0};
};

// StatisticsConverterService::convert for the physics columns; systemClock is the given wall
// clock in seconds since the Unix epoch.
DataPointCollection convertStatistics(TimestepStatistics const& data, double time, double systemClock);

// StatisticsService: the history keeps at most 1000 samples; a sample averages the data points
// taken since the previous one, at least longtermTimestepDelta timesteps apart (10 at first,
// doubled whenever the history is halved by averaging neighbouring samples).
class StatisticsHistory {
public:
    static constexpr double kDefaultTimestepDelta = 10.0;
    static constexpr size_t kMaxSamples = 1000;

    void addDataPoint(TimestepStatistics const& raw, uint64_t timestep, double systemClock);
    void resetTime(uint64_t timestep);
    void rewrite(StatisticsHistoryData const& data, uint64_t timestep);
    StatisticsHistoryData const& data() const { return _data; }
    double longtermTimestepDelta() const { return _longtermTimestepDelta; }

private:
    StatisticsHistoryData _data;
    double _longtermTimestepDelta = kDefaultTimestepDelta;
    std::optional<DataPointCollection> _accumulated;
    int _numDataPoints = 0;
    bool _haveLast = false;
};

// The .statistics.csv text: a header row (per colour-dependent column seven "name (color i)"
// entries and "name (accumulated)"), then one row per sample, numbers in fixed notation with nine
// decimals.  Reading maps columns by name: unknown columns are skipped, missing ones stay 0.
void serializeStatistics(StatisticsHistoryData const& history, std::ostream& out);
StatisticsHistoryData deserializeStatistics(std::istream& in);

}  // namespace alien
