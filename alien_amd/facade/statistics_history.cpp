// statistics_history.cpp — see statistics_history.hpp.
#include "statistics_history.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace alien {

namespace {

constexpr double kNearZero = double(1.0e-4f);  // NEAR_ZERO (Base/Definitions.h:27), a float

template <class T>
DataPoint bySummation(const T (&values)[kStatColors]) {
    DataPoint r;
    r.summedValues = 0;
    for (int i = 0; i < kStatColors; ++i) {
        r.values[i] = double(values[i]);
        r.summedValues += r.values[i];
    }
    return r;
}

// the columns of the file and the fields they hold (SerializerService.cpp:1224-1310; columns 11
// and 12 hold the defender and the transmitter activities, in that order, as in the reference)
struct Column {
    const char* name;
    bool colorDependent;
};
constexpr Column kColumns[] = {{"Time step", false},
                               {"Cells", true},
                               {"Self-replicators", true},
                               {"Viruses", true},
                               {"Free cells", true},
                               {"Energy particles", true},
                               {"Average genome cells", true},
                               {"Total energy", true},
                               {"Created cells", true},
                               {"Attacks", true},
                               {"Muscle activities", true},
                               {"Transmitter activities", true},
                               {"Defender activities", true},
                               {"Injection activities", true},
                               {"Completed injections", true},
                               {"Nerve pulses", true},
                               {"Neuron activities", true},
                               {"Sensor activities", true},
                               {"Sensor matches", true},
                               {"Reconnector creations", true},
                               {"Reconnector deletions", true},
                               {"Detonations", true},
                               {"Colonies", true},
                               {"Genome complexity average", true},
                               {"Genome complexity Maximum", true},
                               {"Genome complexity variance", true},
                               {"System clock", false}};
constexpr int kNumColumns = int(sizeof(kColumns) / sizeof(kColumns[0]));

template <class D>  // DataPointCollection, const or not
auto dataPointOf(int col, D& d) -> decltype(&d.numCells) {
    switch (col) {
        case 1: return &d.numCells;
        case 2: return &d.numSelfReplicators;
        case 3: return &d.numViruses;
        case 4: return &d.numFreeCells;
        case 5: return &d.numParticles;
        case 6: return &d.averageGenomeCells;
        case 7: return &d.totalEnergy;
        case 8: return &d.numCreatedCells;
        case 9: return &d.numAttacks;
        case 10: return &d.numMuscleActivities;
        case 11: return &d.numDefenderActivities;
        case 12: return &d.numTransmitterActivities;
        case 13: return &d.numInjectionActivities;
        case 14: return &d.numCompletedInjections;
        case 15: return &d.numNervePulses;
        case 16: return &d.numNeuronActivities;
        case 17: return &d.numSensorActivities;
        case 18: return &d.numSensorMatches;
        case 19: return &d.numReconnectorCreated;
        case 20: return &d.numReconnectorRemoved;
        case 21: return &d.numDetonations;
        case 22: return &d.numColonies;
        case 23: return &d.averageGenomeComplexity;
        case 24: return &d.maxGenomeComplexityOfColonies;
        case 25: return &d.varianceGenomeComplexity;
        default: return nullptr;
    }
}
template <class D>
auto scalarOf(int col, D& d) -> decltype(&d.time) {
    if (col == 0) return &d.time;
    if (col == kNumColumns - 1) return &d.systemClock;
    return nullptr;
}

std::string toString(double v) {
    std::ostringstream ss;
    ss << std::fixed << std::setprecision(9) << v;
    return ss.str();
}

std::vector<std::string> splitComma(std::string const& line) {
    std::vector<std::string> out;
    size_t start = 0;
    while (true) {
        const size_t p = line.find(',', start);
        out.push_back(line.substr(start, p == std::string::npos ? std::string::npos : p - start));
        if (p == std::string::npos) break;
        start = p + 1;
    }
    return out;
}

// getPrincipalPart (SerializerService.cpp:1387-1396): leading blanks dropped, cut at " ("
std::string principalPart(std::string const& name) {
    size_t b = 0;
    while (b < name.size() && std::isspace(static_cast<unsigned char>(name[b]))) ++b;
    std::string t = name.substr(b);
    const size_t p = t.find(" (");
    return p == std::string::npos ? t : t.substr(0, p);
}

int columnIndex(std::string const& principal) {
    for (int k = 0; k < kNumColumns; ++k)
        if (principal == kColumns[k].name) return k;
    return -1;
}

// std::stod, as the reference reads its entries (throws std::invalid_argument on a non-number)
double parseNumber(std::string const& s) { return std::stod(s); }

}  // namespace

DataPoint DataPoint::operator+(DataPoint const& o) const {
    DataPoint r;
    for (int i = 0; i < kStatColors; ++i) r.values[i] = values[i] + o.values[i];
    r.summedValues = summedValues + o.summedValues;
    return r;
}

DataPoint DataPoint::operator/(double divisor) const {
    DataPoint r;
    for (int i = 0; i < kStatColors; ++i) r.values[i] = values[i] / divisor;
    r.summedValues = summedValues / divisor;
    return r;
}

DataPointCollection DataPointCollection::operator+(DataPointCollection const& o) const {
    DataPointCollection r;
    r.time = time + o.time;
    r.systemClock = systemClock + o.systemClock;
    for (int col = 1; col < kNumColumns - 1; ++col) *dataPointOf(col, r) = *dataPointOf(col, *this) + *dataPointOf(col, o);
    return r;
}

DataPointCollection DataPointCollection::operator/(double divisor) const {
    DataPointCollection r;
    r.time = time / divisor;
    r.systemClock = systemClock / divisor;
    for (int col = 1; col < kNumColumns - 1; ++col) *dataPointOf(col, r) = *dataPointOf(col, *this) / divisor;
    return r;
}

DataPointCollection convertStatistics(TimestepStatistics const& data, double time, double systemClock) {
    DataPointCollection r;
    r.time = time;
    r.systemClock = systemClock;
    r.numCells = bySummation(data.device.numCells);
    r.numFreeCells = bySummation(data.numFreeCells);
    r.numParticles = bySummation(data.device.numParticles);
    r.totalEnergy = bySummation(data.device.totalEnergy);
    // replicators, viruses, colonies, genome statistics: no cell functions on this path; the
    // averages over zero replicators are the summed zeros (getDataPointByAveraging), and the
    // process properties of unchanged accumulated counters are 0 (getDataPointForProcessProperty)
    return r;
}

void StatisticsHistory::addDataPoint(TimestepStatistics const& raw, uint64_t timestep, double systemClock) {
    const double t = double(timestep);
    if (!_data.empty() && _data.back().time > t + kNearZero) _data.clear();

    if (!_haveLast || _data.empty() || t - _data.back().time > _longtermTimestepDelta / 100 * (_numDataPoints + 1)) {
        DataPointCollection point;
        if (!_haveLast && !_data.empty()) {
            // no raw statistics since the history was rewritten: reuse the last entry
            point = _data.back();
            point.time = t;
        } else {
            point = convertStatistics(raw, t, systemClock);
        }
        _haveLast = true;
        _accumulated = _accumulated ? *_accumulated + point : point;
        ++_numDataPoints;
    }

    if (_accumulated && (_data.empty() || t - _data.back().time > _longtermTimestepDelta)) {
        DataPointCollection point = *_accumulated / double(_numDataPoints);
        _numDataPoints = 0;
        _accumulated.reset();
        if (!_data.empty() && std::abs(_data.back().time - t) < kNearZero) _data.pop_back();
        _data.push_back(point);
        if (_data.size() > kMaxSamples) {
            StatisticsHistoryData halved;
            halved.reserve(_data.size() / 2);
            for (size_t i = 0; i < (_data.size() - 1) / 2; ++i) {
                DataPointCollection p = (_data[i * 2] + _data[i * 2 + 1]) / 2.0;
                p.time = _data[i * 2].time;
                halved.push_back(p);
            }
            halved.push_back(_data.back());
            _data.swap(halved);
            _longtermTimestepDelta *= 2.0;
        }
    }
}

void StatisticsHistory::resetTime(uint64_t timestep) {
    if (_data.empty()) return;
    const double prev = _data.back().time;
    if (prev > 0) {
        _longtermTimestepDelta *= double(timestep) / prev;
        if (_longtermTimestepDelta < kDefaultTimestepDelta) _longtermTimestepDelta = kDefaultTimestepDelta;
    } else {
        _longtermTimestepDelta = kDefaultTimestepDelta;
    }
    StatisticsHistoryData kept;
    for (auto const& p : _data)
        if (p.time < double(timestep)) kept.push_back(p);
    _data.swap(kept);
    _accumulated.reset();
    _numDataPoints = 0;
}

void StatisticsHistory::rewrite(StatisticsHistoryData const& data, uint64_t timestep) {
    _accumulated.reset();
    _numDataPoints = 0;
    _haveLast = false;
    _longtermTimestepDelta =
        data.empty() ? kDefaultTimestepDelta
                     : std::max(kDefaultTimestepDelta, (double(timestep) - data.front().time) / double(data.size()));
    _data = data;
}

void serializeStatistics(StatisticsHistoryData const& history, std::ostream& out) {
    for (int col = 0; col < kNumColumns; ++col) {
        if (col != 0) out << ", ";
        if (!kColumns[col].colorDependent) {
            out << kColumns[col].name;
        } else {
            for (int i = 0; i < kStatColors; ++i) {
                if (i != 0) out << ",";
                out << kColumns[col].name << " (color " << i << ")";
            }
            out << "," << kColumns[col].name << " (accumulated)";
        }
    }
    out << "\n";
    for (auto const& d : history) {
        std::string line;
        for (int col = 0; col < kNumColumns; ++col) {
            if (const double* v = scalarOf(col, d)) {
                line += (col ? "," : "") + toString(*v);
            } else {
                DataPoint const& dp = *dataPointOf(col, d);
                for (int i = 0; i < kStatColors; ++i) line += "," + toString(dp.values[i]);
                line += "," + toString(dp.summedValues);
            }
        }
        out << line << "\n";
    }
}

StatisticsHistoryData deserializeStatistics(std::istream& in) {
    StatisticsHistoryData result;
    std::string header;
    if (!std::getline(in, header)) return result;
    // consecutive header entries with the same principal part form one column
    struct Info {
        std::string name;
        int col;
        int size;
    };
    std::vector<Info> infos;
    for (auto const& entry : splitComma(header)) {
        const std::string pp = principalPart(entry);
        if (!infos.empty() && infos.back().name == pp)
            ++infos.back().size;
        else
            infos.push_back({pp, columnIndex(pp), 1});
    }
    std::string line;
    while (std::getline(in, line)) {
        const std::vector<std::string> e = splitComma(line);
        DataPointCollection d;
        size_t start = 0;
        for (auto const& info : infos) {
            if (info.col >= 0) {
                if (double* v = scalarOf(info.col, d)) {
                    if (start < e.size()) *v = parseNumber(e[start]);
                } else {
                    DataPoint& dp = *dataPointOf(info.col, d);
                    for (int i = 0; i < kStatColors; ++i)
                        if (start + i < e.size()) dp.values[i] = parseNumber(e[start + i]);
                    if (start + kStatColors < e.size()) dp.summedValues = parseNumber(e[start + kStatColors]);
                }
            }
            start += size_t(info.size);
        }
        result.push_back(d);
    }
    return result;
}

}  // namespace alien
