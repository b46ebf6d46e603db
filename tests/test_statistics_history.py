"""Statistics history of the facade and its .statistics.csv form (SerializerService.cpp:1212-1484,
StatisticsService.cu:12-110, StatisticsConverterService.cpp:78-140), host only.

The file layout is pinned by ALIEN's own `resources/autosave.statistics.csv` (tests/golden/), a header
written by an earlier ALIEN whose fifth column group is "Cell connections" where the current writer
has "Free cells"; the sampling rules are pinned by hand-computed sequences."""
import os

import pytest

F = pytest.importorskip("alien_amd._facade")

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "autosave.statistics.csv")
Z7 = [0] * 7


def history(points):
    """A StatisticsHistory fed with (timestep, cells of colour 0) points."""
    h = F.StatisticsHistory()
    for t, n in points:
        h.addDataPoint([n] + [0] * 6, Z7, Z7, [0.0] * 7, t, 0.0)
    return h


def column_groups(header):
    """(name, entries) per column of a header row, as the reference's reader groups them."""
    groups = []
    for entry in header.rstrip("\n").split(","):
        name = entry.lstrip().split(" (")[0]
        if groups and groups[-1][0] == name:
            groups[-1][1] += 1
        else:
            groups.append([name, 1])
    return groups


def test_header_has_the_reference_layout():
    fixture = open(FIXTURE).read()
    ours = F.serializeStatistics([])
    assert ours.endswith("\n") and ours.count("\n") == 1
    # the fixture's 22 columns, with the renamed one, are the first 22 of ours, entry for entry;
    # the older file separates every entry with ", ", the current writer (SerializerService.cpp:
    # 1412-1418) the colours of a column with "," and the columns with ", "
    old = column_groups(fixture.replace("Cell connections", "Free cells"))
    new = column_groups(ours)
    assert len(old) == 22 and new[:22] == old
    assert [g[0] for g in new[22:]] == ["Colonies", "Genome complexity average", "Genome complexity Maximum",
                                        "Genome complexity variance", "System clock"]
    assert all(g[1] == 8 for g in new[1:-1]) and new[0] == ["Time step", 1] and new[-1] == ["System clock", 1]
    assert ours.startswith("Time step, Cells (color 0),Cells (color 1),")
    assert ",Cells (color 6),Cells (accumulated), Self-replicators (color 0)," in ours
    assert [e.strip() for e in fixture.rstrip("\n").split(",")][:10] == [
        e.strip() for e in ours.split(",")][:10]


def test_reading_the_reference_header_gives_an_empty_history():
    assert F.deserializeStatistics(open(FIXTURE).read()) == []


def test_columns_are_read_by_name():
    header = open(FIXTURE).read().rstrip("\n")
    n = len(header.split(","))
    row = ",".join(str(float(k)) for k in range(n))
    (p,) = F.deserializeStatistics(header + "\n" + row + "\n")
    assert p.time == 0.0
    assert p.numCells.values == [1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0] and p.numCells.summedValues == 8.0
    assert p.numViruses.values[0] == 17.0
    # "Cell connections" (entries 25..32) is not a column of the current format: skipped
    assert p.numFreeCells.values == [0.0] * 7 and p.numFreeCells.summedValues == 0.0
    assert p.numParticles.values[0] == 33.0 and p.numParticles.summedValues == 40.0
    assert p.totalEnergy.values[6] == 55.0
    # columns 11 and 12 hold the defender and the transmitter activities, as in the reference
    assert p.numDefenderActivities.values[0] == 81.0
    assert p.numTransmitterActivities.values[0] == 89.0
    assert p.systemClock == 0.0  # absent from the older file


def test_round_trip_and_number_format():
    h = F.StatisticsHistory()
    h.addDataPoint([3, 0, 1, 0, 0, 0, 2], [1, 0, 0, 0, 0, 0, 0], [0, 5, 0, 0, 0, 0, 0],
                   [12.5, 0.25, 4e-10, 0.0, 0.0, 0.0, 1234.5], 7, 1700000000.0)
    text = F.serializeStatistics(h.data())
    row = text.split("\n")[1].split(",")
    assert row[0] == "7.000000000"
    assert row[1:9] == ["3.000000000", "0.000000000", "1.000000000", "0.000000000", "0.000000000",
                        "0.000000000", "2.000000000", "6.000000000"]
    assert row[-1] == "1700000000.000000000"
    (p,) = F.deserializeStatistics(text)
    assert p.numCells.values == [3.0, 0.0, 1.0, 0.0, 0.0, 0.0, 2.0] and p.numCells.summedValues == 6.0
    assert p.numFreeCells.values[0] == 1.0 and p.numParticles.summedValues == 5.0
    assert p.totalEnergy.values[0] == 12.5 and p.totalEnergy.values[6] == 1234.5
    assert p.totalEnergy.values[2] == 0.0  # 4e-10 written with nine decimals
    assert F.serializeStatistics(F.deserializeStatistics(text)) == text


def test_sampling_averages_since_the_last_sample():
    # t=0 starts the history; t=1..11 accumulate until 11 - 0 > 10 (the long-term delta), and the
    # sample is their average (time 66/11 = 6, cells 6); t=12..17 likewise until 17 - 6 > 10
    h = history([(t, t) for t in range(18)])
    d = h.data()
    assert [p.time for p in d] == [0.0, 6.0, 14.5]
    assert [p.numCells.values[0] for p in d] == [0.0, 6.0, 14.5]
    assert d[1].numCells.summedValues == 6.0


def test_a_point_before_the_last_sample_starts_a_new_history():
    # samples at 0, 6, 14.5, 21.5 (18..25) and 29 (26..32); 33..39 are pending (39 - 29 is not
    # above 10).  A point at timestep 3 clears the history but not the pending points, as in the
    # reference: the new first sample averages all eight (time 255 / 8, cells (7 + 5) / 8).
    h = history([(t, 1) for t in range(40)])
    assert [p.time for p in h.data()] == [0.0, 6.0, 14.5, 21.5, 29.0]
    h.addDataPoint([5] + [0] * 6, Z7, Z7, [0.0] * 7, 3, 0.0)
    assert [p.time for p in h.data()] == [31.875]
    assert h.data()[0].numCells.values[0] == 1.5


def test_history_is_halved_beyond_1000_samples():
    # one sample every 11 timesteps (each point is more than 10 after the last sample)
    h = history([(11 * k, k) for k in range(1000)])
    assert len(h.data()) == 1000 and h.longtermTimestepDelta() == 10.0
    h.addDataPoint([1000] + [0] * 6, Z7, Z7, [0.0] * 7, 11 * 1000, 0.0)
    d = h.data()
    # 1001 samples: pairs (0,1) .. (998,999) averaged at the first one's time, then the last
    assert len(d) == 501 and h.longtermTimestepDelta() == 20.0
    assert d[0].time == 0.0 and d[0].numCells.values[0] == 0.5
    assert d[499].time == 11 * 998.0 and d[499].numCells.values[0] == 998.5
    assert d[500].time == 11 * 1000.0 and d[500].numCells.values[0] == 1000.0


def test_rewrite_and_reset_time():
    h = history([(11 * k, k) for k in range(5)])
    data = h.data()
    h2 = F.StatisticsHistory()
    h2.rewrite(data, 1000)
    assert h2.longtermTimestepDelta() == (1000 - 0) / 5
    h2.rewrite(data, 20)
    assert h2.longtermTimestepDelta() == 10.0  # at least the default
    # the next point after a rewrite reuses the last sample with the new time
    h2.addDataPoint([99] + [0] * 6, Z7, Z7, [0.0] * 7, 60, 0.0)
    assert h2.data()[-1].time == 60.0 and h2.data()[-1].numCells.values[0] == 4.0
    # reset to timestep 22: samples at or after it go, the delta scales by 22 / 60 (not below 10)
    h2.resetTime(22)
    assert [p.time for p in h2.data()] == [0.0, 11.0]
    assert h2.longtermTimestepDelta() == 10.0
