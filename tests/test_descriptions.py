"""Data descriptions of the facade and their conversion to and from transfer objects
(DescriptionConverter.cpp:132-711), host only: bonds by id <-> by index, cells and particles without
an id, invalid energies, bonds to cells that are absent, clustering by bonds, pass-through data."""
import numpy as np
import pytest

from alien_amd.abi import CELL_DTYPE, PARTICLE_DTYPE

F = pytest.importorskip("alien_amd._facade")


def cell(cid, x, y, bonds=(), energy=100.0, **kw):
    c = F.CellDescription()
    c.id = cid
    c.pos = F.RealVector2D(x, y)
    c.energy = energy
    c.maxConnections = 6
    conns = []
    for other, dist, ang in bonds:
        cd = F.ConnectionDescription()
        cd.cellId, cd.distance, cd.angleFromPrevious = other, dist, ang
        conns.append(cd)
    c.connections = conns
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def particle(pid, x, y, e):
    p = F.ParticleDescription()
    p.id, p.pos, p.energy = pid, F.RealVector2D(x, y), e
    return p


def to_arrays(pair):
    cb, pb = pair
    return np.frombuffer(cb, dtype=CELL_DTYPE), np.frombuffer(pb, dtype=PARTICLE_DTYPE)


def chain_data():
    d = F.DataDescription()
    d.cells = [
        cell(10, 1.0, 1.0, [(11, 1.0, 360.0)]),
        cell(11, 2.0, 1.0, [(10, 1.0, 180.0), (12, 1.0, 180.0)]),
        cell(12, 3.0, 1.0, [(11, 1.0, 360.0)]),
        cell(20, 9.0, 9.0),
    ]
    d.particles = [particle(30, 5.0, 5.0, 2.5)]
    return d


def test_bonds_by_id_become_bonds_by_index_and_back():
    cells, parts = to_arrays(F.dataToTO(chain_data()))
    assert list(cells["id"]) == [10, 11, 12, 20]
    assert list(cells["numConnections"]) == [1, 2, 1, 0]
    assert cells["connections"][1][0]["cellIndex"] == 0 and cells["connections"][1][1]["cellIndex"] == 2
    assert cells["connections"][0][0]["cellIndex"] == 1 and cells["connections"][2][0]["cellIndex"] == 1
    assert cells["connections"][1][1]["angleFromPrevious"] == 180.0
    assert len(parts) == 1 and parts["id"][0] == 30 and parts["energy"][0] == 2.5
    back = F.toDataDescription(cells, parts)
    assert [c.id for c in back.cells] == [10, 11, 12, 20]
    assert [[b.cellId for b in c.connections] for c in back.cells] == [[11], [10, 12], [11], []]
    assert back.cells[1].pos.x == 2.0 and back.particles[0].energy == 2.5


def test_absent_partner_is_dropped_and_its_angle_moves_on():
    d = F.DataDescription()
    # bonds 0 (absent) and 12 then 0 (absent): the first absent angle goes to the next kept bond,
    # the trailing one to the first kept bond
    d.cells = [cell(11, 0.0, 0.0, [(0, 1.0, 40.0), (12, 1.0, 100.0), (13, 1.0, 150.0), (0, 1.0, 70.0)]),
               cell(12, 1.0, 0.0, [(11, 1.0, 360.0)]), cell(13, 0.0, 1.0, [(11, 1.0, 360.0)])]
    cells, _ = to_arrays(F.dataToTO(d))
    assert cells["numConnections"][0] == 2
    angles = [cells["connections"][0][b]["angleFromPrevious"] for b in range(2)]
    assert angles == [140.0 + 70.0, 150.0]
    assert sum(angles) == 360.0


def test_bond_to_a_missing_cell_and_too_many_bonds_are_rejected():
    d = F.DataDescription()
    d.cells = [cell(1, 0.0, 0.0, [(99, 1.0, 360.0)])]
    with pytest.raises(ValueError, match="not in the data"):
        F.dataToTO(d)
    d.cells = [cell(1, 0.0, 0.0, [(k, 1.0, 360.0 / 7) for k in range(2, 9)])] + [cell(k, 1.0, 0.0) for k in range(2, 9)]
    with pytest.raises(ValueError, match="more than 6"):
        F.dataToTO(d)


def test_missing_ids_are_generated_and_invalid_energies_cleared():
    d = F.DataDescription()
    d.cells = [cell(0, 0.0, 0.0, energy=float("nan")), cell(0, 1.0, 0.0, energy=-3.0), cell(5, 2.0, 0.0, energy=2e12)]
    d.particles = [particle(0, 1.0, 1.0, 1.0)]
    cells, parts = to_arrays(F.dataToTO(d))
    ids = [int(x) for x in cells["id"]] + [int(parts["id"][0])]
    assert ids[2] == 5
    generated = [ids[0], ids[1], ids[3]]
    assert all(i >> 48 == 1 for i in generated) and len(set(generated)) == 3
    assert list(cells["energy"]) == [0.0, 0.0, 0.0]


def test_clusters_are_the_connected_components():
    d = F.DataDescription()
    # two bodies and a lone cell, interleaved in the input
    d.cells = [
        cell(1, 0.0, 0.0, [(3, 1.0, 360.0)]),
        cell(2, 5.0, 5.0, [(4, 1.0, 360.0)]),
        cell(3, 1.0, 0.0, [(1, 1.0, 180.0), (5, 1.0, 180.0)]),
        cell(4, 6.0, 5.0, [(2, 1.0, 360.0)]),
        cell(5, 2.0, 0.0, [(3, 1.0, 360.0)]),
        cell(6, 9.0, 9.0),
    ]
    cells, parts = to_arrays(F.dataToTO(d))
    cl = F.toClusteredDataDescription(cells, parts)
    got = [[c.id for c in k.cells] for k in cl.clusters]
    # ordered by first cell in transfer order, breadth first in bond order
    assert got == [[1, 3, 5], [2, 4], [6]]
    flat = F.DataDescription(cl)
    assert [c.id for c in flat.cells] == [1, 3, 5, 2, 4, 6]
    # clustered data converts back to the same bonds
    cells2, _ = to_arrays(F.clusteredToTO(cl))
    by_id = {int(c["id"]): c for c in cells2}
    ids2 = [int(x) for x in cells2["id"]]
    for c in cells:
        c2 = by_id[int(c["id"])]
        assert c2["numConnections"] == c["numConnections"]
        for b in range(int(c["numConnections"])):
            assert ids2[c2["connections"][b]["cellIndex"]] == cells["id"][c["connections"][b]["cellIndex"]]
            assert c2["connections"][b]["angleFromPrevious"] == c["connections"][b]["angleFromPrevious"]


def test_pass_through_data_survives_the_round_trip():
    d = chain_data()
    d.cells[1].metadata.name = "spine"
    d.cells[1].metadata.description = "middle cell"
    fn = F.CellFunctionDescription()
    fn.type = 4
    fn.state = [1, 2, 3]
    fn.genome = list(range(40))
    d.cells[1].cellFunction = fn
    d.cells[2].creatureId = 77
    cl = F.roundTripClustered(d)
    c11 = [c for k in cl.clusters for c in k.cells if c.id == 11][0]
    assert c11.metadata.name == "spine" and c11.metadata.description == "middle cell"
    assert c11.cellFunction.type == 4 and list(c11.cellFunction.state) == [1, 2, 3]
    assert list(c11.cellFunction.genome) == list(range(40))
    assert [c.creatureId for k in cl.clusters for c in k.cells if c.id == 12] == [77]
