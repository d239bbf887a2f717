"""Periodic boundary pairs on the host side (no GPU): the grid file's NCon
lines round-trip through the writer and reader, the set-up refuses what the
reference would stop on, and the oracle's periodic set-up is the reference's
(the reference's own set-up dumps of the periodic fixtures)."""
import numpy as np
import pytest

from helpers import case_from_fixture, load
from rocketcombustor_amd import deck, nozzle
from rocketcombustor_amd.native import RcError
from rocketcombustor_amd.solver import SteadyCase

oracle = pytest.importorskip("oracle.pyoracle")


@pytest.mark.parametrize("direction", ["x", "y"])
def test_grid_file_round_trip(tmp_path, direction):
    case = nozzle.periodic_channel_case(12, 4, direction=direction)
    m = case.mesh
    p = str(tmp_path / "grid1.in")
    deck.write_grid(p, m.coor, m.cells, m.bedges, m.btype, m.nface, periodic=m.periodic)
    text = open(p).read().split("\n")
    # boundaries are written by type, the periodic ones with type 100
    assert "100 %d" % int((m.btype == 101).sum()) in text
    coor, cells, bedges, btype, nface, periodic = deck.read_grid(p, with_periodic=True)
    assert np.array_equal(cells, m.cells) and nface == m.nface
    assert np.array_equal(coor, m.coor)
    # boundaries come back one per type, in type order, each in its face
    # order; labels are 100 + the boundary's 1-based index in the file
    nplain = len(set(m.btype[m.btype < 100]))
    lab = {101: 101 + nplain, 102: 102 + nplain}
    for t in sorted(set(int(t) for t in m.btype)):
        assert np.array_equal(bedges[btype == lab.get(t, t)], m.bedges[m.btype == t])
    assert len(bedges) == len(m.bedges)
    assert periodic == tuple((lab[a], lab[b], x, y) for a, b, x, y in m.periodic)
    with pytest.raises(ValueError):
        deck.read_grid(p)


def test_case_from_dir_keeps_periodic_pairs(tmp_path):
    case = nozzle.periodic_channel_case(12, 4, direction="y")
    nozzle.write_reference_case(str(tmp_path), case)
    back = nozzle.case_from_dir(str(tmp_path))
    assert len(back.mesh.periodic) == 2
    # the same set-up as the original labels (oracle set-up arrays)
    a, b = oracle.OracleCase(case), oracle.OracleCase(back)
    for n in ("Neighbor", "NeighF"):
        assert np.array_equal(a.get(n), b.get(n))


def test_unpaired_periodic_boundary_is_refused():
    """Only one direction of the pair: the target boundary's faces keep type
    100, on which the reference's flux_lax stops; the set-up refuses them
    (before any device work)."""
    case = nozzle.periodic_channel_case(12, 4, direction="x")
    one_way = case.mesh.periodic[:1]
    bad = nozzle.Case(**{**case.__dict__, "mesh": nozzle.Mesh(case.mesh.coor, case.mesh.cells, case.mesh.bedges,
                                                              case.mesh.btype, case.mesh.nface, periodic=one_way)})
    with pytest.raises(RcError, match="no partner"):
        SteadyCase(bad)


def test_unmatched_translation_is_refused():
    case = nozzle.periodic_channel_case(12, 4, direction="x")
    a, b, x, y = case.mesh.periodic[0]
    wrong = ((a, b, x * 1.001, y), case.mesh.periodic[1])
    bad = nozzle.Case(**{**case.__dict__, "mesh": nozzle.Mesh(case.mesh.coor, case.mesh.cells, case.mesh.bedges,
                                                              case.mesh.btype, case.mesh.nface, periodic=wrong)})
    with pytest.raises(RcError, match="not find target"):
        SteadyCase(bad)


@pytest.mark.parametrize("name", ["n24x6_periodic_x", "n24x6_periodic_y"])
def test_periodic_faces_in_reference_setup(name):
    """The reference's own dumps: a periodic face's neighbour is the owner
    of its partner, it is listed only in its owner's NeighF, and the
    oracle forms the same arrays."""
    d = load(name)
    case = case_from_fixture(d)
    o = oracle.OracleCase(case)
    assert np.array_equal(o.get("Neighbor"), d["ref_Neighbor"])
    assert np.array_equal(o.get("NeighF"), d["ref_NeighF"])
    nb, fn, nf = d["ref_Neighbor"], d["ref_FNode"], d["ref_NeighF"]
    labels = {int(p[0]) for p in d["periodic"]} | {int(p[1]) for p in d["periodic"]}
    pedges = {frozenset(e) for e, t in zip(case.mesh.bedges.tolist(), case.mesh.btype) if int(t) in labels}
    assert pedges
    nper = 0
    for f in range(len(nb)):
        if frozenset((fn[f] - 1).tolist()) not in pedges:
            continue
        nper += 1
        nl, nr = nb[f]
        # a cell neighbour across, listed in its owner's faces only
        assert nr > 0 and nr != nl
        assert f + 1 in nf[nl - 1] and f + 1 not in nf[nr - 1]
    assert nper == len(pedges)
    # no boundary face is left of the periodic boundaries
    assert (nb[:, 1] < 0).sum() == len(case.mesh.btype) - len(pedges)
