"""The reference's own unit tests for this path, restated against the oracle and the product
builder (CPU only):
  crates/sims-field/src/fdtd/run.rs:775-823   a_port_loop_circles_only_its_own_columns
  crates/sims-field/src/fdtd/run.rs:825-858   coefficient_sets_give_back_every_edge
  crates/sims-field/tests/fdtd.rs:176-187     a_run_stops_when_progress_says_so
  crates/sims-field/tests/fdtd.rs:271-286     a_thin_wire_outside_its_radius_range_is_refused
plus the builder parity: the product's C++ builder must produce bit-identical tables to the
oracle's restatement of Sim::try_new + pack.
"""
import ctypes as C
import math

import numpy as np
import pytest

from fieldgrid_amd import abi
from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W
from oracle import oracle as O

from .fixtures import everything, everything_extras, everything_parts


def line(n):
    """run.rs:771-773"""
    return np.array([i * 1e-3 * (1.0 + 0.1 * i) for i in range(n)])


def ring_sim():
    grid = F.Grid(line(9), line(8), line(7), 0)
    mats = F.Materials.new(grid)

    def column(i, j):
        return [F.Edge(2, (i, j, 3))]

    ring = [column(i, j) for i in range(2, 5) for j in range(2, 5) if (i, j) != (3, 3)]
    port = F.PortDef("p", ring, 50.0, False)
    return grid, mats, ring, port


def test_a_port_loop_circles_only_its_own_columns():
    grid, mats, ring, port = ring_sim()
    sim = F.Sim.try_new(grid, mats, None, [], [port], [], F.Media())
    n = sim.dims()
    got = O.port_loop(sim.desc, 0, 2, 3)

    def single(i, j):
        one = F.PortDef("q", [[F.Edge(2, (i, j, 3))]], 50.0, False)
        return O.port_loop(F.Sim.try_new(grid, mats, None, [], [one], [], F.Media()).desc, 0, 2, 3)

    want = {}
    for c in ring:
        for id_, comp, w in single(c[0].at[0], c[0].at[1]):
            want[(id_, comp)] = want.get((id_, comp), 0.0) + w
    want = {k: v for k, v in want.items() if abs(v) > 1e-12}
    assert len(got) == 16 and len(want) == 16
    for id_, comp, w in got:
        assert abs(want[(id_, comp)] - w) < 1e-15
    centre = {(i, c): w for i, c, w in got}
    for id_, comp, w in single(3, 3):
        assert centre[(id_, comp)] == -w
    sq = lambda i, j: F.idx(n, i, j, 3)
    sx = single(3, 3)
    assert len(sx) == 4
    assert any(s[0] == sq(3, 3) and s[1] == 1 and s[2] > 0 for s in sx)
    assert any(s[0] == sq(2, 3) and s[1] == 1 and s[2] < 0 for s in sx)
    assert any(s[0] == sq(3, 3) and s[1] == 0 and s[2] < 0 for s in sx)
    assert any(s[0] == sq(3, 2) and s[1] == 0 and s[2] > 0 for s in sx)


def test_the_product_builder_packs_the_same_port_loop():
    """The loop terms the product packs for a ring port equal the restated port_loop, each
    scaled by polarity*d/height (run.rs:322-329)."""
    grid, mats, ring, port = ring_sim()
    sim = F.Sim.try_new(grid, mats, None, [], [port], [], F.Media())
    p = sim.problem()
    want = O.port_loop(sim.desc, 0, 2, 3)
    got = [(p.loops[q].id, p.loops[q].comp, p.loops[q].w) for q in range(p.n_loops)]
    assert [(a, b) for a, b, _ in got] == [(a, b) for a, b, _ in want]
    for (_, _, w), (_, _, w0) in zip(got, want):
        assert w == np.float32(w0 * 1.0)  # one layer: share = d/height = 1


def coef_fixture():
    """run.rs:825-842"""
    grid = F.Grid(line(9), line(8), line(7), 0)
    mats = F.Materials.new(grid)
    q = np.arange(mats.eps.size)
    mats.eps = (1.0 + (q % 5)).astype(np.float32)
    mats.sigma = ((q % 3) * 0.01).astype(np.float32)
    port = F.PortDef("p", [[F.Edge(2, (4, 4, 3))]], 50.0, False)

    def pec(c, i, j, k):
        return (c == 0) & (k == 2)

    return grid, mats, pec, port


@pytest.mark.parametrize("narrow", [0, 1], ids=["narrow", "wide"])
def test_coefficient_sets_give_back_every_edge(narrow):
    grid, mats, pec, port = coef_fixture()
    sim = F.Sim.try_new(grid, mats, pec, [], [port], [], F.Media())
    ca, cb = sim.coefficients()
    nn = ca.shape[1]
    packed = O.Packed(sim.desc, narrow)  # owns the oracle's tables while they are read
    for problem, who in ((sim.problem(narrow), "product"), (packed.problem, "oracle")):
        assert problem.coef_wide == (narrow == 1), who
        sets = np.ctypeslib.as_array(problem.coef_sets, shape=(2 * problem.n_sets,)).reshape(-1, 2)
        assert len(sets) > 2
        words = 3 * nn if problem.coef_wide else (3 * nn + 1) // 2
        index = np.ctypeslib.as_array(problem.coef_index, shape=(words,))
        flat = index if problem.coef_wide else index.view(np.uint16)[: 3 * nn]
        got = sets[flat.astype(np.int64)]
        assert np.array_equal(got[:, 1], cb.ravel()), who
        nz = cb.ravel() != 0
        assert np.array_equal(got[nz, 0], ca.ravel()[nz]), who


def test_a_run_stops_when_progress_says_so():
    sim = everything()
    ex = everything_extras(sim, 5000)
    with pytest.raises(O.OracleError) as e:
        O.run_cpu(sim.desc, 0, F.Pulse(5e9, 5e9), ex, 32, cancel_at=2, fields=False)
    assert str(e.value) == "cancelled"


@pytest.mark.parametrize("radius", [0.6e-3, 0.01e-3])
def test_a_thin_wire_outside_its_radius_range_is_refused(radius):
    grid = F.Grid(*(np.arange(11) * 1e-3 for _ in range(3)), 0)
    mats = F.Materials.new(grid)
    media = F.Media(wires=[(F.Edge(2, (5, 5, 4)), radius)])
    with pytest.raises(abi.FieldgridError) as e:
        F.Sim.try_new(grid, mats, None, [], [], [], media)
    assert "from a 55th to half of its cell" in str(e.value)
    d = F.SimDesc(grid, mats, None, [], [], [], media)
    assert O.check(d) == str(e.value)


def _tables(problem, nn):
    """Every table of an fg_problem as numpy arrays (for bit comparison)."""
    def arr(ptr, count, dtype):
        if count == 0:
            return np.zeros(0, dtype)
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_uint8)), shape=(count * np.dtype(dtype).itemsize,)).view(dtype).copy()
    p = problem
    n = list(p.n)
    words = 3 * nn if p.coef_wide else (3 * nn + 1) // 2
    out = {
        "n": np.array(n), "pml": np.array([p.pml]), "dt": np.array([p.dt]), "wide": np.array([p.coef_wide]),
        "axes": arr(p.axes, 10 * sum(n), np.float32),
        "sets": arr(p.coef_sets, 2 * p.n_sets, np.float32),
        "index": arr(p.coef_index, words, np.uint32),
        "ports": arr(p.ports, p.n_ports * C.sizeof(abi.fg_port), np.uint8),
        "edges": arr(p.port_edges, p.n_port_edges * C.sizeof(abi.fg_port_edge), np.uint8),
        "loops": arr(p.loops, p.n_loops * C.sizeof(abi.fg_loop_term), np.uint8),
        "sources": arr(p.sources, p.n_sources * C.sizeof(abi.fg_source), np.uint8),
        "inductors": arr(p.inductors, p.n_inductors * C.sizeof(abi.fg_inductor), np.uint8),
        "debye": arr(p.debye_edges, p.n_debye * C.sizeof(abi.fg_debye_edge), np.uint8),
        "dtab": arr(p.debye_table, max(p.n_pole_sets, 1) * p.n_poles * 2, np.float32),
        "sheets": arr(p.sheets, 16 * p.n_sheets, np.uint32),
        "branches": arr(p.branches, p.n_branches, np.float32),
        "wires": arr(p.wires, p.n_wires * C.sizeof(abi.fg_wire_h), np.uint8),
    }
    return out


def _same_tables(sim):
    nn = int(np.prod(sim.dims()))
    mine = _tables(sim.problem(), nn)
    ref = O.Packed(sim.desc)
    theirs = _tables(ref.problem, nn)
    for key in mine:
        a, b = np.asarray(mine[key]), np.asarray(theirs[key])
        assert a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8)), key
    ca, cb = sim.coefficients()
    oca, ocb = ref.coefficients()
    assert np.array_equal(ca.view(np.uint32), oca.view(np.uint32))
    assert np.array_equal(cb.view(np.uint32), ocb.view(np.uint32))


def test_the_builder_matches_the_oracle_on_every_feature():
    """Sim::try_new + pack on tests/fdtd.rs:43-112 (Debye band, PEC, C/R/L, two ports, six
    sheets, eight thin-wire edges): every table bit-identical."""
    sim = everything()
    assert sim.sheets == 6 and sim.inductors == 1 and sim.debye_edges > 0 and sim.wires > 30
    _same_tables(sim)


@pytest.mark.parametrize("per_box_tau", [False, True])
def test_the_builder_matches_the_oracle_on_boxes(per_box_tau):
    _same_tables(W.boxes(nodes=40, per_box_tau=per_box_tau, nboxes=12, ndebye=4))


def test_the_builder_matches_the_oracle_on_layers():
    _same_tables(W.layered(dims=(40, 36, 48), pml=6, port_i=(12, 20)))


def test_patches_match_the_oracle():
    sim = everything()
    p = F.patches(sim, 2)
    g, pos, area = O.patches(sim.grid, 2)
    assert np.array_equal(p.gpu, g) and np.array_equal(p.pos, pos) and np.array_equal(p.area, area)
    n = sim.dims()
    inner = [d - 2 * (4 + 2) - 1 for d in n]
    assert len(p.area) == 2 * (inner[1] * inner[2] + inner[2] * inner[0] + inner[0] * inner[1])
