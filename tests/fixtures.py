"""Problem builders shared by the tests: the reference's own fixtures restated
(crates/sims-field/tests/fdtd.rs:25-125) and seeded versions of the benchmark workloads."""
import math

import numpy as np

from fieldgrid_amd import fdtd as F


def graded(n, start, grow):
    """tests/fdtd.rs:25-33"""
    out = [0.0]
    d = start
    for _ in range(1, n):
        out.append(out[-1] + d)
        d *= grow
    return np.array(out)


def sheet_table():
    """tests/fdtd.rs:35-41"""
    return F.Sheets([1.0, 1.0, 2.0, 0.0, 0.9, 0.05, 0.8, 0.1, 0.02, 0.01, 1.0, 0.2, 0.1, 0.05, 0.02, 1.2, 0.1,
                     0.08, 0.03, 0.01], 3)


def everything_parts(sheets=True, wires=True):
    """tests/fdtd.rs:43-112: every feature on a 26x22x20 graded grid with pml 4."""
    grid = F.Grid(graded(26, 1e-3, 1.01), graded(22, 1e-3, 1.0), graded(20, 0.8e-3, 1.02), 4)
    n = grid.dims()
    mats = F.Materials.new(grid)
    m = np.zeros((n[0] - 1, n[1] - 1, n[2] - 1), bool)
    m[:, :, 6:9] = True
    m = m.ravel()
    mats.eps[m], mats.inf[m], mats.debye[m], mats.sigma[m] = 4.0, 3.6, 0.4, 0.002
    debye = F.Debye.band(1e9, 2e10)

    def pec(c, i, j, k):
        return (c != 2) & (k == 6) & (i < 7)

    lumped = [F.Lumped("c", [F.Edge(2, (14, 11, 6))], F.Element.Capacitor(1e-12)),
              F.Lumped("r", [F.Edge(2, (16, 11, 6))], F.Element.Resistor(30.0)),
              F.Lumped("l", [F.Edge(2, (18, 11, 6))], F.Element.Inductor(2e-9))]

    def column(i, j):
        return [F.Edge(2, (i, j, k)) for k in range(6, 9)]

    ports = [F.PortDef("a", [column(10, 10), column(10, 11)], 50.0, False),
             F.PortDef("b", [column(20, 12)], 75.0, True)]
    resistive = [F.Resistive(0, F.idx(n, i, 8, 9), 0.4, i % 2 == 0, (0, 1)) for i in range(8, 14)] if sheets else []
    wires = ([(F.Edge(2, (8, 16, k)), 0.2e-3) for k in range(10, 14)] +
             [(F.Edge(0, (i, 16, 14)), 0.2e-3) for i in range(8, 12)]) if wires else []
    media = F.Media(sheets=sheet_table() if sheets else F.Sheets(), debye=debye, wires=wires)
    return grid, mats, pec, lumped, ports, resistive, media


def everything(sheets=True, wires=True):
    return F.Sim.try_new(*everything_parts(sheets, wires))


def everything_extras(sim, steps, fused_e=False):
    """tests/fdtd.rs:114-125"""
    return F.Extras(max_steps=steps, decay_db=math.inf, freqs=[3e9, 7e9], plane_k=7,
                    ntff=F.patches(sim, 2).gpu, fused_e=fused_e, check_every=250)


def worst(a, b):
    """tests/fdtd.rs:127-132: max|a-b| / max|a|"""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    assert a.shape == b.shape
    scale = np.abs(a).max()
    assert scale > 0
    return float(np.abs(a - b).max() / scale)


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(a))


def plain_vacuum(cells=63, dx=1e-3, pml=10, seed=1, nprobes=8):
    """The plain CPU-runnable workload: a cells^3 vacuum grid (cells+1 nodes per axis),
    CPML on all faces, a soft Ez point source at the centre (a 1e7-ohm single-edge port,
    the reference's soft-source idiom, tests/fdtd.rs:206-211) and Ez/Hy point probes at
    seeded interior nodes."""
    grid = F.Grid.uniform([0.0] * 3, [cells * dx] * 3, [cells] * 3, pml)
    n = grid.dims()
    mats = F.Materials.new(grid)
    c = n[0] // 2
    port = F.PortDef("src", [[F.Edge(2, (c, c, c))]], 1e7, False)
    sim = F.Sim.try_new(grid, mats, None, [], [port], [], F.Media())
    rng = np.random.default_rng(seed)
    cells_ = rng.integers(pml + 1, n[0] - pml - 1, size=(nprobes, 3))
    probes = [(2, *map(int, p)) for p in cells_] + [(4, *map(int, p)) for p in cells_]
    return sim, probes
