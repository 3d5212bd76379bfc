"""Pins for the oracle's resistive sheets (cpu.rs:169-237) and thin wires (sim.rs:556-635,
cpu.rs:270-273), CPU only. The reference holds no golden vectors for them, so they are pinned by
what the mathematics fixes:

* a sheet face with no branches and no history term is a trapezoidal resistor of r * g0 ohms,
  which is exactly the reference's lumped resistor (sim.rs:484-491): the FP64 runs agree to
  rounding;
* a complex branch whose pole has no imaginary part is two real branches (residues cr and -ci):
  the FP64 runs agree to rounding, which fixes the complex branch's index and sign layout
  against the real one;
* the edges around a thin wire of radius r in cells of size d carry eps * f and the H nodes
  around it mu / f, f = 2 / ln(d / r) (the thin-wire subcell model): the builder's coefficients
  and its H correction terms are those closed forms;
* the reference's own thin-wire stability test (tests/fdtd.rs:238-270) on the FP32 path.

The branches' values and the face weights are pinned in tests/test_oracle_pins.py: the sheet's
voltage and current follow the closed-form impedance of its table row, and the weights meet
their special cases.
"""
import math

import numpy as np
import pytest

from fieldgrid_amd import fdtd as F
from oracle import oracle as O

from .fixtures import worst

EPS0 = 8.8541878128e-12
MU0 = 1.25663706212e-6


def _uniform(n, d=1e-3):
    return np.arange(n) * d


def _sheet_case(extra_sheets=(), table=None, lumped=(), sheet_r=25.0):
    """A 21x19x17-node 1 mm grid (CPML 4) with a 50-ohm z-port and an x edge beside it."""
    grid = F.Grid(_uniform(21), _uniform(19), _uniform(17), 4)
    n = grid.dims()
    mats = F.Materials.new(grid)
    cells = (n[0] - 1, n[1] - 1, n[2] - 1)
    mats.sigma.reshape(cells)[8:13, 7:11, 6:10] = 0.02  # a lossy neighbourhood: x_sig != 0 on the edge
    port = F.PortDef("p", [[F.Edge(2, (8, 9, k)) for k in range(7, 10)]], 50.0, False)
    resistive = [F.Resistive(0, F.idx(n, 10, 9, 8), sheet_r, False, (0, 1))] if table is not None else []
    resistive += list(extra_sheets)
    media = F.Media(sheets=F.Sheets(table, 4) if table is not None else F.Sheets())
    return F.Sim.try_new(grid, mats, None, list(lumped), [port], resistive, media)


def _run64(sim, steps=400):
    ex = F.Extras(max_steps=steps, decay_db=math.inf, probes=[(0, 10, 9, 8), (2, 8, 9, 8), (4, 10, 9, 8)])
    return O.run_cpu(sim.desc, 0, F.Pulse(10e9, 10e9), ex, 64)


def test_a_resistive_sheet_face_is_a_lumped_resistor():
    """nr = nc = 0, g1 = 0: ibar (r g0 + len c / 2) = len (E_pre + E_old) / 2 and E = E_pre - c ibar
    is the trapezoidal resistor of r g0 ohms; the lumped resistor's (ca, cb) with
    beta = dt len / (2 R eps A) + x (sim.rs:484-491) is the same update solved in closed form.
    The sheet's record (len, c, r, E_old, weights) is f32 even on the FP64 path (cpu.rs:169-237),
    so the two agree to f32 rounding of the sheet current, accumulated over the run (< 5e-6 of
    the fields' peak); the resistor itself moves the fields by ~5e-2."""
    sheet = _sheet_case(table=[0.0, 0.0, 0.0, 0.0, 2.0, 0.0, 2.0, 0.0], sheet_r=25.0)
    lumped = _sheet_case(lumped=[F.Lumped("r", [F.Edge(0, (10, 9, 8))], F.Element.Resistor(50.0))])
    bare = _sheet_case()
    assert sheet.sheets == 1
    a, b, c = _run64(sheet), _run64(lumped), _run64(bare)
    for name in ("e", "h", "probes", "series"):
        x, y = getattr(a, name), getattr(b, name)
        print(f"{name}: sheet vs lumped resistor {worst(y, x):.2e}, vs no resistor {worst(getattr(c, name), x):.2e}")
        assert worst(y, x) < 5e-6
    assert worst(c.probes, a.probes) > 1e-2  # the resistor matters at its own edge


def test_a_complex_branch_with_no_imaginary_part_is_two_real_branches():
    """One complex branch (pr, pi = 0, q1, q2) with face residues (cr, ci) against two real
    branches (pr, q1), (pr, q2) with residues (cr, -ci), on two adaptive and two plain sheets."""
    n = (21, 19, 17)
    sheets = [F.Resistive(0, F.idx(n, i, 9, 8), 30.0, i % 2 == 0, (0, 1)) for i in range(10, 14)]
    cplx = [0.0, 1.0, 0.0, 0.0, 0.8, 0.0, 0.1, 0.02,
            1.0, 0.2, 0.05, 0.03,
            1.2, 0.1, 0.02, 0.04]
    real = [2.0, 0.0, 0.0, 0.0, 0.8, 0.1, 0.8, 0.02,
            1.0, 0.2, 0.05, -0.03,
            1.2, 0.1, 0.02, -0.04]
    a = _run64(_sheet_case(extra_sheets=sheets, table=cplx))
    b = _run64(_sheet_case(extra_sheets=sheets, table=real))
    bare = _run64(_sheet_case())
    for name in ("e", "h", "probes", "series"):
        w = worst(getattr(b, name), getattr(a, name))
        print(f"{name}: complex vs two real {w:.2e}")
        assert w < 1e-10
    assert worst(bare.probes, a.probes) > 1e-2
    # a sign slip in the complex residue would show
    wrong = list(cplx)
    wrong[11], wrong[15] = -cplx[11], -cplx[15]
    c = _run64(_sheet_case(extra_sheets=sheets, table=wrong))
    assert worst(b.probes, c.probes) > 1e-6


def _wire_case(radius, d=1e-3):
    grid = F.Grid(_uniform(24, d), _uniform(22, d), _uniform(20, d), 4)
    mats = F.Materials.new(grid)
    port = F.PortDef("p", [[F.Edge(2, (7, 10, k)) for k in range(8, 11)]], 50.0, False)
    wire = [(F.Edge(2, (12, 10, k)), radius) for k in range(7, 11)]
    return F.Sim.try_new(grid, mats, None, [], [port], [], F.Media(wires=wire)), wire


@pytest.mark.parametrize("radius", [0.05e-3, 0.2e-3, 0.45e-3])
def test_the_cells_around_a_thin_wire_carry_eps_times_f_and_mu_over_f(radius):
    d = 1e-3
    sim, wire = _wire_case(radius, d)
    n = sim.dims()
    nn = n[0] * n[1] * n[2]
    f = 2.0 / math.log(d / radius)
    ref = O.Packed(sim.desc)
    ca, cb = ref.coefficients()
    dt = ref.dt
    key = lambda c, p: c * nn + F.idx(n, *p)
    on_wire = {key(e.comp, e.at) for e, _ in wire}
    # E edges of the two other axes at both ends of every wire edge and one cell back: eps * f
    scaled = set()
    for e, _ in wire:
        p = list(e.at)
        end = list(p)
        end[e.comp] += 1
        for axis in ((e.comp + 1) % 3, (e.comp + 2) % 3):
            for node in (p, end):
                back = list(node)
                back[axis] -= 1
                for at in (node, back):
                    if key(axis, at) not in on_wire:
                        scaled.add(key(axis, at))
    flat_ca, flat_cb = ca.reshape(-1), cb.reshape(-1)
    want = dt / (EPS0 * f)
    for q in scaled:
        assert flat_ca[q] == np.float32(1.0)
        assert abs(flat_cb[q] - want) <= 2e-7 * want
    others = np.ones(3 * nn, bool)
    others[list(scaled)] = False
    interior = flat_cb[others][flat_cb[others] != 0]
    assert abs(np.median(interior) - dt / EPS0) <= 2e-7 * dt / EPS0
    # H nodes circling the wire: H += (f - 1) k_mu (minus the off-wire curl), so that with the
    # plain update the node sees mu / f. Expected terms from the Yee curl of component c at q:
    # -k_mu [(E_v(q+u) - E_v(q)) / d_u - (E_u(q+v) - E_u(q)) / d_v]
    k_mu = dt / MU0
    p = ref.problem
    assert p.n_wires == 4 * 4  # four H nodes around each of the four edges (none shared)
    for w in range(p.n_wires):
        entry = p.wires[w]
        c, id_ = divmod(int(entry.h), nn)
        q = [id_ // (n[1] * n[2]), (id_ // n[2]) % n[1], id_ % n[2]]
        u, v = (c + 1) % 3, (c + 2) % 3
        qu, qv = list(q), list(q)
        qu[u] += 1
        qv[v] += 1
        s = (f - 1.0) * k_mu
        expect = {key(v, qu): -s / d, key(v, q): s / d, key(u, qv): s / d, key(u, q): -s / d}
        expect = {e: x for e, x in expect.items() if e not in on_wire}
        got = {int(entry.e[t]): float(entry.c[t]) for t in range(4) if entry.c[t] != 0.0}
        assert set(got) == set(expect)
        for e, x in expect.items():
            assert abs(got[e] - x) <= 2e-7 * abs(x)


@pytest.mark.parametrize("radius", [0.09e-3, 1.7e-3])
def test_thin_wires_at_either_end_of_their_radius_stay_stable(radius):
    """tests/fdtd.rs:238-270 on the FP32 path: 40 000 steps, the late peak of |V| stays below
    twice the early one."""
    sim, ex, pulse = thin_wire_case(radius)
    rec = O.run_cpu(sim.desc, 0, pulse, ex, 32, fields=False)
    v = np.abs(np.asarray(rec.series).reshape(-1, 2)[:, 0])
    early, late = v[5000:10000].max(), v[35000:40000].max()
    print(f"radius {radius}: early {early:.3e} late {late:.3e}")
    assert late < 2.0 * early


def thin_wire_case(radius, steps=40_000):
    """The reference's thin-wire stability problem (tests/fdtd.rs:240-262)."""
    a, b, d = 0.10, 0.07, 0.05
    grid = F.Grid(np.array([a * i / 22.0 for i in range(23)]), np.array([b * i / 18.0 for i in range(19)]),
                  np.array([d * i / 13.0 for i in range(14)]), 0)
    n = grid.dims()
    port = F.PortDef("probe", [[F.Edge(2, (n[0] // 3, n[1] // 3, n[2] // 3))]], 1e7, False)
    wires = [(F.Edge(2, (11, 9, k)), radius) for k in range(3, 9)] + \
            [(F.Edge(0, (i, 9, 9)), radius) for i in range(11, 15)]
    sim = F.Sim.try_new(grid, F.Materials.new(grid), None, [], [port], [], F.Media(wires=wires))
    return sim, F.Extras(max_steps=steps, decay_db=math.inf), F.Pulse(3.0e9, 2.5e9)
