"""Pins for the oracle's port, lumped-element, plane-DFT, Debye and sheet arithmetic (CPU only).

The reference holds no golden vectors for these parts of its FDTD path. Each is pinned by an
identity or closed form that the reference's update must satisfy, evaluated from the oracle's
own records at FP64 (so a wrong term shows as a residual far above rounding), and each check is
shown to fail on a planted mistake (oracle_set_mutation):

* plane DFT (cpu.rs:336-361): the accumulators equal a direct f64 DFT of point-probe series at
  the same nodes, E at t = (n+1) dt and H at t = (n+1/2) dt, slot (q nf + kf) 10 + 2c;
  planted 8 (E phase at the half step) fails;
* port (sim.rs:484-523, cpu.rs:291-334): the resistive-source edge obeys the discrete circuit
  law of a Thevenin source of EMF s((n+1/2) dt) behind R feeding the edge's capacitance eps A/l
  and conductance sigma A/l, with the port's loop current I^{n+1/2}:
      (C_e/dt)(V^{n+1}-V^n) + (1/R + G)(V^{n+1}+V^n)/2 + I^{n+1/2} = s((n+1/2)dt)/R
  (one edge), and its per-edge form with R_e = R cols / m for columns of m edges; planted 6
  (source at (n+1) dt) fails;
* lumped capacitor (sim.rs:416-430): C adds count*C*l/A to the edge permittivity, i.e. the
  series edges' capacitances in series make C; planted 10 (A/l for l/A) fails;
* lumped inductor (sim.rs:524-534, cpu.rs:298-303): with the edge current recovered from
  Ampere's law at the edge, I^n = circ H^{n+1/2} - A (E^{n+1} - ca E^n) / cb, each of the count
  edges obeys (L/count)(I^{n+1} - I^n)/dt = l E^{n+1}; planted 11 (L count) fails;
* Debye ADE (cpu.rs:239-260) away from its limits: a cavity filled with a one-pole medium at
  w tau ~ 1, and one with the reference's wideband fit (Debye::band + fit, sim.rs:249-282) at
  eps_r 4, tan 0.02, rings at the complex frequency that Debye::response (sim.rs:264-273)
  predicts through the Yee scheme's own dispersion; planted 9 (tau halved) fails while the
  static and optical limits (test_oracle_physics) still hold;
* resistive sheet (cpu.rs:169-237): a sheet whose two faces share one branch row is linear and
  time-invariant, and its voltage/current sequences obey V(w) = Z(e^{i w dt}) I(w) with Z the
  closed-form impedance of the table's row (derived from cpu.rs:195-237):
      Z(z) = r [a0 - 2 g1/(z+1) + sum_real res q/(z-p) + sum_complex Re-pair of c Q/(z-P)];
  planted 13 (history sign) and 14 (real pole p, q swapped) fail;
* sheet face weights (cpu.rs:169-193): special cases (a plain sheet's weights sum to 1 and are
  the faces' energy shares; equal faces weigh equally; a dark face weighs nothing; opposite
  faces of an adaptive sheet weigh a quarter each; the adaptive scale saturates at 4); planted
  12 (adaptive scale dropped) fails.
"""
import cmath
import math

import numpy as np
import pytest

from fieldgrid_amd import fdtd as F
from oracle import oracle as O

from .fixtures import graded

EPS0 = 8.8541878128e-12
MU0 = 1.25663706212e-6
C0 = 299_792_458.0


def axis_d(line):
    """sim.rs:29-44: primal spacing d and dual spacing dd per node."""
    line = np.asarray(line, np.float64)
    n = len(line)
    d = np.empty(n)
    d[:-1] = line[1:] - line[:-1]
    d[-1] = line[-1] - line[-2]
    dd = np.empty(n)
    dd[0], dd[-1] = d[0], d[n - 2]
    dd[1:-1] = 0.5 * (line[2:] - line[:-2])
    return d, dd


def pulse(f0, fc, t):
    """sim.rs:646-657"""
    tau = math.sqrt(math.log(10.0)) / (math.pi * fc)
    t0 = 4.0 * tau
    x = (t - t0) / tau
    return np.exp(-x * x) * np.cos(2.0 * math.pi * f0 * (t - t0))


def mutated(m, fn, *a, **k):
    O.set_mutation(m)
    try:
        return fn(*a, **k)
    finally:
        O.set_mutation(0)


# ---- a graded test bed ---------------------------------------------------------------------------

def bed(eps=2.5, sigma=0.01):
    """A 22x20x18-node graded grid (CPML 4) whose cells are eps_r `eps`, sigma `sigma` in a block
    around the middle, vacuum elsewhere."""
    grid = F.Grid(graded(22, 1e-3, 1.02), graded(20, 0.9e-3, 1.03), graded(18, 1.1e-3, 0.99), 4)
    n = grid.dims()
    mats = F.Materials.new(grid)
    cells = (n[0] - 1, n[1] - 1, n[2] - 1)
    sl = (slice(6, 16), slice(5, 15), slice(5, 13))
    mats.eps.reshape(cells)[sl] = eps
    mats.inf.reshape(cells)[sl] = eps
    mats.sigma.reshape(cells)[sl] = sigma
    return grid, mats


def edge_geometry(grid, comp, at):
    """(length l, area A, H probes of the dual loop with their weights) of the edge (comp, at):
    A curl H = (Hv(p) - Hv(p-u)) dd_v - (Hu(p) - Hu(p-v)) dd_u (cpu.rs:109-167 times A)."""
    u, v = (comp + 1) % 3, (comp + 2) % 3
    dd = [axis_d(grid.axis(a))[1] for a in range(3)]
    d = axis_d(grid.axis(comp))[0]
    p = list(at)
    pu, pv = list(at), list(at)
    pu[u] -= 1
    pv[v] -= 1
    loop = [((3 + v, *p), dd[v][p[v]]), ((3 + v, *pu), -dd[v][p[v]]),
            ((3 + u, *p), -dd[u][p[u]]), ((3 + u, *pv), dd[u][p[u]])]
    return d[at[comp]], dd[u][at[u]] * dd[v][at[v]], loop


def probes_for(grid, edges):
    """Point probes: E of every edge and the four H values of its dual loop."""
    out = []
    for comp, at in edges:
        out.append((comp, *at))
        for pr, _ in edge_geometry(grid, comp, at)[2]:
            out.append(pr)
    return out


def edge_series(grid, rec, edges, nprobes):
    """Per edge: E^{n+1} after every step n and the loop sum A curl H^{n+1/2}."""
    pr = rec.probes.reshape(rec.steps, nprobes)
    out, col = [], 0
    for comp, at in edges:
        e = pr[:, col]
        loop = edge_geometry(grid, comp, at)[2]
        circ = sum(pr[:, col + 1 + q] * w for q, (_, w) in enumerate(loop))
        out.append((e, circ))
        col += 5
    return out


# ---- plane DFT -----------------------------------------------------------------------------------

PLANE_K = 9
PLANE_FREQS = [4e9, 9e9, 13e9]
PLANE_NODES = [(6, 7), (11, 10), (14, 5), (9, 13)]


def plane_run():
    grid, mats = bed()
    n = grid.dims()
    port = F.PortDef("p", [[F.Edge(2, (10, 9, k)) for k in (8, 9)]], 50.0, False)
    sim = F.Sim.try_new(grid, mats, None, [], [port], [], F.Media())
    probes = [(c, i, j, PLANE_K) for (i, j) in PLANE_NODES for c in range(5)]
    ex = F.Extras(max_steps=300, decay_db=math.inf, freqs=PLANE_FREQS, plane_k=PLANE_K, probes=probes,
                  check_every=100)
    rec = O.run_cpu(sim.desc, 0, F.Pulse(9e9, 7e9), ex, 64)
    return sim, rec, probes


def plane_error():
    """max |plane accumulator - direct DFT of the probe series| / max |accumulator|."""
    sim, rec, probes = plane_run()
    n = sim.dims()
    nf = len(PLANE_FREQS)
    pr = rec.probes.reshape(rec.steps, len(probes))
    s = np.arange(rec.steps, dtype=np.float64)
    dt = rec.dt
    worst, scale = 0.0, 0.0
    for q, (i, j) in enumerate(PLANE_NODES):
        slot = i * n[1] + j
        for kf, f in enumerate(PLANE_FREQS):
            w = 2 * math.pi * f
            for c in range(5):
                t = (s + (1.0 if c < 3 else 0.5)) * dt  # E after step n at (n+1) dt, H at (n+1/2) dt
                direct = np.sum(pr[:, q * 5 + c] * np.exp(-1j * w * t)) * dt
                o = (slot * nf + kf) * 10 + 2 * c
                acc = complex(rec.plane[o], rec.plane[o + 1])
                worst = max(worst, abs(acc - direct))
                scale = max(scale, abs(direct))
    return worst / scale


def test_point_probes_read_the_fields_they_name():
    """The probe extension is itself pinned: its last samples are the final fields at its nodes."""
    sim, rec, probes = plane_run()
    n = sim.dims()
    nn = n[0] * n[1] * n[2]
    last = rec.probes.reshape(rec.steps, len(probes))[-1]
    for q, (c, i, j, k) in enumerate(probes):
        f = rec.e if c < 3 else rec.h
        assert last[q] == f[(c % 3) * nn + F.idx(n, i, j, k)]


def test_the_plane_dft_is_the_direct_dft_of_the_field_series():
    err = plane_error()
    print("plane DFT vs direct DFT of probes", err)
    assert err < 1e-11


def test_a_plane_dft_with_the_e_phase_at_the_half_step_is_caught():
    err = mutated(8, plane_error)
    print("planted 8:", err)
    assert err > 1e-3


# ---- port identity ------------------------------------------------------------------------------

R_PORT = 50.0
PORT_PULSE = F.Pulse(9e9, 7e9)


def port_single(cap=None):
    """One single-edge 50-ohm z-port at (10, 9, 8), optionally with a lumped capacitor on the same
    edge, and a second (undriven, 75-ohm) two-column port. Returns (grid, record)."""
    grid, mats = bed()
    ports = [F.PortDef("a", [[F.Edge(2, (10, 9, 8))]], R_PORT, False),
             F.PortDef("b", [[F.Edge(2, (13, 12, 8))], [F.Edge(2, (13, 13, 8))]], 75.0, True)]
    lumped = [F.Lumped("c", [F.Edge(2, (10, 9, 8))], F.Element.Capacitor(cap))] if cap else []
    sim = F.Sim.try_new(grid, mats, None, lumped, ports, [], F.Media())
    ex = F.Extras(max_steps=400, decay_db=math.inf, check_every=200)
    return grid, O.run_cpu(sim.desc, 0, PORT_PULSE, ex, 64, fields=False)


def single_port_residual(cap=None):
    """The discrete Thevenin law of the driven one-edge port, max residual / max |s/R|."""
    grid, rec = port_single(cap)
    l, A, _ = edge_geometry(grid, 2, (10, 9, 8))
    eps, sigma = EPS0 * 2.5, 0.01
    if cap:
        eps += cap * l / A  # the lumped capacitor folded into the edge (count 1)
    ser = rec.series.reshape(rec.steps, 2, 2)
    V, I = ser[:, 0, 0], ser[:, 0, 1]
    Vp = np.concatenate([[0.0], V[:-1]])
    dt = rec.dt
    t = (np.arange(rec.steps) + 0.5) * dt
    lhs = (eps * A / l) * (V - Vp) / dt + (1.0 / R_PORT + sigma * A / l) * 0.5 * (V + Vp) + I
    rhs = pulse(PORT_PULSE.f0, PORT_PULSE.fc, t) / R_PORT
    return float(np.abs(lhs - rhs).max() / np.abs(rhs).max())


def test_a_driven_port_edge_obeys_its_thevenin_law():
    r = single_port_residual()
    print("one-edge port law residual", r)
    assert r < 1e-9


def test_a_capacitor_on_the_port_edge_adds_its_capacitance():
    r = single_port_residual(cap=0.2e-12)
    print("port + 0.2 pF law residual", r)
    assert r < 1e-9


def test_a_late_source_half_step_is_caught():
    r = mutated(6, single_port_residual)
    print("planted 6:", r)
    assert r > 1e-3


def test_a_misfolded_capacitor_is_caught():
    r = mutated(10, single_port_residual, cap=0.2e-12)
    print("planted 10:", r)
    assert r > 1e-3


COLUMNS = [[(2, (9, 8, 7)), (2, (9, 8, 8))], [(2, (10, 8, 7)), (2, (10, 8, 8))]]


def multi_port_residuals():
    """A 2-column port of 2 edges per column: each edge's law with R_e = R cols / m and
    EMF share s/m, the V readout -sum E l / cols and the I readout: at each edge of the middle
    column, l/height times the ring around all columns at that level, i.e. the sum of their
    A curl H (run.rs:305-340). Returns the three max residuals, relative."""
    grid, mats = bed()
    cols, m = len(COLUMNS), len(COLUMNS[0])
    port = F.PortDef("a", [[F.Edge(c, at) for c, at in col] for col in COLUMNS], R_PORT, False)
    sim = F.Sim.try_new(grid, mats, None, [], [port], [], F.Media())
    edges = [e for col in COLUMNS for e in col]
    probes = probes_for(grid, edges)
    ex = F.Extras(max_steps=400, decay_db=math.inf, check_every=200, probes=probes)
    rec = O.run_cpu(sim.desc, 0, PORT_PULSE, ex, 64, fields=False)
    dt = rec.dt
    t = (np.arange(rec.steps) + 0.5) * dt
    s = pulse(PORT_PULSE.f0, PORT_PULSE.fc, t)
    eps, sigma = EPS0 * 2.5, 0.01
    Re = R_PORT * cols / m
    worst_edge = 0.0
    series = edge_series(grid, rec, edges, len(probes))
    for (comp, at), (e, circ) in zip(edges, series):
        l, A, _ = edge_geometry(grid, comp, at)
        ep = np.concatenate([[0.0], e[:-1]])
        lhs = eps * A * (e - ep) / dt + (l / (2 * Re) + sigma * A / 2) * (e + ep)
        rhs = circ - s / (Re * m)
        worst_edge = max(worst_edge, float(np.abs(lhs - rhs).max() / np.abs(s / (Re * m)).max()))
    ser = rec.series.reshape(rec.steps, 2)
    V = -sum(e * edge_geometry(grid, c, at)[0] / cols for (c, at), (e, _) in zip(edges, series))
    mid = COLUMNS[cols // 2]
    height = sum(edge_geometry(grid, c, at)[0] for c, at in mid)
    # the loop at each middle-column edge's level rings every column (run.rs:102-127), so it reads
    # the sum of the columns' circulations at that level
    I = sum(edge_geometry(grid, c, at)[0] / height *
            sum(series[q][1] for q, (c2, at2) in enumerate(edges) if at2[c] == at[c]) for c, at in mid)
    return (worst_edge, float(np.abs(ser[:, 0] - V).max() / np.abs(V).max()),
            float(np.abs(ser[:, 1] - I).max() / np.abs(I).max()))


def test_every_edge_of_a_multi_column_port_obeys_its_share_of_the_law():
    edge, v, i = multi_port_residuals()
    print("per-edge law", edge, "V readout", v, "I readout", i)
    assert edge < 1e-9 and v < 1e-12 and i < 1e-12


# ---- lumped inductor -----------------------------------------------------------------------------

L_IND = 3e-9


def inductor_residual():
    """Two series edges carrying an inductor of L_IND (each L/2): the edge current recovered from
    Ampere's law obeys (L/2) dI/dt = l E, and starts at zero."""
    grid, mats = bed()
    edges = [(2, (8, 11, 7)), (2, (8, 11, 8))]
    port = F.PortDef("a", [[F.Edge(2, (11, 9, 8))]], R_PORT, False)
    lumped = [F.Lumped("l", [F.Edge(c, at) for c, at in edges], F.Element.Inductor(L_IND))]
    sim = F.Sim.try_new(grid, mats, None, lumped, [port], [], F.Media())
    probes = probes_for(grid, edges)
    ex = F.Extras(max_steps=500, decay_db=math.inf, check_every=250, probes=probes)
    rec = O.run_cpu(sim.desc, 0, PORT_PULSE, ex, 64, fields=False)
    dt = rec.dt
    eps, sigma = EPS0 * 2.5, 0.01
    x = sigma * dt / (2 * eps)
    ca, cb = (1 - x) / (1 + x), dt / (eps * (1 + x))  # sim.rs:431-441
    worst, start = 0.0, 0.0
    for (comp, at), (e, circ) in zip(edges, edge_series(grid, rec, edges, len(probes))):
        l, A, _ = edge_geometry(grid, comp, at)
        ep = np.concatenate([[0.0], e[:-1]])
        cur = circ - A * (e - ca * ep) / cb   # I^n from step n's E^{n+1}, E^n and H^{n+1/2}
        lhs = (L_IND / len(edges)) * (cur[1:] - cur[:-1]) / dt
        rhs = l * e[:-1]
        worst = max(worst, float(np.abs(lhs - rhs).max() / np.abs(rhs).max()))
        start = max(start, abs(cur[0]) / np.abs(cur).max())
    return worst, start


def test_an_inductor_obeys_v_equals_l_di_dt():
    r, start = inductor_residual()
    print("inductor law residual", r, "initial current", start)
    assert r < 1e-8 and start < 1e-12


def test_a_misshared_inductance_is_caught():
    r, _ = mutated(11, inductor_residual)
    print("planted 11:", r)
    assert r > 1e-2


# ---- Debye media away from their limits ----------------------------------------------------------

CAV_CELLS = (25, 14, 8)
CAV_SIDES = (0.10, 0.07, 0.05)


def pencil_poles(y, order, width):
    """Matrix-pencil poles z_k of a sum of damped exponentials y_n = sum c_k z_k^n."""
    Y = np.lib.stride_tricks.sliding_window_view(np.asarray(y, np.float64), width + 1)
    _, _, vh = np.linalg.svd(Y, full_matrices=False)
    V = vh[:order].T
    return np.linalg.eigvals(np.linalg.pinv(V[:-1]) @ V[1:])


def debye_cavity_mode(poles, inf, dd, f_guess, steps, decimate):
    """Mode 110 of a PEC cavity (25x14x8 cells of 0.1 x 0.07 x 0.05 m) filled with a Debye medium
    (eps_inf `inf`, delta `dd`, `poles`), rung by a soft source. Returns (measured, predicted)
    complex angular frequencies (field ~ e^{i w t}, Im w > 0 decays). The prediction solves
    (2/dt)^2 sin^2(w dt/2) eps_r(w) = (2/dt)^2 sin^2(w_v dt/2), the Yee scheme's vacuum mode w_v
    (the dispersion relation of test_oracle_physics) with eps_r(w) = inf + dd (re - i im) from
    Debye::response (sim.rs:264-273)."""
    lines = [np.linspace(0.0, L, n + 1) for L, n in zip(CAV_SIDES, CAV_CELLS)]
    grid = F.Grid(*lines, 0)
    n = grid.dims()
    mats = F.Materials.new(grid)
    mats.eps[:], mats.inf[:], mats.debye[:] = inf + dd, inf, dd
    debye = F.Debye(poles)
    port = F.PortDef("probe", [[F.Edge(2, (n[0] // 3, n[1] // 3, n[2] // 3))]], 1e7, False)
    sim = F.Sim.try_new(grid, mats, None, [], [port], [], F.Media(debye=debye))
    ex = F.Extras(max_steps=steps, decay_db=math.inf, check_every=steps)
    pl = F.Pulse(f_guess, 0.8 * f_guess)
    rec = O.run_cpu(sim.desc, 0, pl, ex, 64, fields=False)
    dt = rec.dt
    t0, _ = F.pulse_shape(pl.fc)
    start = int(2 * t0 / dt) + 10  # after the source pulse has ended
    y = rec.series[0::2][start::decimate][:3000]
    z = pencil_poles(y, 40, 1000)
    ssum = sum(math.sin(math.pi * (L / c) / (2 * L)) ** 2 / (L / c) ** 2 * m
               for m, L, c in zip((1, 1, 0), CAV_SIDES, CAV_CELLS))
    om2 = 4 * C0 ** 2 * ssum  # (2/dt)^2 sin^2(w_v dt / 2) of mode 110

    def eps_r(w):
        re, im = 0.0, 0.0
        for x, a in debye.poles:  # Debye::response at complex w
            r = w / x
            re, im = re + a / (1 + r * r), im + a * r / (1 + r * r)
        return inf + dd * (re - 1j * im)

    g = lambda w: (2 / dt) ** 2 * cmath.sin(w * dt / 2) ** 2 * eps_r(w) - om2
    w = complex(math.sqrt(om2 / (inf + 0.5 * dd)), 0.0)
    for _ in range(60):  # Newton with a central-difference derivative
        h = 1e-4 * abs(w)
        w -= g(w) / ((g(w + h) - g(w - h)) / (2 * h))
    zp = cmath.exp(1j * w * dt * decimate)
    zk = z[np.argmin(np.abs(z - zp))]
    return cmath.log(zk) / (1j * dt * decimate), w


def debye_mid_band_errors():
    """(df/f, dgamma/gamma) of a one-pole medium at w tau = 1 (eps_inf 2, delta 1, Q ~ 5) and
    of the reference's wideband fit at eps_r 4, tan 0.02 (Q ~ 50)."""
    f110 = C0 / 2 * math.sqrt(1 / CAV_SIDES[0] ** 2 + 1 / CAV_SIDES[1] ** 2)
    f1 = f110 / math.sqrt(2.5)
    out = []
    m, p = debye_cavity_mode([(2 * math.pi * f1, 1.0)], 2.0, 1.0, f1, 4000, 1)
    out.append((abs(m.real / p.real - 1), abs(m.imag / p.imag - 1)))
    band = F.Debye.band(0.2e9, 10e9)
    f2 = f110 / 2
    inf, dd = band.fit(4.0, 0.02, f2)
    m, p = debye_cavity_mode(band.poles, inf, dd, f2, 12000, 3)
    out.append((abs(m.real / p.real - 1), abs(m.imag / p.imag - 1)))
    q = m.real / (2 * m.imag)
    return out, q


def test_debye_media_ring_at_the_frequency_their_response_predicts():
    errs, q = debye_mid_band_errors()
    print("Debye mid band (df/f, dgamma/gamma):", errs, "wideband fit Q", q)
    for df, dg in errs:
        assert df < 2e-4 and dg < 1e-2
    assert abs(q - 50.0) < 1.0  # tan delta 0.02 at the fit frequency


def test_a_halved_debye_relaxation_time_is_caught():
    errs, _ = mutated(9, debye_mid_band_errors)
    print("planted 9:", errs)
    assert max(max(df / 2e-4, dg / 1e-2) for df, dg in errs) > 10


# ---- resistive sheets ----------------------------------------------------------------------------

# [nr, nc, -, -, p, q, (Pr, Pi, Qr, Qi), face row: a0, g1, res, (cr, ci)]: one real and one complex
# branch; both faces of the sheet below read row 0, so its weights sum to one and it is LTI
SHEET_ROW = dict(p=0.9, q=0.05, P=complex(0.8, 0.1), Q=complex(0.02, 0.01), a0=1.0, g1=0.2, res=0.1,
                 c=complex(0.05, 0.02))
SHEET_R = 40.0


def sheet_table(row):
    return [1.0, 1.0, 0.0, 0.0, row["p"], row["q"], row["P"].real, row["P"].imag, row["Q"].real, row["Q"].imag,
            row["a0"], row["g1"], row["res"], row["c"].real, row["c"].imag]


def sheet_impedance(row):
    """The closed-form impedance Z(z)/r of a sheet face row as a rational function of z^-1,
    (numerator, denominator) coefficients, derived from cpu.rs:195-237: with ibar the
    half-step current and i_{n+1} = 2 ibar_n - i_n, the edge's mean voltage l (E^{n+1}+E^n)/2 is
    r (a0 ibar - g1 i_n + res x_n + Re(c u_n)), where x_{n+1} = p x_n + q ibar_n and
    u_{n+1} = P u_n + Q ibar_n; so
        Z(z) = r [a0 - 2 g1/(z+1) + res q/(z-p) + (c Q/(z-P) + conj(c Q)/(z-conj P))/2]."""
    P = np.polynomial.polynomial
    one = np.array([1.0 + 0j])
    zi = np.array([0.0, 1.0 + 0j])  # z^-1
    f = {"g": np.array([1.0, 1.0 + 0j]), "p": np.array([1.0, -row["p"] + 0j]),
         "P": np.array([1.0, -row["P"]]), "Pc": np.array([1.0, -np.conj(row["P"])])}
    den = P.polymul(P.polymul(f["g"], f["p"]), P.polymul(f["P"], f["Pc"]))
    def others(*skip):
        out = one
        for k, v in f.items():
            if k not in skip:
                out = P.polymul(out, v)
        return out
    num = row["a0"] * den
    num = P.polyadd(num, -2 * row["g1"] * P.polymul(zi, others("g")))
    num = P.polyadd(num, row["res"] * row["q"] * P.polymul(zi, others("p")))
    num = P.polyadd(num, 0.5 * row["c"] * row["Q"] * P.polymul(zi, others("P")))
    num = P.polyadd(num, 0.5 * np.conj(row["c"] * row["Q"]) * P.polymul(zi, others("Pc")))
    assert np.abs(num.imag).max() < 1e-15 and np.abs(den.imag).max() < 1e-15
    return num.real, den.real


def sheet_case(row):
    """A 24x22x20-node 1 mm grid with CPML 6 (everything decays), a driven 50-ohm z-port and a
    non-adaptive x-directed sheet edge beside it whose faces share row 0."""
    grid = F.Grid(*(np.arange(m) * 1e-3 for m in (24, 22, 20)), 6)
    n = grid.dims()
    mats = F.Materials.new(grid)
    cells = (n[0] - 1, n[1] - 1, n[2] - 1)
    mats.sigma.reshape(cells)[9:15, 8:14, 7:13] = 0.02
    mats.eps.reshape(cells)[9:15, 8:14, 7:13] = 2.0
    mats.inf.reshape(cells)[9:15, 8:14, 7:13] = 2.0
    port = F.PortDef("p", [[F.Edge(2, (10, 11, k)) for k in (9, 10)]], 50.0, False)
    edge = (0, (11, 11, 10))
    sheet = F.Resistive(0, F.idx(n, *edge[1]), SHEET_R, False, (0, 0))
    media = F.Media(sheets=F.Sheets(sheet_table(row), 3))
    sim = F.Sim.try_new(grid, mats, None, [], [port], [sheet], media)
    return grid, sim, edge


def sheet_residual(row=SHEET_ROW):
    """max |V_n - r (Z/r filtering ibar)_n| / max |V_n| over a run: V_n = l (E^{n+1} + E^n)/2 is the
    sheet edge's mean voltage and ibar_n its current, recovered from Ampere's law at the edge,
    ibar = A curl H + A (ca E^n - E^{n+1}) / cb (the sheet's correction E -= c ibar, c = cb/A);
    the closed-form Z is applied as one rational filter from rest."""
    from scipy.signal import lfilter
    grid, sim, (comp, at) = sheet_case(row)
    probes = probes_for(grid, [(comp, at)])
    ex = F.Extras(max_steps=1500, decay_db=math.inf, check_every=500, probes=probes)
    rec = O.run_cpu(sim.desc, 0, F.Pulse(10e9, 8e9), ex, 64, fields=False)
    dt = rec.dt
    (e, circ), = edge_series(grid, rec, [(comp, at)], len(probes))
    l, A, _ = edge_geometry(grid, comp, at)
    eps, sigma = 2.0 * EPS0, 0.02
    x = sigma * dt / (2 * eps)
    ca, cb = (1 - x) / (1 + x), dt / (eps * (1 + x))  # sim.rs:431-441
    ep = np.concatenate([[0.0], e[:-1]])
    ibar = circ + A * (ca * ep - e) / cb
    vbar = 0.5 * l * (e + ep)
    num, den = sheet_impedance(row)
    pred = SHEET_R * lfilter(num, den, ibar)
    return float(np.abs(vbar - pred).max() / np.abs(vbar).max())


def test_a_sheet_edge_has_its_rows_closed_form_impedance():
    r = sheet_residual()
    print("sheet V - Z I residual", r)
    assert r < 1e-5


@pytest.mark.parametrize("mutation", [13, 14])
def test_a_wrong_sheet_branch_is_caught(mutation):
    try:
        r = mutated(mutation, sheet_residual)
    except O.OracleError as e:  # a wrong history sign can make the run blow up outright
        r = str(e)
    print(f"planted {mutation}:", r)
    assert isinstance(r, str) or r > 1e-3


def test_sheet_face_weights_in_their_special_cases():
    rng = np.random.default_rng(3)
    ha, hb = rng.standard_normal(200).astype(np.float32), rng.standard_normal(200).astype(np.float32)
    zero = np.zeros(200, np.float32)
    # a plain sheet: the weights are the faces' shares of the accumulated H^2, summing to one
    ma, mb = O.sheet_weights(False, ha, hb)
    aa, bb = np.cumsum(ha.astype(np.float64) ** 2), np.cumsum(hb.astype(np.float64) ** 2)
    assert np.allclose(ma, aa / (aa + bb), rtol=1e-5) and np.allclose(ma + mb, 1.0, rtol=1e-6)
    # equal faces weigh equally; an adaptive sheet keeps its starting sum while the faces agree
    for adaptive in (False, True):
        ma, mb = O.sheet_weights(adaptive, ha, ha)
        assert np.array_equal(ma, mb) and np.allclose(ma, 0.5, rtol=1e-6)
    # a dark face weighs nothing
    for adaptive in (False, True):
        ma, mb = O.sheet_weights(adaptive, ha, zero)
        assert np.allclose(ma, 1.0, rtol=1e-6) and np.all(mb == 0)
        ma, mb = O.sheet_weights(adaptive, zero, hb)
        assert np.all(ma == 0) and np.allclose(mb, 1.0, rtol=1e-6)
    # no field yet: the starting weights stay
    ma, mb = O.sheet_weights(True, zero, zero, start=(0.3, 0.9))
    assert np.all(ma == np.float32(0.3)) and np.all(mb == np.float32(0.9))
    # opposite faces of an adaptive sheet: sum / (sum - 2 ab) = 1/2, a quarter each
    ma, mb = O.sheet_weights(True, ha, -ha)
    assert np.allclose(ma, 0.25, rtol=1e-6) and np.allclose(mb, 0.25, rtol=1e-6)
    # nearly equal faces of an adaptive sheet: the scale saturates at 4
    ma, mb = O.sheet_weights(True, ha, ha * np.float32(1.001))
    assert np.allclose(ma + mb, 4.0, rtol=1e-6)
    # uncorrelated faces: the scale is sum / (sum - 2 ab), clamped to [1/2, 4]
    ma, mb = O.sheet_weights(True, ha, hb)
    ab = np.cumsum(ha.astype(np.float64) * hb)
    scale = np.clip((aa + bb) / (aa + bb - 2 * ab), 0.5, 4.0)
    assert np.allclose(ma + mb, scale, rtol=1e-5)


def test_a_dropped_adaptive_scale_is_caught():
    rng = np.random.default_rng(3)
    ha = rng.standard_normal(50).astype(np.float32)
    ma, mb = mutated(12, O.sheet_weights, True, ha, -ha)
    assert not np.allclose(ma, 0.25, rtol=1e-3)
