"""Pins for the oracle from physics the reference's algorithm must satisfy (CPU only).

The reference ships no numeric golden files for its FDTD path, so the oracle's restatement of
cpu.rs is pinned by checks chosen so that a plausible mistake fails one of them:
  * the reference's own analytic cavity test (crates/sims-field/tests/fdtd.rs:189-236);
  * the exact Yee dispersion of a uniform anisotropic PEC cavity (catches axis mix-ups);
  * the Courant limit: stable just below it, unstable just above (catches curl sign errors);
  * CPML reflection below -40 dB against a 3x larger grid (catches psi and b/c mistakes);
  * Debye limits: tau -> 0 behaves as eps_inf + delta, tau -> inf as eps_inf;
  * a short dipole: NTFF radiated power equals the port's input power and its directivity
    lies between the Hertzian 1.5 and the half-wave 1.64 (crates/sims-antenna/tests/fdtd.rs:45-55).
test_planted_mistakes_are_caught plants one mistake at a time in the oracle and shows the
matching check fails.
"""
import math

import numpy as np
import pytest

from fieldgrid_amd import fdtd as F
from oracle import oracle as O

C0 = F.C0


def spectrum_peak(v, dt, lo, hi, points=801, hann=False):
    """|sum v_s e^{-i w s}| maximised over [lo, hi] (tests/fdtd.rs:217-232), refined by a
    parabola through the best sample and its neighbours. `hann` tapers the record first,
    which suppresses leakage from neighbouring modes for sub-1e-4 peak positions."""
    v = np.asarray(v, np.float64)
    s = np.arange(len(v))
    if hann:
        v = v * (0.5 - 0.5 * np.cos(2 * math.pi * s / (len(v) - 1)))
    fs = np.linspace(lo, hi, points)
    mags = np.array([abs(np.dot(v, np.exp(-1j * 2 * math.pi * f * dt * s))) for f in fs])
    q = int(np.argmax(mags))
    q = min(max(q, 1), points - 2)
    a, b, c = mags[q - 1], mags[q], mags[q + 1]
    shift = 0.5 * (a - c) / (a - 2 * b + c) if (a - 2 * b + c) != 0 else 0.0
    return fs[q] + shift * (fs[1] - fs[0])


def cavity(xs, ys, zs, eps=1.0, debye=None, r=1e7, steps=40000, pulse=(3.0e9, 2.5e9)):
    grid = F.Grid(xs, ys, zs, 0)
    n = grid.dims()
    mats = F.Materials.new(grid)
    media = F.Media()
    if debye is not None:
        inf, dd, tau = debye
        mats.eps[:], mats.inf[:], mats.debye[:] = inf + dd, inf, dd
        media = F.Media(debye=F.Debye([(1.0 / tau, 1.0)]))
    else:
        mats.eps[:], mats.inf[:] = eps, eps
    port = F.PortDef("probe", [[F.Edge(2, (n[0] // 3, n[1] // 3, n[2] // 3))]], r, False)
    sim = F.Sim.try_new(grid, mats, None, [], [port], [], media)
    ex = F.Extras(max_steps=steps, decay_db=math.inf, check_every=5000)
    rec = O.run_cpu(sim.desc, 0, F.Pulse(*pulse), ex, 32, fields=False)
    return sim, rec.series[0::2]


# ---- the reference's own cavity test ---------------------------------------------------------

def reference_cavity_errors():
    """tests/fdtd.rs:189-236: graded x, modes 110, 210, 111 within 0.4 %."""
    a, b, d = 0.10, 0.07, 0.05
    xs, t, s = [], 0.0, 0.0035
    while t < a - 1e-9:
        xs.append(t)
        t += s
        s = min(s * 1.06, 0.0055)
    xs.append(a)
    ys = [b * i / 18 for i in range(19)]
    zs = [d * i / 13 for i in range(14)]
    sim, v = cavity(xs, ys, zs)
    errs = []
    for m, nn, p in [(1, 1, 0), (2, 1, 0), (1, 1, 1)]:
        fa = C0 / 2.0 * math.sqrt((m / a) ** 2 + (nn / b) ** 2 + (p / d) ** 2)
        peak = spectrum_peak(v, sim.dt, 0.97 * fa, 1.03 * fa, 401)
        errs.append(abs(peak / fa - 1.0))
    return errs


@pytest.mark.slow
def test_an_empty_cavity_rings_at_its_analytic_modes():
    errs = reference_cavity_errors()
    print("mode errors", errs)
    assert max(errs) < 0.004


# ---- exact Yee dispersion ---------------------------------------------------------------------

CELLS = (25, 14, 8)
SIDES = (0.10, 0.07, 0.05)


def dispersion_errors():
    """Peak of each mode vs the Yee scheme's own discrete eigenfrequency:
    sin^2(w dt/2)/(c dt)^2 = sum_a sin^2(m_a pi d_a / (2 L_a)) / d_a^2 (uniform, anisotropic)."""
    lines = [np.linspace(0.0, L, n + 1) for L, n in zip(SIDES, CELLS)]
    sim, v = cavity(*lines)
    dt = sim.dt
    out = []
    for mode in [(1, 1, 0), (2, 1, 0), (1, 1, 1)]:
        ssum = 0.0
        for m, L, n in zip(mode, SIDES, CELLS):
            d = L / n
            ssum += math.sin(m * math.pi * d / (2 * L)) ** 2 / d ** 2
        fd = 2.0 / dt * math.asin(C0 * dt * math.sqrt(ssum)) / (2 * math.pi)
        fc = C0 / 2.0 * math.sqrt(sum((m / L) ** 2 for m, L in zip(mode, SIDES)))
        peak = spectrum_peak(v, dt, 0.995 * fd, 1.005 * fd, 801, hann=True)
        out.append((abs(peak / fd - 1.0), abs(peak / fc - 1.0)))
    return out


@pytest.mark.slow
def test_a_uniform_cavity_rings_at_the_yee_schemes_own_frequencies():
    errs = dispersion_errors()
    print("discrete / continuous errors", errs)
    for disc, cont in errs:
        assert disc < 1e-5          # the discrete eigenfrequency, to 1e-5
        assert cont > 30 * disc     # and measurably off the continuous one


# ---- Courant limit ------------------------------------------------------------------------------

def courant_run(factor, steps=3000):
    """A 20^3-node vacuum box with a soft source; dt set to factor x the 3D Courant limit.
    The E coefficients keep the built dt, as when the reference's pub `dt` is edited
    (sim.rs:176), so the effective Courant number is sqrt(dt_run * dt_built)."""
    grid = F.Grid.uniform([0.0] * 3, [19e-3] * 3, [19] * 3, 0)
    n = grid.dims()
    port = F.PortDef("s", [[F.Edge(2, (10, 10, 10))]], 1e7, False)
    sim = F.Sim.try_new(grid, F.Materials.new(grid), None, [], [port], [], F.Media())
    limit = 1.0 / (C0 * math.sqrt(3.0) / 1e-3)
    dt_run = (factor * limit) ** 2 / sim.dt
    ex = F.Extras(max_steps=steps, decay_db=math.inf, check_every=100)
    try:
        O.run_cpu(sim.desc, 0, F.Pulse(10e9, 10e9), ex, 32, dt=dt_run, fields=False)
        return "stable"
    except O.OracleError as e:
        return str(e)


def test_the_courant_limit_separates_stable_from_unstable():
    assert courant_run(0.99) == "stable"
    assert courant_run(1.02) == "the FDTD run went unstable"


# ---- CPML reflection --------------------------------------------------------------------------

def cpml_reflection(nodes=32, pml=8, scale=3, steps=150):
    """Ez at probes two cells inside the CPML interface on three faces, against the same
    offsets from the source in a `scale` times larger grid, over a window that ends before
    the larger grid's own boundary echo returns. Returns max|a-b|/max|b|."""
    def run(n_nodes):
        grid = F.Grid.uniform([0.0] * 3, [(n_nodes - 1) * 1e-3] * 3, [n_nodes - 1] * 3, pml)
        c = n_nodes // 2
        port = F.PortDef("s", [[F.Edge(2, (c, c, c))]], 1e7, False)
        sim = F.Sim.try_new(grid, F.Materials.new(grid), None, [], [port], [], F.Media())
        reach = nodes // 2 - pml - 3
        probes = [(2, c + reach, c, c), (2, c, c - reach, c), (2, c, c, c + reach), (2, c + reach, c + reach, c)]
        ex = F.Extras(max_steps=steps, decay_db=math.inf, check_every=steps, probes=probes)
        return O.run_cpu(sim.desc, 0, F.Pulse(20e9, 20e9), ex, 32, fields=False).probes.reshape(steps, -1)
    a, b = run(nodes), run(scale * nodes)
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_the_cpml_reflects_less_than_minus_40_db():
    r = cpml_reflection()
    print("CPML reflection", r, 20 * math.log10(r), "dB")
    assert r < 0.01


# ---- Debye limits -----------------------------------------------------------------------------

def debye_limit_errors():
    """Mode 110 of a uniform cavity filled with a one-pole Debye medium (eps_inf 2, delta 2):
    with tau = 0.2 ps it must ring like eps 4, with tau = 100 ns like eps 2."""
    lines = [np.linspace(0.0, L, n + 1) for L, n in zip(SIDES, CELLS)]
    f110 = C0 / 2.0 * math.sqrt((1 / SIDES[0]) ** 2 + (1 / SIDES[1]) ** 2)
    out = []
    for tau, eps in ((0.2e-12, 4.0), (100e-9, 2.0)):
        f = f110 / math.sqrt(eps)
        sim_d, vd = cavity(*lines, debye=(2.0, 2.0, tau), steps=30000, pulse=(f, 0.8 * f))
        sim_p, vp = cavity(*lines, eps=eps, steps=30000, pulse=(f, 0.8 * f))
        pd = spectrum_peak(vd, sim_d.dt, 0.9 * f, 1.1 * f, 801)
        pp = spectrum_peak(vp, sim_p.dt, 0.9 * f, 1.1 * f, 801)
        out.append(abs(pd / pp - 1.0))
    return out


@pytest.mark.slow
def test_debye_media_reach_their_static_and_optical_limits():
    errs = debye_limit_errors()
    print("Debye limit errors", errs)
    assert max(errs) < 5e-4


# ---- dipole: NTFF radiated power vs port input power ------------------------------------------

def dipole_power(nodes=56, pml=8, inset=3, half=6, f=6e9, steps=2500):
    """A centre-fed z dipole (PEC edges either side of a 50-ohm port edge). Returns
    (radiated/p_in, directivity) at f: p_in = Re(V I*)/2 from the port series with the phase
    convention of sims-antenna's finish (crates/sims-antenna/src/fdtd/mod.rs:229-249), radiated
    power from the NTFF accumulators through Far (ntff.rs:61-137)."""
    grid = F.Grid.uniform([0.0] * 3, [(nodes - 1) * 1e-3] * 3, [nodes - 1] * 3, pml)
    c = nodes // 2

    def pec(comp, i, j, k):
        return (comp == 2) & (i == c) & (j == c) & (k >= c - half) & (k <= c + half) & (k != c)

    port = F.PortDef("feed", [[F.Edge(2, (c, c, c))]], 50.0, False)
    sim = F.Sim.try_new(grid, F.Materials.new(grid), pec, [], [port], [], F.Media())
    pt = F.patches(sim, inset)
    ex = F.Extras(max_steps=steps, decay_db=math.inf, freqs=[f], ntff=pt.gpu, check_every=500)
    rec = O.run_cpu(sim.desc, 0, F.Pulse(f, f), ex, 32, fields=False)
    dt = sim.dt
    w = 2 * math.pi * f
    s = np.arange(rec.steps)
    V = np.sum(rec.series[0::2] * np.exp(-1j * w * (s + 1.0) * dt)) * dt
    I = np.sum(rec.series[1::2] * np.exp(-1j * w * (s + 0.5) * dt)) * dt
    p_in = 0.5 * (V * np.conj(I)).real
    far = O.Far.load(pt, rec.ntff, 1, 0, f)
    total, peak = radiated(far, 36)
    return total / p_in, 4 * math.pi * peak / total


def radiated(far, steps):
    """Far::radiated (ntff.rs:122-136), vectorised over directions."""
    k = 2 * math.pi * far.freq / C0
    th = math.pi * (np.arange(steps) + 0.5) / steps
    ph = math.pi * np.arange(2 * steps) / steps
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    phase = far.area[None, :] * np.exp(1j * k * (d @ far.pos.T))
    nvec = phase @ far.j
    lvec = phase @ far.m
    rn = np.sum(nvec * d, -1, keepdims=True)
    perp = nvec - rn * d
    rl = np.cross(d, lvec)
    e = (rl - perp * F.ETA0) * (k / (4 * math.pi))
    u = np.sum(np.abs(e) ** 2, -1) / (2 * F.ETA0)
    total = float(np.sum(u * np.sin(T).ravel()) * (math.pi / steps) ** 2)
    return total, float(u.max())


def test_a_dipole_radiates_what_its_port_delivers():
    ratio, directivity = dipole_power()
    print("radiated/p_in", ratio, "directivity", directivity, 10 * math.log10(directivity), "dBi")
    assert abs(ratio - 1.0) < 0.03
    assert 1.5 <= directivity <= 1.7


# ---- planted mistakes ---------------------------------------------------------------------------

def _fails(mutation, check):
    O.set_mutation(mutation)
    try:
        return check()
    finally:
        O.set_mutation(0)


@pytest.mark.slow
@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5, 7])
def test_planted_mistakes_are_caught(mutation):
    """Each planted mistake (oracle_set_mutation) makes its pinning check fail:
    1 psi dropped, 3 CPML b/c transposed -> CPML reflection; 2 H curl sign -> Courant;
    4 wrong axis spacing table -> Yee dispersion; 5 Debye ka for kb -> Debye limits;
    7 conjugated NTFF E phase -> dipole power balance. Mutations 6 and 8-14 (source half-step,
    plane-DFT phase, Debye relaxation time, capacitor and inductor folding, sheet branches and
    weights) are caught in tests/test_oracle_pins.py."""
    if mutation in (1, 3):
        def check():
            try:
                return cpml_reflection() >= 0.01
            except O.OracleError:
                return True
    elif mutation == 2:
        def check():
            return courant_run(0.99) != "stable"
    elif mutation == 4:
        def check():
            try:
                return max(d for d, _ in dispersion_errors()) >= 1e-4
            except O.OracleError:
                return True
    elif mutation == 5:
        def check():
            try:
                return max(debye_limit_errors()) >= 5e-4
            except O.OracleError:
                return True
    else:
        def check():
            ratio, directivity = dipole_power()
            return abs(ratio - 1.0) >= 0.03 or not (1.5 <= directivity <= 1.7)
    assert _fails(mutation, check)
