"""Seeded synthetic FDTD workloads (problem builders shared by the benchmark and the tests).

  plain(...)       64^3-node vacuum, 1 mm, CPML 10, soft Ez point source, Ez/Hy point probes
  boxes(...)       256^3, 0.5 mm, CPML 12, 40 dielectric boxes (10 Debye), one 50-ohm z-port,
                   NTFF box 4 cells inside the CPML at 2, 5, 10 GHz
  layered(...)     768x768x384, 1 mm, CPML 16, eps_r ~ U[1,4] per 8-cell z-layer, two ports
  sweep_runs(...)  the batched sweep's per-run problems (uneven grids, 1-4 ports)
  design(...)      the adjoint inverse-design problem (192x192x112, design box 64x64x16 cells)
Every builder takes a size scale so tests can run the same recipe at oracle-sized grids.
"""
import math

import numpy as np

from . import fdtd as F


def _grid(nodes, dx, pml):
    cells = [n - 1 for n in nodes]
    return F.Grid.uniform([0.0] * 3, [c * dx for c in cells], cells, pml)


def plain(nodes=64, dx=1e-3, pml=10, seed=1, nprobes=8):
    """A vacuum grid of nodes^3 with a 1e7-ohm single-edge z-port at the centre as a soft
    source (the reference's idiom, crates/sims-field/tests/fdtd.rs:206-211) and Ez, Hy
    point probes at `nprobes` seeded nodes outside the CPML."""
    grid = _grid([nodes] * 3, dx, pml)
    n = grid.dims()
    c = n[0] // 2
    port = F.PortDef("src", [[F.Edge(2, (c, c, c))]], 1e7, False)
    sim = F.Sim.try_new(grid, F.Materials.new(grid), None, [], [port], [], F.Media())
    rng = np.random.default_rng(seed)
    at = rng.integers(pml + 1, n[0] - pml - 1, size=(nprobes, 3))
    probes = [(2, *map(int, p)) for p in at] + [(4, *map(int, p)) for p in at]
    return sim, probes


def plain_pulse():
    """10 GHz centre, 5 GHz -20 dB half-bandwidth (sim.rs:646-657)."""
    return F.Pulse(10e9, 5e9)


def _box_materials(grid, rng, count, debye_count, lo, hi, max_size, per_box_tau):
    """Axis-aligned boxes with eps_r ~ U[1,10], sigma ~ U[0,0.05]; the first `debye_count`
    are Debye media with delta_eps ~ U[1,5], tau ~ U[5,50] ps (eps_inf = eps_r)."""
    n = grid.dims()
    cells = (n[0] - 1, n[1] - 1, n[2] - 1)
    mats = F.Materials.new(grid)
    eps = mats.eps.reshape(cells)
    sig = mats.sigma.reshape(cells)
    inf = mats.inf.reshape(cells)
    dd = mats.debye.reshape(cells)
    dset = np.zeros(cells, np.uint8)
    taus = []
    for b in range(count):
        size = rng.integers(4, max_size + 1, size=3)
        a = [int(rng.integers(lo, hi - size[q])) for q in range(3)]
        sl = tuple(slice(a[q], a[q] + int(size[q])) for q in range(3))
        er = rng.uniform(1.0, 10.0)
        s = rng.uniform(0.0, 0.05)
        eps[sl], sig[sl], inf[sl] = er, s, er
        if b < debye_count:
            dd[sl] = rng.uniform(1.0, 5.0)
            tau = rng.uniform(5e-12, 50e-12)
            taus.append(tau)
            dset[sl] = b + 1 if per_box_tau else 0
    return mats, dset, taus


def boxes(nodes=256, dx=0.5e-3, pml=12, seed=2, per_box_tau=False, nboxes=40, ndebye=10):
    """The boxes workload. With per_box_tau=False every Debye box shares one pole at the median
    tau (the reference's single global pole table, run.rs:296-301); with True each Debye box
    keeps its own tau through the per-cell pole-set extension. Returns (sim, extras-kwargs)."""
    grid = _grid([nodes] * 3, dx, pml)
    n = grid.dims()
    rng = np.random.default_rng(seed)
    inset = 4
    lo, hi = pml + inset + 1, n[0] - 1 - pml - inset - 1
    mats, dset, taus = _box_materials(grid, rng, nboxes, ndebye, lo, hi, max(4, (hi - lo) // 6), per_box_tau)
    if per_box_tau:
        sets = [F.Debye([])] + [F.Debye([(1.0 / t, 1.0)]) for t in taus]
        mats.debye_set = dset.ravel()
        media = F.Media(debye=sets[0], pole_sets=sets)
    else:
        media = F.Media(debye=F.Debye([(1.0 / float(np.median(taus)), 1.0)] if taus else []))
    c = n[0] // 2
    # one 50-ohm z-port: one column of 4 edges, in vacuum near the centre
    port = F.PortDef("p1", [[F.Edge(2, (c, c - 3 * nodes // 16, k)) for k in range(c - 2, c + 2)]], 50.0, False)
    sim = F.Sim.try_new(grid, mats, None, [], [port], [], media)
    return sim


def boxes_pulse():
    """0-20 GHz Gaussian pulse: 10 GHz centre, 10 GHz -20 dB half-bandwidth."""
    return F.Pulse(10e9, 10e9)


def boxes_extras(sim, steps=2000, inset=4, freqs=(2e9, 5e9, 10e9), **kw):
    return F.Extras(max_steps=steps, decay_db=math.inf, freqs=list(freqs), ntff=F.patches(sim, inset).gpu,
                    check_every=kw.pop("check_every", 500), **kw)


def layered(dims=(768, 768, 384), dx=1e-3, pml=16, seed=3, layer=8, port_i=(40, 52)):
    """The single-run throughput workload: eps_r ~ U[1,4] per `layer`-cell z-layer, two
    50-ohm single-column z-ports at x-index <= 55 (so the reference's f32 port indices,
    run.rs:341-343, could still express them)."""
    grid = _grid(list(dims), dx, pml)
    n = grid.dims()
    cells = (n[0] - 1, n[1] - 1, n[2] - 1)
    rng = np.random.default_rng(seed)
    nl = (cells[2] + layer - 1) // layer
    er = rng.uniform(1.0, 4.0, size=nl).astype(np.float32)
    mats = F.Materials.new(grid)
    mats.eps = np.ascontiguousarray(np.broadcast_to(np.repeat(er, layer)[: cells[2]], cells).ravel())
    mats.inf = mats.eps
    kc = n[2] // 2
    ports = [F.PortDef(f"p{q + 1}", [[F.Edge(2, (port_i[q], n[1] * (q + 1) // 3, k)) for k in range(kc - 2, kc + 2)]],
                       50.0, False) for q in range(2)]
    return F.Sim.try_new(grid, mats, None, [], ports, [], F.Media())


def layered_pulse():
    return F.Pulse(10e9, 10e9)


def sweep_runs(count=96, seed=4, lo=48, hi=200, dx=0.5e-3, pml=8, nboxes=8, ndebye=2, steps=(800, 3000),
               nfreq=64):
    """The batched sweep: `count` independent runs, each with dims ~ U[lo, hi] per axis, CPML
    `pml`, `nboxes` dielectric boxes as in boxes() (eps_r ~ U[1,10], sigma ~ U[0,0.05], the first
    `ndebye` of them Debye with one pole at their median tau), 1-4 single-column 50-ohm z-ports
    of which one is driven, steps ~ U[steps], and port V/I spectra at `nfreq` frequencies over
    the 0-20 GHz pulse band. Returns a list of fdtd.BatchRun."""
    rng = np.random.default_rng(seed)
    pulse = boxes_pulse()
    freqs = [(q + 1) * 20e9 / nfreq for q in range(nfreq)]
    runs = []
    for _ in range(count):
        dims = [int(v) for v in rng.integers(lo, hi + 1, size=3)]
        grid = _grid(dims, dx, pml)
        n = grid.dims()
        cells = (n[0] - 1, n[1] - 1, n[2] - 1)
        mats = F.Materials.new(grid)
        eps = mats.eps.reshape(cells)
        sig = mats.sigma.reshape(cells)
        inf = mats.inf.reshape(cells)
        dd = mats.debye.reshape(cells)
        taus = []
        for b in range(nboxes):
            a, e = [], []
            for q in range(3):
                span = cells[q] - 2 * (pml + 2)
                size = int(rng.integers(2, max(3, span // 4) + 1))
                start = int(rng.integers(pml + 2, pml + 2 + max(1, span - size)))
                a.append(start), e.append(start + size)
            sl = tuple(slice(a[q], e[q]) for q in range(3))
            er, s = rng.uniform(1.0, 10.0), rng.uniform(0.0, 0.05)
            eps[sl], sig[sl], inf[sl] = er, s, er
            if b < ndebye:
                dd[sl] = rng.uniform(1.0, 5.0)
                taus.append(rng.uniform(5e-12, 50e-12))
        media = F.Media(debye=F.Debye([(1.0 / float(np.median(taus)), 1.0)] if taus else []))
        nports = int(rng.integers(1, 5))
        kc = n[2] // 2
        ports = []
        for q in range(nports):
            i = int(rng.integers(pml + 2, n[0] - pml - 2))
            j = int(rng.integers(pml + 2, n[1] - pml - 2))
            ports.append(F.PortDef(f"p{q + 1}", [[F.Edge(2, (i, j, k)) for k in range(kc - 2, kc + 2)]], 50.0, False))
        sim = F.Sim.try_new(grid, mats, None, [], ports, [], media)
        driven = int(rng.integers(0, nports))
        k = int(rng.integers(steps[0], steps[1] + 1))
        runs.append(F.BatchRun(sim, driven, pulse, F.Extras(max_steps=k, decay_db=math.inf), freqs))
    return runs


def design(dims=(192, 192, 112), dx=0.5e-3, pml=12, box=((64, 128), (64, 128), (48, 64)), port=(30, 96),
           probe=(162, 96), steps=3000, freq=8e9, seed=5, sigma=0.0):
    """The inverse-design problem: vacuum with a design box of cells whose eps_r ~ U[1,4] (seed),
    a 50-ohm single-column z-port at (port[0], port[1]) on one side, driven by a Gaussian pulse
    f0 = fc = `freq`, and the observed Ez at node (probe[0], probe[1]) at mid-height on the other
    side. `sigma` > 0 gives the box cells a conductivity ~ U[0, sigma] too. Returns a dict of
    fdtd.Design arguments plus the objective frequency."""
    grid = _grid(list(dims), dx, pml)
    n = grid.dims()
    cells = (n[0] - 1, n[1] - 1, n[2] - 1)
    rng = np.random.default_rng(seed)
    mats = F.Materials.new(grid)
    lo = tuple(b[0] for b in box)
    hi = tuple(b[1] for b in box)
    shape = tuple(h - l for l, h in zip(lo, hi))
    eps_box = rng.uniform(1.0, 4.0, size=shape).astype(np.float32)
    sl = tuple(slice(l, h) for l, h in zip(lo, hi))
    mats.eps.reshape(cells)[sl] = eps_box
    mats.inf.reshape(cells)[sl] = eps_box
    if sigma > 0:
        mats.sigma.reshape(cells)[sl] = rng.uniform(0.0, sigma, size=shape).astype(np.float32)
    kc = n[2] // 2
    ports = [F.PortDef("p1", [[F.Edge(2, (port[0], port[1], k)) for k in range(kc - 2, kc + 2)]], 50.0, False)]
    return dict(grid=grid, mats=mats, pec=None, lumped=[], ports=ports, resistive=[], media=F.Media(),
                lo=lo, hi=hi, obs=[(2, probe[0], probe[1], kc)], driven=0, pulse=F.Pulse(freq, freq),
                steps=steps), freq


def design_small(steps=300, sigma=0.02, seed=5, media=False):
    """The design recipe at oracle size: 30x28x26 nodes, 1 mm, CPML 5, a 6x5x4-cell box. With
    `media`, the wave from the port also crosses a block of Debye medium between the box and the
    observation (eps_inf 2, delta 1.5, tau 20 ps), a 3 nH lumped inductor over two z edges and a
    thin z wire of 0.1 mm radius over four edges, all outside the box."""
    args, freq = design(dims=(30, 28, 26), dx=1e-3, pml=5, box=((11, 17), (11, 16), (10, 14)), port=(7, 14),
                        probe=(23, 13), steps=steps, freq=8e9, seed=seed, sigma=sigma)
    if media:
        grid, mats = args["grid"], args["mats"]
        n = grid.dims()
        cells = (n[0] - 1, n[1] - 1, n[2] - 1)
        sl = (slice(18, 22), slice(9, 19), slice(8, 18))
        mats.eps.reshape(cells)[sl] = 3.5
        mats.inf.reshape(cells)[sl] = 2.0
        mats.debye.reshape(cells)[sl] = 1.5
        args["media"] = F.Media(debye=F.Debye([(1.0 / 20e-12, 1.0)]),
                                wires=[(F.Edge(2, (9, 21, k)), 0.1e-3) for k in range(11, 15)])
        args["lumped"] = [F.Lumped("l", [F.Edge(2, (9, 8, k)) for k in (12, 13)], F.Element.Inductor(3e-9))]
    return args, freq


def dft_objective(series, dt, freq):
    """J = |D|^2 with D = sum_n s_n dt e^{-i w (n+1) dt} (the E-field phase of the reference's
    running DFTs, cpu.rs:350-353) over one observation's series s_n (E after step n), and
    dJ/ds_n = 2 Re(conj(D) dt e^{-i w (n+1) dt}). Returns (J, dJ/ds)."""
    s = np.asarray(series, np.float64).reshape(-1)
    w = dt * np.exp(-1j * 2.0 * np.pi * freq * (np.arange(s.size) + 1.0) * dt)
    D = np.sum(s * w)
    return float(abs(D) ** 2), 2.0 * np.real(np.conj(D) * w)


def bytes_per_node_step(sim, extras=None, core=False):
    """Algorithmic HBM bytes of one step per node: 24 B of E/H read + 24 B written + 6 B of
    u16 coefficient-set indices (12 B when wide) + 8 B per CPML psi element, plus (unless
    `core`) the sparse extras (Debye edges, NTFF patches). Ports and probes are negligible."""
    n = sim.dims()
    nn = n[0] * n[1] * n[2]
    psi = 0
    if sim.pml:
        for c in range(3):
            for a in range(3):
                if a != c:
                    psi += F.psi_len(n, a, sim.pml)
    psi *= 2  # E and H stores
    p = sim.problem()
    per = 48.0 + (12.0 if p.coef_wide else 6.0) + 8.0 * psi / nn
    if core:
        return per
    if sim.debye_edges:
        per += sim.debye_edges * (8.0 * p.n_poles + 8.0) / nn
    if extras is not None and extras.ntff is not None:
        nf = min(len(extras.freqs), 4)
        per += (len(extras.ntff) // 13) * (48.0 + 96.0 * nf) / nn
    return per
