"""The sims-antenna caller of the FDTD path, restated for the reference's own antenna validation
(crates/sims-antenna/tests/fdtd.rs:6-55): an axis-aligned metal + lumped-port model is meshed
(crates/sims-antenna/src/fdtd/mesh.rs:6-128 with crates/sims-field/src/grid.rs:66-145), built
(crates/sims-antenna/src/fdtd/engine.rs:44-130), run with the caller's extras
(crates/sims-antenna/src/fdtd/mod.rs:184-227) and post-processed by finish() (mod.rs:229-266).
Only what an Aabb-and-port model uses is restated (no triangle parts, dielectrics or wires).
Test infrastructure: the product's run and far-field calls are what the GPU test measures."""
import bisect
import math
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np

from fieldgrid_amd import fdtd as F

C0 = 299_792_458.0
ETA0 = 376.730_313_668


@dataclass
class Aabb:
    lo: Tuple[float, float, float]
    hi: Tuple[float, float, float]

    @staticmethod
    def new(a, b):
        """mod.rs:20-22"""
        return Aabb(tuple(min(a[i], b[i]) for i in range(3)), tuple(max(a[i], b[i]) for i in range(3)))


@dataclass
class Port:
    a: Tuple[float, float, float]
    b: Tuple[float, float, float]
    r: float


@dataclass
class Model:
    metal: List[Aabb]
    port: Port
    f0: float
    span: float
    fine: float

    def bounds(self):
        """mod.rs:131-146 (metal and port only)"""
        boxes = list(self.metal) + [Aabb.new(self.port.a, self.port.b)]
        lo = tuple(min(b.lo[i] for b in boxes) for i in range(3))
        hi = tuple(max(b.hi[i] for b in boxes) for i in range(3))
        return Aabb(lo, hi)


@dataclass
class Options:
    """mod.rs:170-182"""
    points: int = 121
    max_steps: int = 200_000
    decay_db: float = 50.0
    pml: int = 8
    far_field: bool = True


def dipole():
    """crates/sims-antenna/tests/fdtd.rs:6-22: a 141 mm dipole with a 3 mm gap at 1 GHz."""
    length, g = 0.141, 0.003
    return Model(metal=[Aabb.new((0.0, 0.0, -length / 2), (0.0, 0.0, -g / 2)),
                        Aabb.new((0.0, 0.0, g / 2), (0.0, 0.0, length / 2))],
                 port=Port((0.0, 0.0, -g / 2), (0.0, 0.0, g / 2), 50.0), f0=1.0e9, span=0.2, fine=g)


# ---- grid.rs ---------------------------------------------------------------------------------------

def _dedup(values, tol):
    out = []
    for v in values:
        if out and abs(v - out[-1]) < tol:
            continue
        out.append(v)
    return out


def graded(fixed, features, coarse, ratio, fill=0.0):
    """grid.rs:73-115"""
    fixed = _dedup(sorted(fixed), 1e-9)

    def size(x):
        s = coarse(x)
        for at, sz in features:
            s = min(s, sz + (ratio - 1.0) * abs(x - at))
        return s

    if not fixed:
        return []
    out = [fixed[0]]
    steps = 256
    for a, b in zip(fixed[:-1], fixed[1:]):
        acc = [0.0] * (steps + 1)
        for k in range(steps):
            x = a + (b - a) * (k + 0.5) / steps
            acc[k + 1] = acc[k] + (b - a) / steps / size(x)
        fine = max(math.ceil(acc[steps]), 1)
        n = fine if acc[steps] / fine >= fill else max(math.floor(acc[steps] / fill), 1)
        for c in range(1, n):
            target = acc[steps] * c / n
            k = min(max(bisect.bisect_left(acc, target), 1), steps)
            t = (target - acc[k - 1]) / max(acc[k] - acc[k - 1], 1e-30)
            out.append(a + (b - a) * ((k - 1) + t) / steps)
        out.append(b)
    return out


def smooth(line, limit):
    """grid.rs:117-134"""
    line = list(line)
    for _ in range(200):
        d = [line[i + 1] - line[i] for i in range(len(line) - 1)]
        bad = next((i for i in range(len(d)) if (i > 0 and d[i] > limit * d[i - 1]) or
                    (i + 1 < len(d) and d[i] > limit * d[i + 1])), None)
        if bad is None:
            break
        i = bad
        small = min([d[j] for j in (i - 1, i + 1) if 0 <= j < len(d)], default=math.inf)
        parts = max(math.ceil(d[i] / (limit * small)), 2)
        a, b = line[i], line[i + 1]
        line[i + 1:i + 1] = [a + (b - a) * k / parts for k in range(1, parts)]
    return line


def pad(line, cells):
    """grid.rs:136-145"""
    if len(line) < 2:
        return line
    first, last = line[1] - line[0], line[-1] - line[-2]
    pre = [line[0] - first * i for i in range(cells, 0, -1)]
    post = [line[-1] + last * i for i in range(1, cells + 1)]
    return pre + list(line) + post


def nearest(line, v):
    """grid.rs:48-55"""
    i = bisect.bisect_left(line, v)
    if i >= 1 and i < len(line):
        return i - 1 if abs(v - line[i - 1]) <= abs(line[i] - v) else i
    if i >= 1:
        return i - 1
    return min(i, len(line) - 1)


# ---- mesh.rs, engine.rs -----------------------------------------------------------------------------

def mesh(model, pml):
    """mesh.rs:6-128 for metal boxes and a port"""
    bounds = model.bounds()
    f_max = model.f0 * (1.0 + model.span)
    lam = C0 / f_max
    coarse_air = lam / 20.0
    margin = C0 / model.f0 / 4.0
    fine = model.fine if model.fine > 0.0 else lam / 80.0
    ratio = 1.3
    axes = []
    for a in range(3):
        fixed, features = [], []
        pinned = [model.port.a[a], model.port.b[a]]
        for mi, m in enumerate(model.metal):
            wide = m.hi[a] - m.lo[a] > 2.0 * fine
            for v, inward in ((m.lo[a], 1.0), (m.hi[a], -1.0)):
                shared = any(abs(p - v) < 1e-9 for p in pinned) or any(
                    oi != mi and all(o.lo[t] <= m.hi[t] and o.hi[t] >= m.lo[t] for t in range(3) if t != a)
                    and (abs(o.lo[a] - v) < 1e-9 or abs(o.hi[a] - v) < 1e-9)
                    for oi, o in enumerate(model.metal))
                if wide and not shared:
                    fixed += [v + inward * fine / 3.0, v - inward * 2.0 * fine / 3.0]
                else:
                    fixed.append(v)
                features.append((v, fine))
        p = model.port
        for v in (p.a[a], p.b[a]):
            fixed.append(v)
            features.append((v, fine))
        length = abs(p.b[a] - p.a[a])
        if length > 0.0:
            features.append((0.5 * (p.a[a] + p.b[a]), min(length / 2.0, fine * 2.0)))
        fixed += [bounds.lo[a] - margin, bounds.hi[a] + margin]
        fixed = _dedup(sorted(fixed), 1e-7)
        axes.append(pad(smooth(graded(fixed, features, lambda x: coarse_air, ratio, 0.0), 1.5), pml))
    return F.Grid(np.array(axes[0]), np.array(axes[1]), np.array(axes[2]), pml)


def metal_edges(grid, metal):
    """engine.rs:43-73: every edge whose both ends lie in a metal box is PEC."""
    n = grid.dims()
    masks = [np.zeros(n, bool) for _ in range(3)]
    for m in metal:
        r = []
        for a in range(3):
            ax = list(grid.axis(a))
            r.append((bisect.bisect_left(ax, m.lo[a] - 1e-9), bisect.bisect_right(ax, m.hi[a] + 1e-9)))
        for c in range(3):
            lim = list(r)
            lim[c] = (lim[c][0], max(lim[c][1] - 1, 0))
            masks[c][lim[0][0]:lim[0][1], lim[1][0]:lim[1][1], lim[2][0]:lim[2][1]] = True
    return np.stack([mk.ravel() for mk in masks])


def port_def(grid, p):
    """engine.rs:75-92"""
    spans = [abs(p.b[t] - p.a[t]) for t in range(3)]
    axis = max(range(3), key=lambda t: (spans[t], t))  # max_by keeps the last of equal maxima
    at = [nearest(list(grid.axis(t)), min(p.a[t], p.b[t])) for t in range(3)]
    end = nearest(list(grid.axis(axis)), max(p.a[axis], p.b[axis]))
    count = max(end - at[axis], 1)
    edges = []
    for s in range(count):
        q = list(at)
        q[axis] += s
        edges.append(F.Edge(axis, tuple(q)))
    return F.PortDef("feed", [edges], p.r, False)


@dataclass
class Setup:
    """mod.rs:184-211"""
    sim: F.Sim
    patches: Optional[F.Patches]
    freqs: List[float]
    f0: float
    fc: float
    grid: F.Grid


def setup(model, opts=Options()):
    grid = mesh(model, opts.pml)
    sim = F.Sim.try_new(grid, F.Materials.new(grid), metal_edges(grid, model.metal), [], [port_def(grid, model.port)],
                        [], F.Media())
    f0 = model.f0
    freqs = [f0 * (1.0 - model.span + 2.0 * model.span * i / max(opts.points - 1, 1)) for i in range(opts.points)]
    patches = F.patches(sim, 3) if opts.far_field else None
    return Setup(sim, patches, freqs, f0, f0 * 0.5, grid)


def extras(s, opts=Options()):
    """mod.rs:213-227"""
    return F.Extras(max_steps=opts.max_steps, decay_db=opts.decay_db, freqs=[s.f0] if s.patches else [],
                    ntff=s.patches.gpu if s.patches else None, check_every=500)


def finish_host(s, series, dt):
    """finish() (mod.rs:229-266) on the host: the f64 phasor recursion of the port series,
    Z = V/I over the sweep and p_in = Re(V I*)/2 at f0. Returns (freqs, z, p_in)."""
    ser = np.asarray(series, np.float64).reshape(-1, 2)
    allf = list(s.freqs) + [s.f0]
    zs = []
    for f in allf:
        w = 2 * math.pi * f
        rot = complex(math.cos(-w * dt), math.sin(-w * dt))
        pe = dt * rot
        ph = dt * complex(math.cos(-w * 0.5 * dt), math.sin(-w * 0.5 * dt))
        v = i = 0j
        for a, b in ser:
            v += pe * a
            i += ph * b
            pe *= rot
            ph *= rot
        zs.append((v, i))
    z = np.array([v / i for v, i in zs])
    v0, i0 = zs[-1]
    return list(s.freqs), z[:-1], 0.5 * (v0 * np.conj(i0)).real


def radiated(intensity):
    """crates/sims-antenna/tests/fdtd.rs:24-39: total radiated power and peak intensity over a
    60 x 90 grid of directions; `intensity` maps an (n, 3) array of directions to W/sr."""
    nt, npp = 60, 90
    th = (np.arange(nt) + 0.5) / nt * math.pi
    ph = (np.arange(npp) + 0.5) / npp * 2.0 * math.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    u = np.asarray(intensity(d), np.float64).reshape(nt, npp)
    total = float(np.sum(u * np.sin(T)) * (math.pi / nt) * (2.0 * math.pi / npp))
    return total, float(u.max())


def resonance(freqs, z):
    """crates/sims-antenna/tests/fdtd.rs:41-44: the first frequency where Im Z crosses zero upward."""
    for q in range(len(freqs) - 1):
        if z[q].imag < 0.0 and z[q + 1].imag >= 0.0:
            return freqs[q]
    return 0.0
