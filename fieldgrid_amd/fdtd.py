"""Python mirror of sims-field's FDTD API (crates/sims-field/src/fdtd/mod.rs:1-11).

Same names, argument meanings and error behaviour as the reference's Rust API:
  Grid, Materials, Edge, Element, Lumped, PortDef, Resistive, Debye, Sheets, Media, Sim,
  Extras, Pulse, Record, run / run_blocking, patches, Patches, Far, time_step, gaussian,
  pulse_shape, idx, unidx, psi_len, and the constants C0, EPS0, MU0, ETA0.
`Sim.try_new` runs the C++ builder in libfieldgrid.so; `run` runs the HIP engine on an
MI355X through the C-ABI of include/fieldgrid.h. Beyond the reference: `run_batch` (many runs
in shared launches, include/fieldgrid_batch.h) with `BatchRun`, `Tune` and `port_spectra`;
`Design` (the discrete adjoint, include/fieldgrid_adjoint.h). Errors raise FieldgridError carrying the
reference's message ("cancelled", "the FDTD run went unstable", ...).

Two differences of form, not of meaning:
  * `pec` may be a boolean mask shaped (3, n0, n1, n2) or a vectorised callable
    pec(c, i, j, k) over numpy index arrays (the Rust closure is called per edge);
  * `run` takes `gpu=None` as an error, not a CPU fallback: the reference's CPU path,
    run_cpu (cpu.rs:424-482), stays with the reference.
"""
import ctypes as C
import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import abi

C0 = 299_792_458.0
EPS0 = 8.854_187_812_8e-12
MU0 = 1.256_637_062_12e-6
ETA0 = 376.730_313_668


# ---- grid (grid.rs:1-46) -------------------------------------------------------------------

def uniform(lo: float, hi: float, cells: int) -> np.ndarray:
    """grid.rs:44-46"""
    return np.array([lo + (hi - lo) * i / cells for i in range(cells + 1)], dtype=np.float64)


@dataclass
class Grid:
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    pml: int = 0

    def __post_init__(self):
        self.x = np.ascontiguousarray(self.x, dtype=np.float64)
        self.y = np.ascontiguousarray(self.y, dtype=np.float64)
        self.z = np.ascontiguousarray(self.z, dtype=np.float64)

    @staticmethod
    def uniform(lo, hi, n, pml):
        """Grid::uniform (grid.rs:10-13): n cells per axis."""
        return Grid(uniform(lo[0], hi[0], n[0]), uniform(lo[1], hi[1], n[1]),
                    uniform(lo[2], hi[2], n[2]), pml)

    def axis(self, a: int) -> np.ndarray:
        return (self.x, self.y, self.z)[a]

    def dims(self) -> Tuple[int, int, int]:
        return (len(self.x), len(self.y), len(self.z))

    def cells(self) -> int:
        return (len(self.x) - 1) * (len(self.y) - 1) * (len(self.z) - 1)


def idx(n, i, j, k):
    """sim.rs:8-10"""
    return (i * n[1] + j) * n[2] + k


def unidx(n, id_):
    """sim.rs:642-644"""
    return (id_ // (n[1] * n[2]), (id_ // n[2]) % n[1], id_ % n[2])


def psi_len(n, axis, pml):
    """sim.rs:139-143"""
    if pml == 0:
        return 0
    m = list(n)
    m[axis] = 2 * pml
    return m[0] * m[1] * m[2]


def time_step(grid: Grid) -> float:
    """sim.rs:133-137"""
    s = 0.0
    for line in (grid.x, grid.y, grid.z):
        mn = min(line[i + 1] - line[i] for i in range(len(line) - 1))
        s += 1.0 / (mn * mn)
    return 0.99 / (C0 * math.sqrt(s))


def pulse_shape(fc: float) -> Tuple[float, float]:
    """sim.rs:646-649: (t0, tau)"""
    tau = math.sqrt(math.log(10.0)) / (math.pi * fc)
    return (4.0 * tau, tau)


def gaussian(f0: float, fc: float) -> Callable[[float], float]:
    """sim.rs:651-657"""
    t0, tau = pulse_shape(fc)

    def f(t):
        x = (t - t0) / tau
        return math.exp(-x * x) * math.cos(2.0 * math.pi * f0 * (t - t0))

    return f


# ---- problem description (sim.rs:145-301) --------------------------------------------------

@dataclass
class Edge:
    comp: int
    at: Tuple[int, int, int]


@dataclass
class Element:
    kind: int
    value: float

    @staticmethod
    def Resistor(v):
        return Element(abi.FG_RESISTOR, v)

    @staticmethod
    def Capacitor(v):
        return Element(abi.FG_CAPACITOR, v)

    @staticmethod
    def Inductor(v):
        return Element(abi.FG_INDUCTOR, v)


@dataclass
class Lumped:
    name: str
    edges: List[Edge]
    element: Element


@dataclass
class PortDef:
    name: str
    columns: List[List[Edge]]
    r: float
    reference_above: bool = False


@dataclass
class Resistive:
    comp: int
    id: int
    r: float
    adaptive: bool
    faces: Tuple[int, int]


@dataclass
class Debye:
    """sim.rs:243-283"""
    poles: List[Tuple[float, float]] = field(default_factory=list)

    @staticmethod
    def band(f_lo: float, f_hi: float) -> "Debye":
        x0, x1 = 2.0 * math.pi * f_lo / 30.0, 2.0 * math.pi * f_hi * 30.0
        span = math.log(x1 / x0)
        n = max(math.ceil(span / 0.7), 1)
        h = span / n
        poles = []
        for k in range(n + 1):
            w = 0.5 * h if k in (0, n) else h
            poles.append((x0 * math.exp(k * h), w / span))
        return Debye(poles)

    def response(self, w: float) -> Tuple[float, float]:
        re = im = 0.0
        for x, a in self.poles:
            r = w / x
            re += a / (1.0 + r * r)
            im += a * r / (1.0 + r * r)
        return (re, im)

    def fit(self, er: float, tan: float, f: float) -> Tuple[float, float]:
        if tan <= 0.0 or not self.poles:
            return (er, 0.0)
        s1, s2 = self.response(2.0 * math.pi * f)
        delta = er * tan / s2
        return (er - delta * s1, delta)


@dataclass
class Sheets:
    table: List[float] = field(default_factory=lambda: [0.0] * 4)
    states: int = 0


@dataclass
class Media:
    sheets: Sheets = field(default_factory=Sheets)
    debye: Debye = field(default_factory=Debye)
    wires: List[Tuple[Edge, float]] = field(default_factory=list)
    # extension: several pole sets, chosen per cell by Materials.debye_set
    pole_sets: Optional[List[Debye]] = None


class Materials:
    """sim.rs:217-222, 285-294: per-cell eps_r, sigma, eps_inf, delta_eps (cell id
    (i*(n1-1)+j)*(n2-1)+k). `debye_set` (extension) picks a Media.pole_sets entry per cell."""

    def __init__(self, grid: Grid):
        cells = grid.cells()
        self.eps = np.ones(cells, np.float32)
        self.sigma = np.zeros(cells, np.float32)
        self.inf = np.ones(cells, np.float32)
        self.debye = np.zeros(cells, np.float32)
        self.debye_set: Optional[np.ndarray] = None

    @staticmethod
    def new(grid: Grid) -> "Materials":
        return Materials(grid)


def _edges(edges: Sequence[Edge]):
    arr = (abi.fg_edge * max(len(edges), 1))()
    for q, e in enumerate(edges):
        arr[q].comp = e.comp
        arr[q].at[:] = [int(v) for v in e.at]
    return arr


def _ptr(a: Optional[np.ndarray], ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    return a.ctypes.data_as(C.POINTER(ctype))


class SimDesc:
    """The fg_sim_desc of one problem, with every array it points at kept alive."""

    def __init__(self, grid, mats, pec, lumped, ports, resistive, media):
        self.keep = []
        d = abi.fg_sim_desc()
        n = grid.dims()
        self.n = n
        for a in range(3):
            line = grid.axis(a)
            self.keep.append(line)
            d.line[a] = line.ctypes.data_as(C.POINTER(C.c_double))
            d.len[a] = len(line)
        d.pml = grid.pml
        for name in ("eps", "sigma", "inf", "debye"):
            arr = np.ascontiguousarray(getattr(mats, name), dtype=np.float32)
            if arr.size != grid.cells():
                raise ValueError(f"Materials.{name} has {arr.size} values for {grid.cells()} cells")
            self.keep.append(arr)
            setattr(d, name, _ptr(arr, C.c_float))
        mask = _pec_mask(pec, n)
        if mask is not None:
            self.keep.append(mask)
        d.pec = _ptr(mask, C.c_uint8)
        lum = (abi.fg_lumped * max(len(lumped), 1))()
        for q, l in enumerate(lumped):
            arr = _edges(l.edges)
            self.keep.append(arr)
            lum[q].kind = l.element.kind
            lum[q].value = l.element.value
            lum[q].n_edges = len(l.edges)
            lum[q].edges = C.cast(arr, C.POINTER(abi.fg_edge))
        self.keep.append(lum)
        d.n_lumped, d.lumped = len(lumped), C.cast(lum, C.POINTER(abi.fg_lumped))
        pds = (abi.fg_port_def * max(len(ports), 1))()
        for q, p in enumerate(ports):
            lens = np.array([len(c) for c in p.columns], np.uint32)
            arr = _edges([e for c in p.columns for e in c])
            self.keep += [lens, arr]
            pds[q].r = p.r
            pds[q].reference_above = int(p.reference_above)
            pds[q].n_columns = len(p.columns)
            pds[q].column_len = _ptr(lens, C.c_uint32)
            pds[q].edges = C.cast(arr, C.POINTER(abi.fg_edge))
        self.keep.append(pds)
        d.n_ports, d.ports = len(ports), C.cast(pds, C.POINTER(abi.fg_port_def))
        res = (abi.fg_resistive * max(len(resistive), 1))()
        for q, r in enumerate(resistive):
            res[q].comp, res[q].adaptive, res[q].id, res[q].r = r.comp, int(r.adaptive), r.id, r.r
            res[q].faces[:] = list(r.faces)
        self.keep.append(res)
        d.n_resistive, d.resistive = len(resistive), C.cast(res, C.POINTER(abi.fg_resistive))
        wires = (abi.fg_wire * max(len(media.wires), 1))()
        for q, (e, radius) in enumerate(media.wires):
            wires[q].edge.comp = e.comp
            wires[q].edge.at[:] = list(e.at)
            wires[q].radius = radius
        self.keep.append(wires)
        d.n_wires, d.wires = len(media.wires), C.cast(wires, C.POINTER(abi.fg_wire))
        sets = media.pole_sets if media.pole_sets else [media.debye]
        poles = np.array([v for s in sets for p in s.poles for v in p] or [0.0], np.float64)
        first = np.cumsum([0] + [len(s.poles) for s in sets]).astype(np.uint32)
        self.keep += [poles, first]
        d.n_poles = int(first[-1]) if len(sets) > 1 else len(sets[0].poles)
        d.n_pole_sets = len(sets)
        d.poles = _ptr(poles, C.c_double)
        d.set_first = _ptr(first, C.c_uint32) if len(sets) > 1 else _ptr(None, C.c_uint32)
        if media.pole_sets and mats.debye_set is not None:
            ds = np.ascontiguousarray(mats.debye_set, dtype=np.uint8)
            self.keep.append(ds)
            d.debye_set = _ptr(ds, C.c_uint8)
        table = np.array(media.sheets.table, np.float32)
        self.keep.append(table)
        d.sheet_states, d.sheet_table_len = media.sheets.states, len(table)
        d.sheet_table = _ptr(table, C.c_float)
        self.desc = d


def _pec_mask(pec, n):
    if pec is None:
        return None
    if isinstance(pec, np.ndarray):
        m = np.ascontiguousarray(pec.reshape(3, *n), dtype=np.uint8)
        return m if m.any() else None
    c, i, j, k = np.meshgrid(np.arange(3), np.arange(n[0]), np.arange(n[1]), np.arange(n[2]),
                             indexing="ij")
    m = np.ascontiguousarray(np.broadcast_to(pec(c, i, j, k), c.shape), dtype=np.uint8)
    return m if m.any() else None


class Sim:
    """sim.rs:173-189; built by the C++ builder (Sim::try_new, sim.rs:352-554)."""

    def __init__(self, grid, desc: SimDesc, handle, ports):
        self.grid = grid
        self.desc = desc
        self._h = handle
        self.ports = ports
        self.pml = grid.pml
        L = abi.lib()
        self.dt = L.fg_sim_dt(handle)
        counts = (C.c_uint64 * 4)()
        L.fg_sim_counts(handle, counts)
        self.debye_edges, self.inductors, self.sheets, self.wires = (int(v) for v in counts)
        self._problem = None

    @staticmethod
    def try_new(grid, mats, pec, lumped, ports, resistive=(), media=None) -> "Sim":
        media = media or Media()
        desc = SimDesc(grid, mats, pec, list(lumped), list(ports), list(resistive), media)
        h = C.c_void_p()
        abi.check(abi.lib().fg_sim_new(C.byref(desc.desc), C.byref(h)))
        return Sim(grid, desc, h, list(ports))

    new = try_new

    def __del__(self):
        if getattr(self, "_h", None):
            abi.lib().fg_sim_free(self._h)
            self._h = None

    def dims(self):
        return self.grid.dims()

    def coefficients(self):
        """(ca, cb) as (3, nn) float32 arrays (sim.rs:177-178)."""
        nn = int(np.prod(self.dims()))
        L = abi.lib()
        ca = np.ctypeslib.as_array(L.fg_sim_ca(self._h), shape=(3 * nn,)).copy()
        cb = np.ctypeslib.as_array(L.fg_sim_cb(self._h), shape=(3 * nn,)).copy()
        return ca.reshape(3, nn), cb.reshape(3, nn)

    def problem(self, narrow: int = 0) -> abi.fg_problem:
        """pack() (run.rs:205-376): the packed device tables, owned by this Sim."""
        if self._problem is not None and self._problem[0] == narrow:
            return self._problem[1]
        p = abi.fg_problem()
        abi.check(abi.lib().fg_sim_pack(self._h, narrow, C.byref(p)))
        self._problem = (narrow, p)
        return p


# ---- run (run.rs:34-59, 378-758) -----------------------------------------------------------

@dataclass
class Extras:
    min_steps: int = 0
    max_steps: int = 0
    decay_db: float = 0.0
    freqs: List[float] = field(default_factory=list)
    plane_k: Optional[int] = None
    ntff: Optional[np.ndarray] = None
    fused_e: bool = False
    check_every: int = 0
    profile: bool = False
    probes: List[Tuple[int, int, int, int]] = field(default_factory=list)  # (field, i, j, k)
    engine: Optional[int] = None  # FG_ENGINE_*; None follows fused_e

    def to_c(self, keep: list) -> abi.fg_extras:
        x = abi.fg_extras()
        x.min_steps, x.max_steps, x.decay_db = self.min_steps, self.max_steps, self.decay_db
        x.nf = min(len(self.freqs), abi.FG_MAX_FREQS)
        for q in range(x.nf):
            x.freqs[q] = self.freqs[q]
        x.plane_k = -1 if self.plane_k is None else self.plane_k
        if self.ntff is not None and len(self.ntff):
            arr = np.ascontiguousarray(self.ntff, dtype=np.uint32)
            keep.append(arr)
            x.n_patches = len(arr) // 13
            x.ntff = arr.ctypes.data_as(C.POINTER(C.c_uint32))
        x.check_every = self.check_every
        x.profile = int(self.profile)
        if self.engine is not None:
            x.engine = self.engine
        else:
            x.engine = abi.FG_ENGINE_FUSED if self.fused_e else abi.FG_ENGINE_AUTO
        if self.probes:
            pr = (abi.fg_point_probe * len(self.probes))()
            for q, (f, i, j, k) in enumerate(self.probes):
                pr[q].field, pr[q].i, pr[q].j, pr[q].k = f, i, j, k
            keep.append(pr)
            x.n_probes = len(self.probes)
            x.probes = C.cast(pr, C.POINTER(abi.fg_point_probe))
        return x


@dataclass
class Pulse:
    f0: float
    fc: float


@dataclass
class Record:
    steps: int
    ports: int
    series: np.ndarray
    decay_db: float
    plane: np.ndarray
    ntff: np.ndarray
    probes: np.ndarray = None
    e: Optional[np.ndarray] = None
    h: Optional[np.ndarray] = None


class Gpu:
    """A HIP device handle (replaces `Gpu`, gpu.rs:36-41)."""

    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        abi.check(abi.lib().fg_open(ordinal, C.byref(h)))
        self._h = h
        self.name = abi.lib().fg_name(h).decode()

    open = staticmethod(lambda ordinal=0: Gpu(ordinal))
    open_blocking = open

    def max_cells(self) -> int:
        return abi.lib().fg_max_cells(self._h)

    def synchronize(self):
        abi.check(abi.lib().fg_device_synchronize(self._h))

    def stream_copy(self, nbytes: int = 1 << 30, iterations: int = 20) -> float:
        """Achieved read+write GB/s of a plain device copy (the bandwidth probe)."""
        v = C.c_double()
        abi.check(abi.lib().fg_stream_copy(self._h, nbytes, iterations, C.byref(v)))
        return v.value

    def __del__(self):
        if getattr(self, "_h", None):
            abi.lib().fg_close(self._h)
            self._h = None


def _record_buffers(sim: Sim, extras: Extras, fields: bool):
    n = sim.dims()
    nn = n[0] * n[1] * n[2]
    nf = min(len(extras.freqs), abi.FG_MAX_FREQS)
    ports = len(sim.ports)
    bufs = {
        "series": np.zeros(max(extras.max_steps * ports * 2, 1), np.float32),
        "plane": np.zeros(max(n[0] * n[1] * nf * 10 if extras.plane_k is not None else 0, 1), np.float32),
        "ntff": np.zeros(max((len(extras.ntff) // 13 if extras.ntff is not None else 0) * nf * 12, 1),
                         np.float32),
        "probes": np.zeros(max(extras.max_steps * len(extras.probes), 1), np.float32),
    }
    rec = abi.fg_record()
    for name, arr in bufs.items():
        setattr(rec, name, arr.ctypes.data_as(C.POINTER(C.c_float)))
        setattr(rec, name + "_cap", arr.size)
    if fields:
        bufs["e"] = np.zeros(3 * nn, np.float32)
        bufs["h"] = np.zeros(3 * nn, np.float32)
        rec.e_out = bufs["e"].ctypes.data_as(C.POINTER(C.c_float))
        rec.h_out = bufs["h"].ctypes.data_as(C.POINTER(C.c_float))
    return rec, bufs


def _finish(rec: abi.fg_record, bufs, sim: Sim, extras: Extras) -> Record:
    steps, ports = int(rec.steps), int(rec.ports)
    nf = min(len(extras.freqs), abi.FG_MAX_FREQS)
    n = sim.dims()
    plane_len = n[0] * n[1] * nf * 10 if extras.plane_k is not None else 0
    ntff_len = (len(extras.ntff) // 13 if extras.ntff is not None else 0) * nf * 12
    return Record(
        steps=steps, ports=ports, series=bufs["series"][: steps * ports * 2].copy(),
        decay_db=float(rec.decay_db), plane=bufs["plane"][:plane_len].copy(),
        ntff=bufs["ntff"][:ntff_len].copy(),
        probes=bufs["probes"][: steps * len(extras.probes)].copy(),
        e=bufs.get("e"), h=bufs.get("h"))


def run(gpu: Gpu, sim: Sim, driven: int, pulse: Pulse, extras: Extras,
        progress: Optional[Callable[[int, float], bool]] = None, fields: bool = False) -> Record:
    """run / run_blocking (run.rs:378-758) on the HIP engine. Raises FieldgridError."""
    if gpu is None:
        raise ValueError("run() needs a Gpu; the CPU path is the reference's run_cpu")
    problem = sim.problem()
    keep = []
    x = extras.to_c(keep)
    pl = abi.fg_pulse(pulse.f0, pulse.fc)
    rec, bufs = _record_buffers(sim, extras, fields)
    errors = []

    def cb(_user, steps, decay):
        try:
            return 1 if (progress is None or progress(int(steps), float(decay))) else 0
        except BaseException as e:  # an exception in the callback cancels the run
            errors.append(e)
            return 0

    fn = abi.PROGRESS_FN(cb)
    status = abi.lib().fg_run(gpu._h, C.byref(problem), driven, C.byref(pl), C.byref(x), fn, None,
                              C.byref(rec))
    if errors:
        raise errors[0]
    abi.check(status)
    return _finish(rec, bufs, sim, extras)


run_blocking = run


@dataclass
class BatchRun:
    """One run of a batched sweep: the arguments of run(), plus the frequencies (Hz) of its
    port V/I spectra (the sims-antenna post-processing, crates/sims-antenna/src/fdtd/mod.rs:229-246)."""
    sim: Sim
    driven: int
    pulse: Pulse
    extras: Extras
    freqs: Sequence[float] = ()


@dataclass
class Tune:
    """March tile, steps per march pass (1 or 2), graph chunk and launch mode (1 per-pass
    launches, 2 persistent launches of many passes) for a batch (0 = autotune), with an optional
    JSON cache. `used` comes back as (tj, seg, chunk, depth) and `used_launch` as 1 or 2."""
    tj: int = 0
    seg: int = 0
    chunk: int = 0
    depth: int = 0
    cache: Optional[str] = None
    launch: int = 0
    used_launch: int = 0
    used: Tuple[int, int, int, int] = (0, 0, 0, 0)
    from_cache: bool = False
    tune_ms: float = 0.0
    step_ms: float = 0.0
    kept_steps: int = 0  # trial steps the runs kept (the batch went on from there)


def run_batch(gpu: Gpu, runs: Sequence[BatchRun], tune: Optional[Tune] = None,
              progress: Optional[Callable[[int, float], bool]] = None, fields: bool = False):
    """Run many independent problems together on the HIP engine (fg_run_batch). Each record is
    what run() gives for that problem. Returns (records, spectra) where spectra[r] is None or
    (V, I), complex arrays of shape (ports, len(freqs))."""
    if gpu is None:
        raise ValueError("run_batch() needs a Gpu")
    count = len(runs)
    keep = []
    problems = (C.POINTER(abi.fg_problem) * max(count, 1))()
    xs = (C.POINTER(abi.fg_extras) * max(count, 1))()
    drv = (C.c_uint32 * max(count, 1))()
    pls = (abi.fg_pulse * max(count, 1))()
    recs = (abi.fg_record * max(count, 1))()
    specs = (abi.fg_spectra * max(count, 1))()
    bufs_all, spec_bufs = [], []
    for r, job in enumerate(runs):
        p = job.sim.problem()
        x = job.extras.to_c(keep)
        keep += [p, x]
        problems[r] = C.pointer(p)
        xs[r] = C.pointer(x)
        drv[r] = job.driven
        pls[r] = abi.fg_pulse(job.pulse.f0, job.pulse.fc)
        rec, bufs = _record_buffers(job.sim, job.extras, fields)
        recs[r] = rec
        bufs_all.append(bufs)
        nf = len(job.freqs)
        if nf:
            f = np.ascontiguousarray(job.freqs, np.float64)
            ports = max(len(job.sim.ports), 1)
            v = np.zeros(ports * nf * 2, np.float64)
            i = np.zeros(ports * nf * 2, np.float64)
            keep += [f, v, i]
            specs[r] = abi.fg_spectra(nf, 0, f.ctypes.data_as(C.POINTER(C.c_double)),
                                      v.ctypes.data_as(C.POINTER(C.c_double)), i.ctypes.data_as(C.POINTER(C.c_double)))
            spec_bufs.append((v, i))
        else:
            spec_bufs.append(None)
    t = abi.fg_tune()
    if tune is not None:
        t.tj, t.seg, t.chunk, t.depth, t.launch = tune.tj, tune.seg, tune.chunk, tune.depth, tune.launch
        t.cache = tune.cache.encode() if tune.cache else None
    errors = []

    def cb(_user, steps, decay):
        try:
            return 1 if (progress is None or progress(int(steps), float(decay))) else 0
        except BaseException as e:
            errors.append(e)
            return 0

    fn = abi.PROGRESS_FN(cb)
    status = abi.lib().fg_run_batch(gpu._h, count, problems, drv, pls, xs, C.byref(t), fn, None, recs, specs)
    if errors:
        raise errors[0]
    abi.check(status)
    if tune is not None:
        tune.used = (t.used_tj, t.used_seg, t.used_chunk, t.used_depth)
        tune.from_cache = bool(t.from_cache)
        tune.tune_ms, tune.step_ms = t.tune_ms, t.step_ms
        tune.kept_steps = int(t.kept_steps)
        tune.used_launch = int(t.used_launch)
    records = [_finish(recs[r], bufs_all[r], runs[r].sim, runs[r].extras) for r in range(count)]
    spectra = []
    for r, sb in enumerate(spec_bufs):
        if sb is None:
            spectra.append(None)
            continue
        nf = len(runs[r].freqs)
        v, i = (a.reshape(-1, nf, 2) for a in sb)
        spectra.append((v[..., 0] + 1j * v[..., 1], i[..., 0] + 1j * i[..., 1]))
    return records, spectra


class Design:
    """d(objective)/d(eps_r) over a box of cells by the discrete adjoint (fg_adjoint_*; an
    extension without counterpart in the reference). The contract of a torch.autograd.Function:
    forward(eps_box) -> series of the observed E components after every step, shape
    (steps, n_obs); backward(dseries) -> dJ/d eps_r for the box, given dJ/d series.
    `obs` = [(field 0..2, i, j, k)]; the box [lo, hi) is in cells. `whole_tape`: the forward
    tapes the design edges of every step and the backward recomputes nothing (else
    checkpoint-and-recompute from snapshots every `every` steps)."""

    def __init__(self, gpu: Gpu, grid, mats, pec, lumped, ports, resistive, media, lo, hi, obs, driven,
                 pulse: Pulse, steps: int, every: int = 100, whole_tape: bool = False):
        if gpu is None:
            raise ValueError("Design needs a Gpu")
        self.gpu = gpu  # the device outlives the adjoint's sessions
        self.desc = SimDesc(grid, mats, pec, list(lumped), list(ports), list(resistive), media or Media())
        self.lo, self.hi = tuple(lo), tuple(hi)
        self.shape = tuple(h - l for l, h in zip(lo, hi))
        self.steps, self.n_obs = steps, len(obs)
        self._obs = (abi.fg_point_probe * max(len(obs), 1))(*[abi.fg_point_probe(*o) for o in obs])
        d = abi.fg_design()
        d.lo[:], d.hi[:] = list(lo), list(hi)
        d.n_obs, d.every, d.obs, d.steps = len(obs), every, self._obs, steps
        d.whole_tape = 1 if whole_tape else 0
        self._design = d
        h = C.c_void_p()
        abi.check(abi.lib().fg_adjoint_new(gpu._h, C.byref(self.desc.desc), C.byref(d), driven,
                                           C.byref(abi.fg_pulse(pulse.f0, pulse.fc)), C.byref(h)))
        self._h = h
        self.dt = None

    def forward(self, eps_box=None) -> np.ndarray:
        series = np.zeros(max(self.steps * self.n_obs, 1), np.float32)
        e = None if eps_box is None else np.ascontiguousarray(eps_box, np.float32).reshape(self.shape)
        abi.check(abi.lib().fg_adjoint_forward(self._h, None if e is None else e.ctypes.data_as(C.POINTER(C.c_float)),
                                               series.ctypes.data_as(C.POINTER(C.c_float))))
        return series[: self.steps * self.n_obs].reshape(self.steps, self.n_obs)

    def backward(self, dseries) -> np.ndarray:
        ds = np.ascontiguousarray(dseries, np.float64).reshape(-1)
        if ds.size != self.steps * self.n_obs:
            raise ValueError("dseries must hold steps x n_obs values")
        grad = np.zeros(self.shape)
        abi.check(abi.lib().fg_adjoint_backward(self._h, ds.ctypes.data_as(C.POINTER(C.c_double)),
                                                grad.ctypes.data_as(C.POINTER(C.c_double))))
        return grad

    def forward64(self, eps_box=None) -> np.ndarray:
        series = np.zeros(max(self.steps * self.n_obs, 1))
        e = None if eps_box is None else np.ascontiguousarray(eps_box, np.float64).reshape(self.shape)
        abi.check(abi.lib().fg_adjoint_forward64(self._h, None if e is None else e.ctypes.data_as(C.POINTER(C.c_double)),
                                                 series.ctypes.data_as(C.POINTER(C.c_double))))
        return series[: self.steps * self.n_obs].reshape(self.steps, self.n_obs)

    def timing(self):
        """(forward ms, backward ms, backward recompute ms, backward adjoint ms), snapshots."""
        ms = (C.c_double * 4)()
        snaps = C.c_uint64()
        abi.check(abi.lib().fg_adjoint_timing(self._h, ms, C.byref(snaps)))
        return tuple(ms), int(snaps.value)

    def close(self):
        if getattr(self, "_h", None):
            abi.lib().fg_adjoint_free(self._h)
            self._h = None

    def __del__(self):
        self.close()


def port_spectra(gpu: Gpu, series: np.ndarray, ports: int, dt: float, freqs: Sequence[float]):
    """The sims-antenna port spectra (crates/sims-antenna/src/fdtd/mod.rs:231-246) of a record's
    series, per port, on the device (fg_port_spectra): V(f) = sum_n V_n dt e^{-iw(n+1)dt},
    I(f) = sum_n I_n dt e^{-iw(n+1/2)dt} by the caller's f64 phasor recursion. Returns (V, I)
    of shape (ports, len(freqs))."""
    if gpu is None:
        raise ValueError("port_spectra() needs a Gpu")
    s = np.ascontiguousarray(series, np.float32)
    steps = s.size // max(2 * ports, 1)
    f = np.ascontiguousarray(freqs, np.float64)
    v = np.zeros(max(ports * len(f) * 2, 1))
    i = np.zeros_like(v)
    dp = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    abi.check(abi.lib().fg_port_spectra(gpu._h, dp(s, C.c_float), steps, ports, float(dt), len(f), dp(f, C.c_double),
                                        dp(v, C.c_double), dp(i, C.c_double)))
    shape = (ports, len(f))
    V = (v[0::2] + 1j * v[1::2])[: ports * len(f)].reshape(shape)
    I = (i[0::2] + 1j * i[1::2])[: ports * len(f)].reshape(shape)
    return V, I


@dataclass
class Outcome:
    """What sims-antenna's finish() (crates/sims-antenna/src/fdtd/mod.rs:229-266) derives from a
    single-port record: the impedance sweep, the impedance and input power at f0."""
    freqs: List[float]
    z: np.ndarray
    z0: complex
    p_in: float


def finish(gpu: Gpu, rec: Record, dt: float, freqs: Sequence[float], f0: float) -> Outcome:
    """finish() (mod.rs:229-266) for port 0 of `rec`: spectra at freqs and f0 on the device,
    Z = V / I, p_in = Re(V(f0) I(f0)*) / 2."""
    allf = list(freqs) + [f0]
    V, I = port_spectra(gpu, rec.series, rec.ports, dt, allf)
    z = V[0] / I[0]
    return Outcome(list(freqs), z[:-1], complex(z[-1]), float(0.5 * (V[0, -1] * np.conj(I[0, -1])).real))


def s_parameters(spectra, z0=50.0) -> np.ndarray:
    """The scattering matrix of an N-port from a port sweep: N runs of one problem, run d
    driving port d, spectra[d] = (V, I) of shape (N, nf) as run_batch or port_spectra give them.
    Power waves at a real reference impedance z0 (scalar, or one per port):
    a = (V + z0 I) / (2 sqrt(z0)), b = (V - z0 I) / (2 sqrt(z0)), with I flowing into the
    structure, as finish() takes it (p_in = Re(V I*) / 2 > 0 at the driven port); then
    S = B A^-1 at each frequency, the columns of A and B being the runs. With z0 equal to the
    ports' own resistances only the driven port carries an incident wave. No counterpart in the
    reference, whose sims-rf networks take S-parameters from elsewhere. Returns (nf, N, N)."""
    runs = len(spectra)
    V = np.stack([np.asarray(v, complex) for v, _ in spectra], axis=2)  # (port, f, run)
    I = np.stack([np.asarray(i, complex) for _, i in spectra], axis=2)
    if V.shape[0] != runs or I.shape != V.shape:
        raise ValueError(f"an {V.shape[0]}-port needs {V.shape[0]} runs, one per driven port, of matching spectra")
    z = np.broadcast_to(np.asarray(z0, float), (runs,)).reshape(runs, 1, 1)
    if np.any(z <= 0):
        raise ValueError("reference impedances must be positive")
    A = (V + z * I) / (2.0 * np.sqrt(z))
    B = (V - z * I) / (2.0 * np.sqrt(z))
    A, B = np.moveaxis(A, 1, 0), np.moveaxis(B, 1, 0)  # (f, port, run)
    return np.linalg.solve(np.swapaxes(A, 1, 2), np.swapaxes(B, 1, 2)).swapaxes(1, 2)


class Session:
    """A resident run (fg_session_*): upload once, step, read. Used by benchmarks."""

    def __init__(self, gpu: Gpu, sim: Sim, driven: int, pulse: Pulse, extras: Extras):
        self.gpu, self.sim, self.extras = gpu, sim, extras
        self.problem = sim.problem()
        self._keep = []
        self._x = extras.to_c(self._keep)
        self._pl = abi.fg_pulse(pulse.f0, pulse.fc)
        h = C.c_void_p()
        abi.check(abi.lib().fg_session_new(gpu._h, C.byref(self.problem), driven, C.byref(self._pl),
                                           C.byref(self._x), C.byref(h)))
        self._h = h

    def step(self, count: int):
        abi.check(abi.lib().fg_session_step(self._h, count))

    def sync(self):
        abi.check(abi.lib().fg_session_sync(self._h))

    def steps_per_pass(self) -> int:
        """Steps one pass of the sweep advances (2 with temporal blocking, else 1)."""
        v = C.c_uint32()
        abi.check(abi.lib().fg_session_steps_per_pass(self._h, C.byref(v)))
        return int(v.value)

    def timing(self, enable: bool):
        """Switch per-kernel HIP-event timing; returns ([sweep, fix, observe] ms, steps) so far."""
        ms = (C.c_double * 3)()
        steps = C.c_uint64()
        abi.check(abi.lib().fg_session_timing(self._h, int(enable), ms, C.byref(steps)))
        return list(ms), int(steps.value)

    def energy(self) -> float:
        v = C.c_double()
        abi.check(abi.lib().fg_session_energy(self._h, C.byref(v)))
        return v.value

    def read(self, fields: bool = False) -> Record:
        rec, bufs = _record_buffers(self.sim, self.extras, fields)
        abi.check(abi.lib().fg_session_read(self._h, C.byref(rec)))
        return _finish(rec, bufs, self.sim, self.extras)

    def close(self):
        if getattr(self, "_h", None):
            abi.lib().fg_session_free(self._h)
            self._h = None

    __del__ = close


# ---- NTFF (ntff.rs) ------------------------------------------------------------------------

@dataclass
class Patches:
    pos: np.ndarray   # (count, 3) f64
    area: np.ndarray  # (count,) f64
    gpu: np.ndarray   # (count*13,) u32


def patches(sim: Sim, inset: int) -> Patches:
    """ntff.rs:11-59"""
    L = abi.lib()
    count = L.fg_patches(sim._h, inset, None, None, None)
    gpu = np.zeros(max(count * 13, 1), np.uint32)
    pos = np.zeros((max(count, 1), 3), np.float64)
    area = np.zeros(max(count, 1), np.float64)
    L.fg_patches(sim._h, inset, gpu.ctypes.data_as(C.POINTER(C.c_uint32)),
                 pos.ctypes.data_as(C.POINTER(C.c_double)), area.ctypes.data_as(C.POINTER(C.c_double)))
    return Patches(pos[:count], area[:count], gpu[: count * 13])


class Far:
    """ntff.rs:66-137 on the device (include/fieldgrid_far.h): the f64 far-field transform of
    one frequency's NTFF accumulators. Needs a Gpu, like run()."""

    def __init__(self, gpu: "Gpu", p: "Patches", acc: np.ndarray, nf: int, k: int, freq: float,
                 scale=(1.0, 0.0)):
        if gpu is None:
            raise ValueError("Far needs a Gpu")
        self.gpu = gpu  # the device outlives this object's buffers
        pos = np.ascontiguousarray(p.pos, np.float64)
        area = np.ascontiguousarray(p.area, np.float64)
        acc = np.ascontiguousarray(acc, np.float32)
        count = len(area)
        if acc.size < count * nf * 12:
            raise ValueError("accumulators shorter than patches x nf x 12")
        h = C.c_void_p()
        dp = lambda a, t: a.ctypes.data_as(C.POINTER(t))
        abi.check(abi.lib().fg_far_new(gpu._h, count, dp(pos, C.c_double), dp(area, C.c_double), dp(acc, C.c_float),
                                       nf, k, float(freq), float(scale[0]), float(scale[1]), C.byref(h)))
        self._h, self.freq, self.count = h, float(freq), count

    @staticmethod
    def load(gpu: "Gpu", p: "Patches", acc: np.ndarray, nf: int, k: int, freq: float, scale=(1.0, 0.0)) -> "Far":
        """Far::load (ntff.rs:76-88)."""
        return Far(gpu, p, acc, nf, k, freq, scale)

    @staticmethod
    def _dirs(d):
        d = np.ascontiguousarray(np.asarray(d, np.float64).reshape(-1, 3))
        return d, d.ctypes.data_as(C.POINTER(C.c_double))

    def field(self, d) -> np.ndarray:
        """Far::field (ntff.rs:90-115): one direction -> (3,) complex, or (n, 3) -> (n, 3)."""
        single = np.asarray(d).ndim == 1
        d, dp = self._dirs(d)
        out = np.zeros((len(d), 6))
        abi.check(abi.lib().fg_far_field(self._h, len(d), dp, out.ctypes.data_as(C.POINTER(C.c_double))))
        z = out[:, 0::2] + 1j * out[:, 1::2]
        return z[0] if single else z

    def intensity(self, d):
        """Far::intensity (ntff.rs:117-119): a float for one direction, else an array."""
        single = np.asarray(d).ndim == 1
        d, dp = self._dirs(d)
        out = np.zeros(len(d))
        abi.check(abi.lib().fg_far_intensity(self._h, len(d), dp, out.ctypes.data_as(C.POINTER(C.c_double))))
        return float(out[0]) if single else out

    def radiated(self, steps: int) -> Tuple[float, float]:
        """Far::radiated (ntff.rs:121-136): (total power, peak intensity)."""
        t, pk = C.c_double(), C.c_double()
        abi.check(abi.lib().fg_far_radiated(self._h, steps, C.byref(t), C.byref(pk)))
        return t.value, pk.value

    def close(self):
        if getattr(self, "_h", None):
            abi.lib().fg_far_free(self._h)
            self._h = None

    __del__ = close
