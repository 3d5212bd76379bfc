"""ctypes binding of the FDTD oracle (oracle/fdtd_oracle.cpp) — test infrastructure only.

Imported by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, always as
the checker or the timed CPU baseline, never as part of the product path.
"""
import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from fieldgrid_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class oracle_out(C.Structure):
    _fields_ = [("steps", C.c_uint64), ("decay_db", C.c_double), ("progress_calls", C.c_uint64),
                ("series", C.POINTER(C.c_double)), ("plane", C.POINTER(C.c_double)),
                ("ntff", C.POINTER(C.c_double)), ("probes", C.POINTER(C.c_double)),
                ("e_out", C.POINTER(C.c_double)), ("h_out", C.POINTER(C.c_double)), ("dt", C.c_double),
                ("step_seconds", C.c_double), ("warm_steps", C.c_uint64), ("build_seconds", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(f"{LIB_PATH} is missing: build it with `python oracle/build.py`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.oracle_run.restype = C.c_int
        L.oracle_run.argtypes = [P(abi.fg_sim_desc), C.c_uint32, P(abi.fg_pulse), P(abi.fg_extras), C.c_int,
                                 C.c_uint64, C.c_double, P(oracle_out)]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_set_mutation.argtypes = [C.c_int]
        L.oracle_pack.restype = C.c_int
        L.oracle_pack.argtypes = [P(abi.fg_sim_desc), C.c_uint64, P(C.c_void_p), P(abi.fg_problem), P(C.c_double)]
        L.oracle_pack_free.argtypes = [C.c_void_p]
        L.oracle_pack_ca.restype = P(C.c_float)
        L.oracle_pack_ca.argtypes = [C.c_void_p, C.c_int]
        L.oracle_pack_cb.restype = P(C.c_float)
        L.oracle_pack_cb.argtypes = [C.c_void_p, C.c_int]
        L.oracle_patches.restype = C.c_uint64
        L.oracle_patches.argtypes = [P(C.c_double), C.c_uint32, P(C.c_double), C.c_uint32, P(C.c_double),
                                     C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_uint32), P(C.c_double),
                                     P(C.c_double)]
        L.oracle_port_loop.restype = C.c_uint64
        L.oracle_port_loop.argtypes = [P(abi.fg_sim_desc), C.c_uint32, C.c_uint32, C.c_uint32, P(C.c_uint64),
                                       P(C.c_uint32), P(C.c_double)]
        L.oracle_axis64.argtypes = [P(C.c_double), C.c_uint32, C.c_uint32, C.c_double, P(C.c_double)]
        L.oracle_threads.restype = C.c_int
        L.oracle_check.restype = C.c_int
        L.oracle_check.argtypes = [P(abi.fg_sim_desc)]
        L.oracle_design_gradient.restype = C.c_int
        L.oracle_design_gradient.argtypes = [P(abi.fg_sim_desc), C.c_uint32, P(abi.fg_pulse), C.c_uint64, C.c_uint32,
                                             P(abi.fg_point_probe), P(C.c_uint32), P(C.c_uint32), P(C.c_double),
                                             P(C.c_double), P(C.c_double)]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(message)
        self.status = status


@dataclass
class OracleRecord:
    steps: int
    decay_db: float
    progress_calls: int
    series: np.ndarray
    plane: np.ndarray
    ntff: np.ndarray
    probes: np.ndarray
    e: Optional[np.ndarray]
    h: Optional[np.ndarray]
    dt: float
    step_seconds: float = 0.0
    build_seconds: float = 0.0


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def run_cpu(desc, driven, pulse, extras, precision=32, cancel_at=0, dt=None, fields=True, warm=0) -> OracleRecord:
    """run_cpu (cpu.rs:424-482) on the oracle: precision 32 = the reference's FP32 path,
    64 = the FP64 oracle. `desc` is a fieldgrid_amd.fdtd.SimDesc; `extras` an Extras.
    The record's step_seconds times the stepping loop from step `warm` on."""
    n = desc.n
    nn = n[0] * n[1] * n[2]
    keep = []
    x = extras.to_c(keep)
    nf = x.nf
    ports = desc.desc.n_ports
    bufs = {
        "series": np.zeros(max(extras.max_steps * ports * 2, 1)),
        "plane": np.zeros(max(n[0] * n[1] * nf * 10 if extras.plane_k is not None else 0, 1)),
        "ntff": np.zeros(max(int(x.n_patches) * nf * 12, 1)),
        "probes": np.zeros(max(extras.max_steps * len(extras.probes), 1)),
    }
    out = oracle_out()
    out.warm_steps = warm
    for name, arr in bufs.items():
        setattr(out, name, _dptr(arr))
    if fields:
        bufs["e"] = np.zeros(3 * nn)
        bufs["h"] = np.zeros(3 * nn)
        out.e_out, out.h_out = _dptr(bufs["e"]), _dptr(bufs["h"])
    pl = abi.fg_pulse(pulse.f0, pulse.fc)
    st = lib().oracle_run(C.byref(desc.desc), driven, C.byref(pl), C.byref(x), precision, cancel_at,
                          dt or 0.0, C.byref(out))
    if st != 0:
        raise OracleError(st, lib().oracle_last_error().decode())
    steps = int(out.steps)
    plane_len = n[0] * n[1] * nf * 10 if extras.plane_k is not None else 0
    return OracleRecord(steps, out.decay_db, int(out.progress_calls), bufs["series"][: steps * ports * 2],
                        bufs["plane"][:plane_len], bufs["ntff"][: int(x.n_patches) * nf * 12],
                        bufs["probes"][: steps * len(extras.probes)], bufs.get("e"), bufs.get("h"), out.dt,
                        out.step_seconds, out.build_seconds)


class Packed:
    """The oracle's FP32 build + pack, as an fg_problem (for table-by-table comparison)."""

    def __init__(self, desc, narrow=0):
        self.h = C.c_void_p()
        self.problem = abi.fg_problem()
        dt = C.c_double()
        st = lib().oracle_pack(C.byref(desc.desc), narrow, C.byref(self.h), C.byref(self.problem), C.byref(dt))
        if st != 0:
            raise OracleError(st, lib().oracle_last_error().decode())
        self.dt = dt.value
        self.nn = desc.n[0] * desc.n[1] * desc.n[2]

    def coefficients(self):
        ca = np.stack([np.ctypeslib.as_array(lib().oracle_pack_ca(self.h, c), shape=(self.nn,)).copy() for c in range(3)])
        cb = np.stack([np.ctypeslib.as_array(lib().oracle_pack_cb(self.h, c), shape=(self.nn,)).copy() for c in range(3)])
        return ca, cb

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_pack_free(self.h)
            self.h = None


def patches(grid, inset):
    """ntff.rs:11-59 restated: (gpu u32 table, pos, area)."""
    lines = [np.ascontiguousarray(grid.axis(a), np.float64) for a in range(3)]
    args = []
    for l in lines:
        args += [_dptr(l), len(l)]
    count = lib().oracle_patches(*args, grid.pml, inset, None, None, None)
    gpu = np.zeros(max(count * 13, 1), np.uint32)
    pos = np.zeros((max(count, 1), 3))
    area = np.zeros(max(count, 1))
    lib().oracle_patches(*args, grid.pml, inset, gpu.ctypes.data_as(C.POINTER(C.c_uint32)), _dptr(pos), _dptr(area))
    return gpu[: count * 13], pos[:count], area[:count]


def port_loop(desc, port, axis, k):
    """run.rs:102-127 restated: list of (id, comp, w)."""
    cap = 4096
    ids = np.zeros(cap, np.uint64)
    comps = np.zeros(cap, np.uint32)
    ws = np.zeros(cap)
    count = lib().oracle_port_loop(C.byref(desc.desc), port, axis, k, ids.ctypes.data_as(C.POINTER(C.c_uint64)),
                                   comps.ctypes.data_as(C.POINTER(C.c_uint32)), _dptr(ws))
    return [(int(ids[q]), int(comps[q]), float(ws[q])) for q in range(count)]


def axis64(line, pml, dt):
    """sim.rs:29-131 at f64: rows of (d, dd, be, ce, bh, ch, slot_e, slot_h)."""
    line = np.ascontiguousarray(line, np.float64)
    rows = np.zeros((len(line), 8))
    lib().oracle_axis64(_dptr(line), len(line), pml, dt, _dptr(rows))
    return rows


def check(desc):
    """Sim::try_new's error text, or None."""
    st = lib().oracle_check(C.byref(desc.desc))
    return None if st == 0 else lib().oracle_last_error().decode()


def threads() -> int:
    """OpenMP threads the oracle's plane loops use."""
    return int(lib().oracle_threads())


def set_mutation(m: int):
    """Plant mistake number m in the oracle (0 = none). Tests only."""
    lib().oracle_set_mutation(m)


def design_gradient(desc, driven, pulse, steps, obs, lo, hi, dseries):
    """The FP64 discrete adjoint: dJ/d eps_r over the design box [lo, hi) of cells (k fastest),
    given dJ/ds_n for the E observations `obs` = [(field, i, j, k)] after every step.
    Returns (grad shaped like the box, series (steps, n_obs))."""
    n_obs = len(obs)
    pr = (abi.fg_point_probe * max(n_obs, 1))(*[abi.fg_point_probe(*o) for o in obs])
    lo_ = (C.c_uint32 * 3)(*lo)
    hi_ = (C.c_uint32 * 3)(*hi)
    box = tuple(int(h - l) for l, h in zip(lo, hi))
    grad = np.zeros(box)
    series = np.zeros((steps, max(n_obs, 1)))
    ds = np.ascontiguousarray(dseries, np.float64).reshape(-1)
    assert ds.size == steps * n_obs
    pl = abi.fg_pulse(pulse.f0, pulse.fc)
    st = lib().oracle_design_gradient(
        C.byref(desc.desc), driven, C.byref(pl), steps, n_obs, pr, lo_, hi_, _dptr(ds), _dptr(grad), _dptr(series))
    if st != 0:
        raise OracleError(st, lib().oracle_last_error().decode())
    return grad, series[:, :n_obs]


class Far:
    """ntff.rs:66-137 restated in numpy f64 (Far::load, field, intensity, radiated): the checker
    of the device far field."""

    def __init__(self, pos, area, j, m, freq):
        self.pos, self.area, self.j, self.m, self.freq = pos, area, j, m, freq

    @staticmethod
    def load(p, acc, nf, k, freq, scale=(1.0, 0.0)):
        """ntff.rs:76-88"""
        count = len(p.pos)
        a = np.asarray(acc, np.float64).reshape(count, nf, 12)[:, k, :]
        z = (a[:, 0::2] + 1j * a[:, 1::2]) * complex(scale[0], scale[1])
        return Far(np.asarray(p.pos, np.float64), np.asarray(p.area, np.float64), z[:, :3], z[:, 3:], freq)

    def field(self, d):
        """ntff.rs:90-115"""
        d = np.asarray(d, np.float64)
        k = 2.0 * np.pi * self.freq / 299_792_458.0
        ph = self.area * np.exp(1j * k * (self.pos @ d))
        nn = (self.j * ph[:, None]).sum(0)
        ll = (self.m * ph[:, None]).sum(0)
        perp = nn - (nn * d).sum() * d
        return (np.cross(d, ll) - perp * 376.730_313_668) * (k / (4.0 * np.pi))

    def intensity(self, d):
        """ntff.rs:117-119"""
        f = self.field(d)
        return float((f.real ** 2 + f.imag ** 2).sum() / (2.0 * 376.730_313_668))

    def radiated(self, steps):
        """ntff.rs:121-136"""
        total, peak = 0.0, 0.0
        for a in range(steps):
            th = np.pi * (a + 0.5) / steps
            for b in range(2 * steps):
                ph = np.pi * b / steps
                u = self.intensity([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
                peak = max(peak, u)
                total += u * np.sin(th) * (np.pi / steps) * (np.pi / steps)
        return total, peak


def port_spectra(series, ports, dt, freqs):
    """The sims-antenna port spectra (crates/sims-antenna/src/fdtd/mod.rs:231-246) by direct
    exponentials in f64: V(f) = sum_n V_n dt e^{-iw(n+1)dt}, I(f) = sum_n I_n dt e^{-iw(n+1/2)dt}.
    Returns (V, I) of shape (ports, len(freqs))."""
    s = np.asarray(series, np.float64).reshape(-1, ports, 2)
    n = np.arange(s.shape[0], dtype=np.float64)
    w = 2.0 * np.pi * np.asarray(freqs, np.float64)
    pe = dt * np.exp(-1j * np.outer(w, n + 1.0) * dt)
    ph = dt * np.exp(-1j * np.outer(w, n + 0.5) * dt)
    return (pe @ s[:, :, 0]).T, (ph @ s[:, :, 1]).T


def sheet_weights(adaptive, ha, hb, start=(0.5, 0.5)):
    """The face weights of cpu.rs:169-193 after each step of a sequence of face fields (ha, hb),
    from a sheet record whose weights start at `start` (pack's 0.5, 0.5). Returns (ma, mb)."""
    L = lib()
    ha = np.ascontiguousarray(ha, np.float32)
    hb = np.ascontiguousarray(hb, np.float32)
    ma = np.zeros_like(ha)
    mb = np.zeros_like(ha)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))
    L.oracle_sheet_weights.argtypes = [C.c_int, C.c_float, C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                       C.c_uint64, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    L.oracle_sheet_weights(int(adaptive), start[0], start[1], fp(ha), fp(hb), ha.size, fp(ma), fp(mb))
    return ma, mb
