"""Writes the golden fixtures of this directory from the FP32 oracle (the reference's f32
operation order, oracle/fdtd_oracle.cpp). The reference holds no golden vectors for this path,
so these pin the oracle and the device against drift; each one names the reference recipe it
follows. Run from the repository root: python -m tests.golden.make_golden

Each .npz holds, per run, the traces (port V/I series, probe samples, plane DFT, NTFF
accumulators where the run has them) as f32, SHA-256 digests of the final E and H (f32
bytes), and a downsampled slice of E and H at mid-height for inspection.
"""
import hashlib
import math
import os

import numpy as np

from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W
from oracle import oracle as O
from tests.fixtures import everything, everything_extras

HERE = os.path.dirname(os.path.abspath(__file__))
BIG = 20000


def digest(a):
    return hashlib.sha256(np.asarray(a, np.float32).tobytes()).hexdigest()


def cases():
    """name -> (sim, driven, pulse, extras). Small enough for the oracle in seconds."""
    out = {}
    # config 1 (the plain 64^3 vacuum run: soft source, 8 Ez/Hy probes, 500 steps)
    sim, probes = W.plain(nodes=64, pml=10)
    out["plain_64"] = (sim, 0, W.plain_pulse(), F.Extras(max_steps=500, decay_db=math.inf, probes=probes))
    # tests/fdtd.rs:43-125: every feature (sheets, thin wires, Debye band, C/R/L, two ports,
    # plane DFT, NTFF), each port driven
    sim = everything()
    for driven in (0, 1):
        out[f"everything_driven{driven}"] = (sim, driven, F.Pulse(5e9, 5e9), everything_extras(sim, 600))
    # the boxes recipe at 56^3 (dielectric and Debye boxes with per-box tau, one port, NTFF)
    sim = W.boxes(nodes=56, pml=6, nboxes=8, ndebye=4, per_box_tau=True)
    out["boxes_56"] = (sim, 0, W.boxes_pulse(), W.boxes_extras(sim, steps=400, inset=2))
    return out


def record_arrays(rec, n):
    """The fixture arrays of one record (device or oracle)."""
    e = np.asarray(rec.e, np.float32).reshape(3, *n)
    h = np.asarray(rec.h, np.float32).reshape(3, *n)
    k = n[2] // 2
    arrays = {
        "steps": np.array([rec.steps], np.int64),
        "series": np.asarray(rec.series, np.float32),
        "probes": np.asarray(rec.probes if rec.probes is not None else [], np.float32),
        "e_slice": e[:, ::2, ::2, k].copy(),
        "h_slice": h[:, ::2, ::2, k].copy(),
    }
    digests = {"e_sha256": digest(rec.e), "h_sha256": digest(rec.h)}
    # DFT accumulators: whole when small, else their digest and first BIG values
    for name in ("plane", "ntff"):
        a = np.asarray(getattr(rec, name), np.float32)
        if a.size <= BIG:
            arrays[name] = a
        else:
            arrays[name + "_head"] = a[:BIG]
            digests[name + "_sha256"] = digest(a)
    return arrays, digests


def main():
    for name, (sim, driven, pulse, ex) in cases().items():
        rec = O.run_cpu(sim.desc, driven, pulse, ex, 32)
        arrays, digests = record_arrays(rec, sim.dims())
        for key, value in digests.items():
            arrays[key] = np.array(value)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **arrays)
        print(f"{name}: {rec.steps} steps, {os.path.getsize(path) / 1024:.0f} KiB")


if __name__ == "__main__":
    main()
