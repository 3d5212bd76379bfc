"""The device's field, port and far-field errors against the FP64 oracle, stated per case.

The device reproduces the reference's FP32 path bit for bit (tests/test_gpu_parity.py,
tests/test_golden.py); these numbers are therefore the FP32 path's own rounding error, which
the tests bound:
  * fields: relative L2 of E and H at mid-run (the pulse still inside the grid);
  * ports: worst |V, I - fp64| over the run, relative to the largest |V|, |I| (fdtd.rs worst());
  * probes, plane DFT, NTFF accumulators: worst() likewise;
  * far field: the device transform (fg_far) of the device's NTFF accumulators against the FP64
    oracle's transform of the FP64 accumulators, over 500 directions, relative to the largest
    |E_far|; and the radiated power.
"""
import math

import numpy as np
import pytest

from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W
from oracle import oracle as O

from .fixtures import everything, everything_extras, rel_l2, worst

pytestmark = pytest.mark.gpu


def test_plain_64_fields_and_probes_against_fp64(gpu):
    """The 64^3 plain run (soft source, 8 Ez/Hy probes) at 250 of its 500 steps."""
    sim, probes = W.plain()
    ex = F.Extras(max_steps=250, decay_db=math.inf, probes=probes)
    rec = F.run(gpu, sim, 0, W.plain_pulse(), ex, fields=True)
    o64 = O.run_cpu(sim.desc, 0, W.plain_pulse(), ex, 64)
    e, h, p = rel_l2(o64.e, rec.e), rel_l2(o64.h, rec.h), worst(o64.probes, rec.probes)
    print(f"\nplain 64^3, 250 steps: E rel L2 {e:.2e}, H rel L2 {h:.2e}, probes worst {p:.2e}")
    # fields within the 1e-5 target; a single probe sample carries the f32 rounding of one node
    # (~1e-5 of the trace's peak here), which the integrated outputs average down
    assert e < 1e-5 and h < 1e-5 and p < 5e-5


def test_every_feature_ports_and_dfts_against_fp64(gpu):
    """tests/fdtd.rs:43-125 with every feature (sheets and wires included), 1500 steps."""
    sim = everything()
    ex = everything_extras(sim, 1500)
    rec = F.run(gpu, sim, 0, F.Pulse(5e9, 5e9), ex, fields=True)
    o64 = O.run_cpu(sim.desc, 0, F.Pulse(5e9, 5e9), ex, 64)
    s = np.asarray(rec.series, np.float64).reshape(-1, 2 * rec.ports)
    r = np.asarray(o64.series).reshape(-1, 2 * rec.ports)
    v = np.max(np.abs(s[:, 0::2] - r[:, 0::2])) / np.max(np.abs(r[:, 0::2]))
    i = np.max(np.abs(s[:, 1::2] - r[:, 1::2])) / np.max(np.abs(r[:, 1::2]))
    pl, nt = worst(o64.plane, rec.plane), worst(o64.ntff, rec.ntff)
    print(f"\neverything, 1500 steps: port V worst {v:.2e}, port I worst {i:.2e}, plane DFT {pl:.2e}, NTFF {nt:.2e}")
    assert v < 1e-4 and i < 1e-4 and pl < 1e-4 and nt < 1e-4


def test_dipole_far_field_against_fp64(gpu):
    """A centre-fed z dipole (the oracle's radiated-power test recipe at 40^3): NTFF at 6 GHz on
    the device, its far field by fg_far, against the FP64 oracle's run and Far."""
    nodes, f = 40, 6e9
    grid = F.Grid.uniform([0.0] * 3, [(nodes - 1) * 1e-3] * 3, [nodes - 1] * 3, 6)
    c = nodes // 2

    def pec(comp, i, j, k):
        return (comp == 2) & (i == c) & (j == c) & (k >= c - 5) & (k <= c + 5) & (k != c)

    sim = F.Sim.try_new(grid, F.Materials.new(grid), pec, [], [F.PortDef("feed", [[F.Edge(2, (c, c, c))]], 50.0, False)],
                        [], F.Media())
    pt = F.patches(sim, 2)
    ex = F.Extras(max_steps=1500, decay_db=math.inf, freqs=[f], ntff=pt.gpu)
    rec = F.run(gpu, sim, 0, F.Pulse(f, f), ex)
    o64 = O.run_cpu(sim.desc, 0, F.Pulse(f, f), ex, 64, fields=False)
    rng = np.random.default_rng(5)
    dirs = rng.normal(size=(500, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    dev = F.Far.load(gpu, pt, rec.ntff, 1, 0, f)
    ref = O.Far.load(pt, o64.ntff, 1, 0, f)
    got = dev.field(dirs)
    want = np.array([ref.field(d) for d in dirs])
    err = np.max(np.abs(got - want)) / np.max(np.abs(want))
    t_dev, _ = dev.radiated(18)
    t_ref, _ = ref.radiated(18)
    p = abs(t_dev - t_ref) / t_ref
    print(f"\ndipole far field: |E_far| worst {err:.2e} over 500 directions, radiated power {p:.2e}")
    assert err < 1e-4 and p < 1e-4
    dev.close()
