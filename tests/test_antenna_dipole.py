"""The reference's own antenna validation (crates/sims-antenna/tests/fdtd.rs:45-55), run through
the restated sims-antenna caller (tests/antenna.py) on the 141 mm dipole with a 3 mm gap:
radiated power within 1 % of the port's input power, directivity 2.15 +- 0.05 dBi, and the first
upward zero of Im Z (resonance) in 960-1010 MHz. On the CPU through the oracle's FP32 path (the
reference's run_cpu, as the reference test runs it), and on the GPU through fg_run, the device
port spectra and the device far field."""
import math

import numpy as np
import pytest

from fieldgrid_amd import fdtd as F
from oracle import oracle as O

from . import antenna as A


def far_intensity(far):
    """Far::field / intensity (ntff.rs:90-119) of the oracle's numpy Far, vectorised over
    directions (n, 3)."""
    def intensity(d):
        k = 2.0 * math.pi * far.freq / A.C0
        ph = far.area[None, :] * np.exp(1j * k * (d @ far.pos.T))
        nvec, lvec = ph @ far.j, ph @ far.m
        perp = nvec - np.sum(nvec * d, -1, keepdims=True) * d
        e = (np.cross(d, lvec) - perp * A.ETA0) * (k / (4.0 * math.pi))
        return np.sum(np.abs(e) ** 2, -1) / (2.0 * A.ETA0)
    return intensity


def check_bars(total, peak, p_in, freqs, z):
    eff = total / p_in
    d = 10.0 * math.log10(4.0 * math.pi * peak / total)
    f = A.resonance(freqs, z)
    print(f"power balance {eff:.5f}, directivity {d:.4f} dBi, resonance {f / 1e6:.1f} MHz")
    assert abs(eff - 1.0) < 0.01, eff
    assert abs(d - 2.15) < 0.05, d
    assert 960e6 <= f < 1010e6, f


def dipole_on_the_oracle():
    s = A.setup(A.dipole())
    rec = O.run_cpu(s.sim.desc, 0, F.Pulse(s.f0, s.fc), A.extras(s), 32, fields=False)
    print(f"dims {s.grid.dims()}, {rec.steps} steps, decay {rec.decay_db:.1f} dB")
    freqs, z, p_in = A.finish_host(s, rec.series, rec.dt)
    far = O.Far.load(s.patches, rec.ntff, 1, 0, s.f0)
    total, peak = A.radiated(far_intensity(far))
    return total, peak, p_in, freqs, z


def test_the_dipole_meets_the_reference_bars_on_the_oracle():
    check_bars(*dipole_on_the_oracle())


def test_a_conjugated_ntff_phase_fails_the_dipole_bars():
    O.set_mutation(7)
    try:
        res = dipole_on_the_oracle()
    finally:
        O.set_mutation(0)
    with pytest.raises(AssertionError):
        check_bars(*res)


@pytest.mark.gpu
def test_the_dipole_meets_the_reference_bars_on_the_gpu(gpu):
    s = A.setup(A.dipole())
    rec = F.run(gpu, s.sim, 0, F.Pulse(s.f0, s.fc), A.extras(s))
    print(f"dims {s.grid.dims()}, {rec.steps} steps, decay {rec.decay_db:.1f} dB")
    out = F.finish(gpu, rec, s.sim.dt, s.freqs, s.f0)
    far = F.Far.load(gpu, s.patches, rec.ntff, 1, 0, s.f0)
    total, peak = A.radiated(far.intensity)
    check_bars(total, peak, out.p_in, out.freqs, out.z)
    far.close()
