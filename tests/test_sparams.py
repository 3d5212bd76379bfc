"""S-parameters of an N-port from a port sweep (fdtd.s_parameters; no counterpart in the reference).

CPU: the power-wave algebra against a network whose S, waves and reference impedances are chosen
first (V = sqrt(z0) (a + b), I = (a - b) / sqrt(z0) for any invertible incident-wave matrix), the
one-port closed form S11 = (Z - z0) / (Z + z0) with Z = V / I, and the input checks.
GPU: a two-port in vacuum, its two runs batched (each port driven once), spectra on the device:
the matrix must be reciprocal (S12 = S21, the Yee update with these media is reciprocal), passive
(no singular value above 1) and show the coupling.
"""
import math

import numpy as np
import pytest

from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W


def test_the_power_wave_algebra_recovers_a_chosen_network():
    rng = np.random.default_rng(0)
    n, nf = 3, 7
    S = rng.normal(size=(nf, n, n)) + 1j * rng.normal(size=(nf, n, n))
    A = rng.normal(size=(nf, n, n)) + 1j * rng.normal(size=(nf, n, n))  # any invertible excitation
    B = S @ A
    z0 = np.array([50.0, 75.0, 25.0])
    r = np.sqrt(z0)[None, :, None]
    V, I = r * (A + B), (A - B) / r
    spectra = [(V[:, :, d].T, I[:, :, d].T) for d in range(n)]
    got = F.s_parameters(spectra, z0)
    assert np.max(np.abs(got - S)) < 1e-12 * np.max(np.abs(S))


def test_a_one_port_gives_the_reflection_of_its_impedance():
    rng = np.random.default_rng(1)
    V = rng.normal(size=(1, 9)) + 1j * rng.normal(size=(1, 9))
    I = rng.normal(size=(1, 9)) + 1j * rng.normal(size=(1, 9))
    z = V[0] / I[0]
    s11 = F.s_parameters([(V, I)], 50.0)[:, 0, 0]
    assert np.allclose(s11, (z - 50.0) / (z + 50.0), rtol=1e-12, atol=0)


def test_the_sweep_must_drive_every_port_once():
    V = np.ones((2, 4), complex)
    with pytest.raises(ValueError):
        F.s_parameters([(V, V)], 50.0)  # a two-port from one run
    with pytest.raises(ValueError):
        F.s_parameters([(V, V), (V, V)], [50.0, -1.0])


@pytest.mark.gpu
def test_a_two_port_in_vacuum_is_reciprocal_and_passive(gpu):
    nodes, dx, pml = 44, 0.5e-3, 8
    grid = F.Grid.uniform([0.0] * 3, [(nodes - 1) * dx] * 3, [nodes - 1] * 3, pml)
    kc = nodes // 2
    # two 8 mm single-column 50-ohm z-ports (near half a wavelength at the top of the band, so
    # they radiate and couple), placed without a symmetry that would force S12 = S21
    spots = [(15, 19), (27, 24)]
    ports = [F.PortDef(f"p{q + 1}", [[F.Edge(2, (i, j, k)) for k in range(kc - 8, kc + 8)]], 50.0, False)
             for q, (i, j) in enumerate(spots)]
    sim = F.Sim.try_new(grid, F.Materials.new(grid), None, [], ports, [], F.Media())
    freqs = [f * 1e9 for f in (8, 10, 12, 14, 16, 18)]
    runs = [F.BatchRun(sim, d, W.boxes_pulse(), F.Extras(max_steps=3000, decay_db=math.inf), freqs) for d in (0, 1)]
    recs, spectra = F.run_batch(gpu, runs, tune=F.Tune(seg=16, chunk=8, depth=2))
    S = F.s_parameters(spectra, 50.0)
    recip = np.max(np.abs(S[:, 0, 1] - S[:, 1, 0])) / np.max(np.abs(S[:, 1, 0]))
    sv = np.max(np.linalg.svd(S, compute_uv=False))
    print(f"\n|S21| {np.round(np.abs(S[:, 1, 0]), 4)}, |S11| {np.round(np.abs(S[:, 0, 0]), 4)}")
    print(f"reciprocity |S12 - S21| / max|S21| = {recip:.2e}, largest singular value {sv:.4f}")
    assert np.min(np.abs(S[:, 1, 0])) > 1e-3  # the ports couple
    assert recip < 1e-3
    assert sv <= 1.0 + 1e-3
