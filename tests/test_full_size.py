"""Parity at the workloads' full sizes.

The oracle comparisons elsewhere run scaled-down grids. Here the 768x768x384 layered run (226 M
nodes, CPML 16, two ports) and the 256^3 boxes run (Debye boxes, one port, NTFF 4 cells inside
the CPML) keep their real shape: block classes, segment lengths, concurrent class streams, XCD
ordering and byte offsets past 2 GiB. They are checked by
  * the FP32 oracle over the first steps: fields, DFT accumulators and port series as in the
    scaled-down tests (fields and accumulators bit-identical);
  * a size-independent property over a long run: two steps per pass against one step per pass
    (each bit-identical to the oracle on the scaled-down grids), every output bit-identical,
    after the pulse has crossed most of the grid and its CPML faces.
"""
import math

import numpy as np
import pytest

from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W
from oracle import oracle as O

from .test_gpu_parity import _exact

pytestmark = pytest.mark.gpu


def _same(a, b, names, what):
    for name in names:
        x, y = getattr(a, name), getattr(b, name)
        diff = int(np.sum(x != y))
        print(f"  {what} {name}: {diff} of {x.size} values differ")
        assert diff == 0, f"{what} {name}"


def _nonzero_share(e):
    return float(np.count_nonzero(e)) / e.size


def _two_steps_per_pass(gpu, sim, pulse, ex):
    s = F.Session(gpu, sim, 0, pulse, ex)
    try:
        return s.steps_per_pass() == 2
    finally:
        s.close()


def test_the_768x768x384_layered_run_matches_the_oracle(gpu):
    sim = W.layered()
    ex = F.Extras(max_steps=12, decay_db=math.inf)
    rec = F.run(gpu, sim, 0, W.layered_pulse(), ex, fields=True)
    o32 = O.run_cpu(sim.desc, 0, W.layered_pulse(), ex, 32)
    print(f"\nlayered 768x768x384, 12 steps: {_nonzero_share(rec.e):.2e} of E nonzero")
    _exact(rec, o32, "layered full size")


def test_the_768x768x384_layered_run_two_steps_per_pass_change_nothing(gpu, monkeypatch):
    """1200 steps: the pulse from the ports at x-index 40 and 52 reaches the low-x, both y and
    both z CPML faces and ~690 cells along x."""
    sim = W.layered()
    ex = F.Extras(max_steps=1200, decay_db=math.inf, check_every=500)
    assert _two_steps_per_pass(gpu, sim, W.layered_pulse(), ex)
    two = F.run(gpu, sim, 0, W.layered_pulse(), ex, fields=True)
    monkeypatch.setenv("FIELDGRID_FUSE2", "0")
    one = F.run(gpu, sim, 0, W.layered_pulse(), ex, fields=True)
    print(f"\nlayered 768x768x384, 1200 steps: {_nonzero_share(two.e):.3f} of E nonzero")
    assert _nonzero_share(two.e) > 0.5
    _same(two, one, ("series", "e", "h"), "layered two vs one step per pass")


def test_planes_near_the_offset_limit_march_two_steps_in_short_segments(gpu, monkeypatch):
    """Planes of 5600 x 5600 nodes (125 MB per field component and plane). A two-step block
    reaches its planes through 32-bit lane offsets from a resource based at its first plane, and
    the host caps the segment so that those stay below 2^31: here at 3 planes, with offsets up
    to ~2.08e9 bytes. 60 steps, two steps per pass against one, every output bit-identical, and
    the pulse in every plane of E but the last (the boundary face, whose E the update leaves at
    0, as the FP32 oracle does on a 12 x 60 x 60 version)."""
    sim = W.layered(dims=(12, 5600, 5600), pml=3, port_i=(4, 7))
    pulse = W.layered_pulse()
    ex = F.Extras(max_steps=60, decay_db=math.inf, check_every=30)
    assert _two_steps_per_pass(gpu, sim, pulse, ex)
    two = F.run(gpu, sim, 0, pulse, ex, fields=True)
    monkeypatch.setenv("FIELDGRID_FUSE2", "0")
    one = F.run(gpu, sim, 0, pulse, ex, fields=True)
    n = sim.dims()
    planes = np.abs(two.e.reshape(3, n[0], n[1] * n[2])).max(axis=(0, 2))
    print(f"\n12x5600x5600, 60 steps: largest |E| per plane {planes}")
    assert np.all(planes[:-1] > 0)
    _same(two, one, ("series", "e", "h"), "12x5600x5600 two vs one step per pass")


def test_planes_past_the_offset_limit_take_one_step_per_pass(gpu):
    """Planes of 7000 x 7000 nodes (196 MB): not even a one-plane segment keeps a two-step
    block's offsets below 2^31, so the session marches one step per pass."""
    sim = W.layered(dims=(8, 7000, 7000), pml=2, port_i=(3, 4))
    ex = F.Extras(max_steps=4, decay_db=math.inf)
    assert not _two_steps_per_pass(gpu, sim, W.layered_pulse(), ex)


def test_the_256_cubed_boxes_run_matches_the_oracle(gpu):
    """240 steps: the pulse from the central port reaches the NTFF faces."""
    sim = W.boxes()
    ex = W.boxes_extras(sim, steps=240, check_every=0)
    rec = F.run(gpu, sim, 0, W.boxes_pulse(), ex, fields=True)
    o32 = O.run_cpu(sim.desc, 0, W.boxes_pulse(), ex, 32)
    print(f"\nboxes 256^3, 240 steps: {_nonzero_share(rec.e):.2e} of E nonzero, "
          f"{_nonzero_share(rec.ntff):.2e} of the NTFF accumulators")
    assert _nonzero_share(rec.ntff) > 0.0
    _exact(rec, o32, "boxes full size")


def test_the_256_cubed_boxes_run_two_steps_per_pass_change_nothing(gpu, monkeypatch):
    """800 steps: the pulse fills the grid, crossing the Debye boxes, the NTFF surface (whose
    points the two-step pass captures) and every CPML face."""
    sim = W.boxes()
    ex = W.boxes_extras(sim, steps=800, check_every=0)
    assert _two_steps_per_pass(gpu, sim, W.boxes_pulse(), ex)
    two = F.run(gpu, sim, 0, W.boxes_pulse(), ex, fields=True)
    monkeypatch.setenv("FIELDGRID_FUSE2", "0")
    one = F.run(gpu, sim, 0, W.boxes_pulse(), ex, fields=True)
    print(f"\nboxes 256^3, 800 steps: {_nonzero_share(two.e):.3f} of E nonzero")
    assert _nonzero_share(two.e) > 0.9
    _same(two, one, ("series", "ntff", "e", "h"), "boxes two vs one step per pass")


def test_the_192x192x112_design_forward_matches_the_oracle(gpu):
    """The inverse-design grid (4.1 M nodes, CPML 12, every design cell its own coefficient set,
    so u32 set indices read from L2): the forward's probe series over 1000 of its 3000 steps,
    bit-identical to the FP32 oracle's point probe."""
    from .test_adjoint import make
    args, _ = W.design(steps=1000)
    d = make(gpu, args, every=100)
    s = d.forward()
    ex = F.Extras(max_steps=args["steps"], decay_db=math.inf, probes=args["obs"])
    o32 = O.run_cpu(d.desc, args["driven"], args["pulse"], ex, 32, fields=False)
    ref = np.asarray(o32.probes, np.float32)
    print(f"\ndesign grid, 1000 steps: {int(np.sum(s[:, 0] != ref))} of {ref.size} probe samples differ, "
          f"peak {np.max(np.abs(ref)):.3e}")
    assert np.max(np.abs(ref)) > 0
    assert np.array_equal(s[:, 0], ref)


def test_sweep_recipe_runs_batched_equal_single_runs(gpu):
    """Eight runs of the 96-run sweep recipe at their real sizes (48-200 nodes per axis, CPML 8,
    Debye boxes, 1-4 ports) cut to 600 steps, batched with the autotuner (its trial steps kept)
    against single runs: every record bit-identical."""
    runs = W.sweep_runs(count=8)
    for r in runs:
        r.extras.max_steps = 600
    tune = F.Tune()
    recs, _ = F.run_batch(gpu, runs, tune=tune, fields=True)
    print(f"\n8 sweep runs {[r.sim.dims() for r in runs]}: tuned {tune.used}, kept {tune.kept_steps} steps")
    assert tune.kept_steps > 0
    for q, r in enumerate(runs):
        single = F.run(gpu, r.sim, r.driven, r.pulse, r.extras, fields=True)
        _same(recs[q], single, ("series", "e", "h"), f"run {q}")


def test_the_96_run_sweep_at_full_step_counts(gpu):
    """The batched sweep as benchmarked: all 96 runs (48-200 nodes per axis, CPML 8, Debye boxes,
    1-4 ports, 800-3000 steps each) in one autotuned fg_run_batch. Every record is bit-identical
    to an fg_run of the same problem, and the device spectra at the 64 frequencies match direct
    f64 exponential sums of the series (the sims-antenna finish(), mod.rs:231-246)."""
    runs = W.sweep_runs()
    assert len(runs) == 96
    recs, spectra = F.run_batch(gpu, runs, tune=F.Tune())
    worst = 0.0
    for q, r in enumerate(runs):
        single = F.run(gpu, r.sim, r.driven, r.pulse, r.extras)
        assert single.steps == r.extras.max_steps
        assert np.array_equal(recs[q].series, single.series), f"run {q}"
        V, I = spectra[q]
        v, i = O.port_spectra(recs[q].series, recs[q].ports, r.sim.dt, r.freqs)
        scale = max(np.abs(v).max(), 1e-300), max(np.abs(i).max(), 1e-300)
        worst = max(worst, float(np.abs(V - v).max() / scale[0]), float(np.abs(I - i).max() / scale[1]))
    print(f"\n96 sweep runs: records bit-identical to single runs; spectra worst {worst:.2e}")
    assert worst < 1e-9


def test_the_design_gradient_at_full_size_matches_central_differences(gpu):
    """The inverse-design gradient as benchmarked: 192x192x112 nodes, 3000 steps, snapshots every
    100 steps. dJ/d eps_r against central differences (h = 1e-3) of the f64 forward at 8 seeded
    design cells, within 1e-3 relative."""
    from .test_adjoint import make
    args, freq = W.design()
    d = make(gpu, args, every=100)
    try:
        dt = F.time_step(args["grid"])
        J, dJ = W.dft_objective(d.forward()[:, 0], dt, freq)
        grad = d.backward(dJ)
        n = args["grid"].dims()
        box = np.array(args["mats"].eps.reshape(tuple(q - 1 for q in n))[
            tuple(slice(l, h) for l, h in zip(args["lo"], args["hi"]))], np.float64)
        rng = np.random.default_rng(7)
        cells = [np.unravel_index(int(q), box.shape) for q in rng.choice(box.size, size=8, replace=False)]
        h = 1e-3
        errs = []
        for c in cells:
            up, dn = box.copy(), box.copy()
            up[c] += h
            dn[c] -= h
            fd = (W.dft_objective(d.forward64(up)[:, 0], dt, freq)[0] -
                  W.dft_objective(d.forward64(dn)[:, 0], dt, freq)[0]) / (2 * h)
            errs.append(abs(grad[c] - fd) / abs(fd))
            print(f"  cell {tuple(int(v) for v in c)}: adjoint {grad[c]:.6e}, central difference {fd:.6e}")
        print(f"192x192x112 x 3000 steps: worst relative error {max(errs):.2e}")
        assert max(errs) < 1e-3
    finally:
        d.close()
