"""The batched sweep (include/fieldgrid_batch.h): many runs in shared launches.

Each batched run must give what a single run of the same problem gives, bit for bit: the
per-node arithmetic is the same march kernel body whatever the tile shape, graph chunk or
batch company. Single runs are themselves checked against the oracle in test_gpu_parity.py;
one batched run is also checked against the FP32 oracle here. Port spectra from the device
are checked against the closed-form sums (direct exponentials, not the phasor recursion the
caller and the device use).
"""
import cmath
import math
import os

import numpy as np
import pytest

from fieldgrid_amd import abi
from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W
from oracle import oracle as O

from .fixtures import worst


def finish_spectra(series, ports, dt, freqs):
    """crates/sims-antenna/src/fdtd/mod.rs:231-246 restated as written: f64 phasors rotated by
    e^{-iw dt} per step, V with pe (starting at dt e^{-iw dt}), I with ph (dt e^{-iw dt/2})."""
    s = np.asarray(series, np.float64).reshape(-1, ports, 2)
    V = np.zeros((ports, len(freqs)), complex)
    I = np.zeros((ports, len(freqs)), complex)
    for p in range(ports):
        for q, f in enumerate(freqs):
            w = 2.0 * math.pi * f
            rot = cmath.rect(1.0, -w * dt)
            pe, ph = cmath.rect(dt, -w * dt), cmath.rect(dt, -w * 0.5 * dt)
            v = i = 0j
            for n in range(s.shape[0]):
                v += pe * s[n, p, 0]
                i += ph * s[n, p, 1]
                pe *= rot
                ph *= rot
            V[p, q], I[p, q] = v, i
    return V, I


def test_port_spectra_match_the_callers_phasor_recursion():
    """The oracle's port spectra (direct exponentials) against the recursion of the sims-antenna
    caller, and both against the closed form for a unit impulse at step n0: V = dt e^{-iw(n0+1)dt}."""
    rng = np.random.default_rng(7)
    ports, steps, dt = 3, 257, 1.3e-12
    freqs = [1e9, 7.5e9, 19e9]
    series = rng.standard_normal(steps * ports * 2).astype(np.float32)
    V, I = O.port_spectra(series, ports, dt, freqs)
    V2, I2 = finish_spectra(series, ports, dt, freqs)
    assert np.max(np.abs(V - V2)) <= 1e-9 * np.max(np.abs(V2))
    assert np.max(np.abs(I - I2)) <= 1e-9 * np.max(np.abs(I2))
    imp = np.zeros((steps, ports, 2), np.float32)
    n0 = 40
    imp[n0, 1, 0] = 1.0
    imp[n0, 1, 1] = 1.0
    V, I = O.port_spectra(imp.ravel(), ports, dt, freqs)
    for q, f in enumerate(freqs):
        w = 2 * math.pi * f
        assert abs(V[1, q] - dt * cmath.exp(-1j * w * (n0 + 1) * dt)) < 1e-12 * dt
        assert abs(I[1, q] - dt * cmath.exp(-1j * w * (n0 + 0.5) * dt)) < 1e-12 * dt
        assert V[0, q] == 0 and I[2, q] == 0


def test_the_sweep_recipe_keeps_its_ranges():
    runs = W.sweep_runs(count=6, lo=30, hi=44, steps=(50, 90), nboxes=3, ndebye=1, nfreq=5)
    assert len(runs) == 6
    for r in runs:
        n = r.sim.dims()
        assert all(30 <= v <= 44 for v in n)
        assert 1 <= len(r.sim.ports) <= 4 and 0 <= r.driven < len(r.sim.ports)
        assert 50 <= r.extras.max_steps <= 90 and len(r.freqs) == 5


# The persistent launches (fg_tune.launch = 2) are switched off: their first GPU run (the
# depth2-seg16-persistent case below) did not finish within its time limit, and the cause is
# not known. Kept, skipped, until it is.
PERSIST_OFF = pytest.mark.skip(reason="persistent launches switched off: their first GPU run did not finish, cause "
                                      "unknown")


def small_sweep(count=5, steps=(60, 150), **kw):
    return W.sweep_runs(count=count, lo=24, hi=46, steps=steps, nboxes=3, ndebye=1, nfreq=6, **kw)


def _same(a, b, what):
    assert a.steps == b.steps, f"{what}: steps {a.steps} vs {b.steps}"
    assert a.decay_db == b.decay_db, what
    for name in ("series", "plane", "ntff", "probes", "e", "h"):
        x, y = getattr(a, name), getattr(b, name)
        if x is None:
            continue
        assert np.array_equal(x, y), f"{what}: {name} differs, worst {worst(y, x) if len(x) else 0}"


@pytest.mark.gpu
@pytest.mark.parametrize("tune", [F.Tune(tj=14, seg=16, chunk=8, depth=1), F.Tune(tj=8, seg=8, chunk=1, depth=1),
                                  F.Tune(seg=16, chunk=8, depth=2, launch=1), F.Tune(seg=3, chunk=1, depth=2, launch=1),
                                  pytest.param(F.Tune(seg=16, depth=2, launch=2), marks=PERSIST_OFF),
                                  pytest.param(F.Tune(seg=3, depth=2, launch=2), marks=PERSIST_OFF), F.Tune()],
                         ids=["tj14-seg16-graph8", "tj8-seg8-direct", "depth2-seg16-graph8", "depth2-seg3-direct",
                              "depth2-seg16-persistent", "depth2-seg3-persistent", "autotuned"])
def test_batched_runs_equal_single_runs(gpu, tune):
    """Every batched record bit-identical to a single fg_run of the same problem, with one step
    per march pass and with two (the runs have Debye boxes: the first step's ADE runs inside the
    pass). Odd step counts and check_every 40 end chunks on single steps and shrink the lists."""
    runs = small_sweep()
    for r in runs:
        r.extras.check_every = 40  # several decay checks and list rebuilds per run
    print(f"\nbatch of {len(runs)}: dims {[r.sim.dims() for r in runs]}, steps {[r.extras.max_steps for r in runs]}")
    recs, spectra = F.run_batch(gpu, runs, tune=tune, fields=True)
    print(f"  tuned: tj, seg, chunk, depth = {tune.used}, tune {tune.tune_ms:.1f} ms, step {tune.step_ms:.3f} ms")
    if tune.depth:
        assert tune.used[3] == tune.depth
    if tune.launch:
        assert tune.used_launch == tune.launch
    if not (tune.tj or tune.seg or tune.chunk or tune.depth):
        assert tune.kept_steps == 0  # a decay check at step 40 falls inside the trials: rolled back
    for q, r in enumerate(runs):
        single = F.run(gpu, r.sim, r.driven, r.pulse, r.extras, fields=True)
        _same(recs[q], single, f"run {q}")
        V, I = spectra[q]
        V2, I2 = O.port_spectra(single.series, len(r.sim.ports), r.sim.dt, r.freqs)
        assert np.max(np.abs(V - V2)) <= 1e-9 * max(np.max(np.abs(V2)), 1e-300)
        assert np.max(np.abs(I - I2)) <= 1e-9 * max(np.max(np.abs(I2)), 1e-300)


@pytest.mark.gpu
def test_the_tuners_trial_steps_are_kept_as_progress(gpu):
    """With no decay check or last step inside the tuner's trials, the runs keep the trial steps
    (one and two steps per pass, direct and graph-replayed, switching between candidates) and
    the batch goes on from there: every record still bit-identical to a single run."""
    runs = small_sweep(count=4, steps=(400, 520))
    tune = F.Tune()
    recs, _ = F.run_batch(gpu, runs, tune=tune, fields=True)
    print(f"\ntuned {tune.used}, kept {tune.kept_steps} trial steps, tune {tune.tune_ms:.1f} ms")
    assert not tune.from_cache and 0 < tune.kept_steps < 400
    for q, r in enumerate(runs):
        single = F.run(gpu, r.sim, r.driven, r.pulse, r.extras, fields=True)
        _same(recs[q], single, f"run {q}")


@pytest.mark.gpu
def test_a_batched_run_matches_the_oracle(gpu):
    runs = small_sweep(count=3, steps=(120, 160))
    recs, _ = F.run_batch(gpu, runs, tune=F.Tune(chunk=8))
    for q, r in enumerate(runs):
        o32 = O.run_cpu(r.sim.desc, r.driven, r.pulse, r.extras, 32, fields=False)
        w = worst(o32.series, recs[q].series)
        print(f"\nrun {q}: series worst vs fp32 oracle {w:.2e}")
        assert recs[q].steps == o32.steps
        assert w < 1e-5


@pytest.mark.gpu
def test_batched_decay_stops_each_run_where_a_single_run_stops(gpu):
    runs = small_sweep(count=4, steps=(2000, 2000))
    for q, r in enumerate(runs):
        r.extras.decay_db = 10.0 + 3 * q
        r.extras.check_every = [50, 100, 70, 64][q]
    recs, _ = F.run_batch(gpu, runs, tune=F.Tune(tj=14, seg=16, chunk=16))
    for q, r in enumerate(runs):
        single = F.run(gpu, r.sim, r.driven, r.pulse, r.extras)
        print(f"\nrun {q}: stopped at {recs[q].steps} (single {single.steps}), decay {recs[q].decay_db:.2f} dB")
        assert recs[q].steps < 2000
        _same(recs[q], single, f"run {q}")


@pytest.mark.gpu
def test_the_tuning_cache_is_written_and_reused(gpu, tmp_path):
    runs = small_sweep(count=3, steps=(40, 60))
    path = str(tmp_path / "tune.json")
    t1 = F.Tune(cache=path)
    recs1, _ = F.run_batch(gpu, runs, tune=t1)
    assert os.path.exists(path) and not t1.from_cache and t1.tune_ms > 0
    t2 = F.Tune(cache=path)
    recs2, _ = F.run_batch(gpu, runs, tune=t2)
    assert t2.from_cache and t2.used == t1.used and t2.tune_ms == 0
    for a, b in zip(recs1, recs2):
        _same(a, b, "cached")


@pytest.mark.gpu
def test_progress_can_cancel_a_batch(gpu):
    runs = small_sweep(count=2, steps=(300, 300))
    for r in runs:
        r.extras.check_every = 100
    seen = []
    with pytest.raises(F.abi.FieldgridError) as e:
        F.run_batch(gpu, runs, tune=F.Tune(tj=14, seg=16, chunk=8),
                    progress=lambda s, d: seen.append(s) or s < 200)
    assert str(e.value) == "cancelled" and seen == [100, 200]


@pytest.mark.gpu
def test_device_port_spectra_and_finish_match_the_direct_sums(gpu):
    """fg_port_spectra (finish()'s recursion on the device) against the oracle's direct sums, and
    finish()'s Z and p_in from them (crates/sims-antenna/src/fdtd/mod.rs:229-266)."""
    rng = np.random.default_rng(9)
    ports, steps, dt = 2, 3001, 1.1e-12
    freqs = list(np.linspace(1e9, 20e9, 64))
    series = rng.standard_normal(steps * ports * 2).astype(np.float32)
    V, I = F.port_spectra(gpu, series, ports, dt, freqs)
    V2, I2 = O.port_spectra(series, ports, dt, freqs)
    assert np.max(np.abs(V - V2)) <= 1e-9 * np.max(np.abs(V2))
    assert np.max(np.abs(I - I2)) <= 1e-9 * np.max(np.abs(I2))
    rec = F.Record(steps, ports, series, 0.0, np.zeros(0), np.zeros(0))
    out = F.finish(gpu, rec, dt, freqs[:10], 7e9)
    V3, I3 = O.port_spectra(series, ports, dt, freqs[:10] + [7e9])
    z = V3[0] / I3[0]
    assert np.max(np.abs(out.z - z[:-1])) <= 1e-9 * np.max(np.abs(z))
    assert abs(out.z0 - z[-1]) <= 1e-9 * abs(z[-1])
    assert abs(out.p_in - 0.5 * (V3[0, -1] * np.conj(I3[0, -1])).real) <= 1e-9 * abs(V3[0, -1] * I3[0, -1])


@pytest.mark.gpu
def test_a_given_segment_outside_the_tuners_list_is_timed_as_given(gpu):
    """With the segment length fixed by the caller (20 planes, not one of the tuner's 32/16/8)
    and the depth left to the tuner, the tuner times that length: a finite step time, and the
    records still equal single runs."""
    runs = small_sweep(count=3)
    tune = F.Tune(seg=20)
    recs, _ = F.run_batch(gpu, runs, tune=tune)
    print(f"\n  used {tune.used}, step {tune.step_ms:.3f} ms")
    assert tune.used[1] == 20 and 0.0 < tune.step_ms < 1e3
    for q, r in enumerate(runs):
        _same(recs[q], F.run(gpu, r.sim, r.driven, r.pulse, r.extras), f"run {q}")


@pytest.mark.gpu
@pytest.mark.parametrize("tune", [F.Tune(seg=16, chunk=8), F.Tune(tj=8, seg=5, chunk=1), F.Tune()],
                         ids=["seg16-graph8", "seg5-direct-asked-tj8", "autotuned"])
def test_sheets_and_wires_in_a_batch_equal_single_runs(gpu, tune):
    """Runs with resistive sheets and thin wires (the reference's everything() fixture, both ports
    driven; and with the sheets alone and the wires alone) batched with plain sweep runs: the
    wire runs take the wire march (14-row tiles whatever the caller asks), the sheet runs their
    correction after the march, every run one step per pass; every record, the plane DFT and
    NTFF accumulators included, bit-identical to a single fg_run."""
    from .fixtures import everything, everything_extras
    runs = []
    for sheets, wires, driven in ((True, True, 0), (True, True, 1), (True, False, 0), (False, True, 1)):
        sim = everything(sheets=sheets, wires=wires)
        runs.append(F.BatchRun(sim, driven, F.Pulse(5e9, 5e9), everything_extras(sim, 150 + 17 * len(runs)),
                               [3e9, 7e9]))
    runs += small_sweep(count=2, steps=(90, 130))
    recs, spectra = F.run_batch(gpu, runs, tune=tune, fields=True)
    print(f"\n  used {tune.used}, step {tune.step_ms:.3f} ms")
    assert tune.used[0] == 14 and tune.used[3] == 1
    for q, r in enumerate(runs):
        single = F.run(gpu, r.sim, r.driven, r.pulse, r.extras, fields=True)
        _same(recs[q], single, f"run {q}")


@PERSIST_OFF
@pytest.mark.gpu
@pytest.mark.parametrize("seg", [32, 5])
def test_persistent_passes_with_probes_equal_single_runs(gpu, seg):
    """Persistent launches (each run on to its next pass as soon as its previous one is done;
    step n's port V/I and probes read from the pass's captures) against single runs, bit for
    bit, with point probes of E and H in every run and odd step counts (a single step ends
    each window)."""
    runs = small_sweep(count=6, steps=(41, 97), seed=11)
    rng = np.random.default_rng(3)
    for r in runs:
        n = r.sim.dims()
        r.extras.check_every = 30
        # at i = pml + 1, clear of the dielectric boxes (they start at pml + 2): no Debye edge
        r.extras.probes = [(int(f), 9, int(rng.integers(2, n[1] - 2)), int(rng.integers(2, n[2] - 2)))
                           for f in (0, 2, 3, 5)]
    tune = F.Tune(seg=seg, depth=2, launch=2)
    recs, _ = F.run_batch(gpu, runs, tune=tune, fields=True)
    assert tune.used[3] == 2 and tune.used_launch == 2
    for q, r in enumerate(runs):
        single = F.run(gpu, r.sim, r.driven, r.pulse, r.extras, fields=True)
        assert np.count_nonzero(single.probes) > 0
        _same(recs[q], single, f"run {q}")


@pytest.mark.gpu
def test_a_persistent_launch_request_is_refused_while_switched_off(gpu):
    """fg_tune.launch = 2 is refused with FG_UNSUPPORTED while the persistent launches are off."""
    runs = small_sweep(count=2, steps=(8, 12))
    with pytest.raises(abi.FieldgridError) as e:
        F.run_batch(gpu, runs, tune=F.Tune(seg=16, depth=2, launch=2))
    assert "persistent launches are not available" in str(e.value)
