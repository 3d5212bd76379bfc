"""The discrete adjoint on the GPU (include/fieldgrid_adjoint.h) against the oracle.

* forward: the FP32 series at the observation is bit-identical to the FP32 oracle's point probe;
* backward: dJ/d eps_r matches the oracle's FP64 discrete adjoint (itself pinned by central
  differences in tests/test_adjoint_oracle.py) within 1e-3 of the largest component, and
  central differences of the device's own f64 forward within 1e-3;
* the snapshot interval changes nothing (the recomputed segments are the same arithmetic).
The objective is J = |sum_n Ez_n dt e^{-iw(n+1)dt}|^2 at 8 GHz (workloads.dft_objective).
"""
import math

import numpy as np
import pytest

from fieldgrid_amd import abi
from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W

from .fixtures import sheet_table
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def make(gpu, args, every=100, whole_tape=False):
    return F.Design(gpu, args["grid"], args["mats"], args["pec"], args["lumped"], args["ports"], args["resistive"],
                    args["media"], args["lo"], args["hi"], args["obs"], args["driven"], args["pulse"], args["steps"],
                    every, whole_tape)


def box_of(args):
    n = args["grid"].dims()
    cells = (n[0] - 1, n[1] - 1, n[2] - 1)
    sl = tuple(slice(l, h) for l, h in zip(args["lo"], args["hi"]))
    return args["mats"].eps.reshape(cells)[sl].copy()


def test_the_forward_is_the_fp32_path(gpu):
    args, freq = W.design_small(steps=300)
    d = make(gpu, args, every=64)
    s = d.forward()
    ex = F.Extras(max_steps=args["steps"], decay_db=math.inf, probes=args["obs"])
    o32 = O.run_cpu(d.desc, args["driven"], args["pulse"], ex, 32, fields=False)
    assert np.array_equal(s[:, 0], np.asarray(o32.probes, np.float32))
    o64 = O.run_cpu(d.desc, args["driven"], args["pulse"], ex, 64, fields=False)
    s64 = d.forward64()
    err = np.max(np.abs(s64[:, 0] - o64.probes)) / np.max(np.abs(o64.probes))
    print(f"\nforward64 vs fp64 oracle: {err:.2e} (f32 axis tables and coefficients outside the box)")
    assert err < 1e-5


@pytest.mark.parametrize("sigma", [0.0, 0.02], ids=["lossless", "lossy"])
def test_the_gradient_matches_the_fp64_adjoint(gpu, sigma):
    args, freq = W.design_small(steps=300, sigma=sigma)
    d = make(gpu, args, every=64)
    s = d.forward()
    J, dJ = W.dft_objective(s[:, 0], F.time_step(args["grid"]), freq)
    g = d.backward(dJ)
    go, so = O.design_gradient(d.desc, args["driven"], args["pulse"], args["steps"], args["obs"], args["lo"],
                               args["hi"], dJ)
    rel = np.max(np.abs(g - go)) / np.max(np.abs(go))
    print(f"\nJ = {J:.4e}; max |grad - fp64 adjoint| / max |fp64 adjoint| = {rel:.2e}")
    assert rel < 1e-3
    big = np.abs(go) > 0.1 * np.max(np.abs(go))
    assert np.max(np.abs(g[big] - go[big]) / np.abs(go[big])) < 1e-3


def test_the_gradient_matches_central_differences_of_the_f64_forward(gpu):
    args, freq = W.design_small(steps=300, sigma=0.02)
    d = make(gpu, args, every=50)
    s = d.forward()
    dt = F.time_step(args["grid"])
    _, dJ = W.dft_objective(s[:, 0], dt, freq)
    g = d.backward(dJ)
    box = box_of(args).astype(np.float64)
    rng = np.random.default_rng(3)
    h = 1e-3
    for _ in range(6):
        c = tuple(int(rng.integers(0, n)) for n in box.shape)
        up, dn = box.copy(), box.copy()
        up[c] += h
        dn[c] -= h
        jp = W.dft_objective(d.forward64(up)[:, 0], dt, freq)[0]
        jm = W.dft_objective(d.forward64(dn)[:, 0], dt, freq)[0]
        fd = (jp - jm) / (2 * h)
        print(f"cell {c}: adjoint {g[c]:.6e}  central difference {fd:.6e}")
        assert abs(g[c] - fd) < 1e-3 * max(abs(fd), 1e-3 * np.max(np.abs(g)))


def test_the_snapshot_interval_does_not_change_the_gradient(gpu):
    args, freq = W.design_small(steps=200)
    grads = []
    for every in (7, 100, 200):
        d = make(gpu, args, every=every)
        s = d.forward()
        _, dJ = W.dft_objective(s[:, 0], F.time_step(args["grid"]), freq)
        grads.append(d.backward(dJ))
        _, snaps = d.timing()
        assert snaps == math.ceil(200 / every) - 1
        d.close()
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[1], grads[2])


def test_a_new_eps_is_used_by_the_next_forward(gpu):
    args, freq = W.design_small(steps=120)
    d = make(gpu, args)
    base = d.forward()
    box = box_of(args)
    box[0, 0, 0] += 0.5
    moved = d.forward(box)
    assert not np.array_equal(base, moved)
    args["mats"].eps.reshape(tuple(n - 1 for n in args["grid"].dims()))[
        tuple(slice(l, h) for l, h in zip(args["lo"], args["hi"]))] = box
    ex = F.Extras(max_steps=args["steps"], decay_db=math.inf, probes=args["obs"])
    o32 = O.run_cpu(F.SimDesc(args["grid"], args["mats"], None, [], args["ports"], [], args["media"]), 0,
                    args["pulse"], ex, 32, fields=False)
    assert np.array_equal(moved[:, 0], np.asarray(o32.probes, np.float32))


def test_problems_outside_the_adjoints_reach_are_refused(gpu):
    args, _ = W.design_small(steps=10)
    near = dict(args, lo=(2, 11, 10))  # reaches into the 5-cell CPML
    with pytest.raises(abi.FieldgridError) as e:
        make(gpu, near)
    assert e.value.status == abi.FG_INVALID and "CPML" in str(e.value)
    port = dict(args, lo=(6, 12, 10), hi=(9, 15, 14))  # the port column is at (7, 14)
    with pytest.raises(abi.FieldgridError) as e:
        make(gpu, port)
    assert e.value.status == abi.FG_INVALID and "port" in str(e.value)
    # resistive sheets weigh their faces by the run's accumulated H^2: not linear, refused
    n = args["grid"].dims()
    sheet = dict(args, resistive=[F.Resistive(0, F.idx(n, 22, 20, 12), 30.0, False, (0, 1))],
                 media=F.Media(sheets=sheet_table()))
    with pytest.raises(abi.FieldgridError) as e:
        make(gpu, sheet)
    assert e.value.status == abi.FG_UNSUPPORTED and "nonlinear" in str(e.value)
    wire = dict(args, media=F.Media(wires=[(F.Edge(2, (12, 12, k)), 0.1e-3) for k in range(14, 17)]))
    with pytest.raises(abi.FieldgridError) as e:
        make(gpu, wire)
    assert e.value.status == abi.FG_INVALID and "thin wire" in str(e.value)


def _backward_both_ways(gpu, args, monkeypatch, every=40):
    d = make(gpu, args, every=every)
    s = d.forward()
    _, dJ = W.dft_objective(s[:, 0], F.time_step(args["grid"]), freq_of(args))
    monkeypatch.setenv("FIELDGRID_ADJ_PAIRS", "0")  # one adjoint step per march pass
    monkeypatch.setenv("FIELDGRID_ADJ_MARCH", "1")
    g_march = d.backward(dJ)
    monkeypatch.setenv("FIELDGRID_ADJ_MARCH", "0")
    g_gather = d.backward(dJ)
    d.close()
    return g_march, g_gather


def freq_of(args):
    return args["pulse"].f0


@pytest.mark.parametrize("tj", ["14", "8"])
def test_the_fused_adjoint_march_equals_the_gather_kernels(gpu, monkeypatch, tj):
    """The one-pass adjoint march (k_adj_march) against the two gather kernels (k_adj_e +
    k_adj_h), bit for bit, on a grid with interior, z-slab and x/y-slab blocks (40x64x140 nodes,
    CPML 5, segments of 4 planes), on the same grid with a 30x40x60-cell box (72 000 distinct
    cells: more than 65 536 coefficient sets, so u32 set indices and sets read from global
    memory) and on the oracle-size grid (slab blocks only)."""
    monkeypatch.setenv("FIELDGRID_TJ", tj)
    big, _ = W.design(dims=(40, 64, 140), dx=1e-3, pml=5, box=((12, 20), (20, 30), (60, 80)), port=(9, 30),
                      probe=(30, 40), steps=130, sigma=0.02)
    wide, _ = W.design(dims=(48, 64, 140), dx=1e-3, pml=5, box=((8, 38), (12, 52), (40, 100)), port=(6, 30),
                       probe=(42, 40), steps=130, sigma=0.02)
    small, _ = W.design_small(steps=130)
    for args in (big, wide, small):
        g_march, g_gather = _backward_both_ways(gpu, args, monkeypatch)
        assert np.max(np.abs(g_gather)) > 0
        assert np.array_equal(g_march, g_gather)


def test_the_fused_adjoint_march_equals_the_gather_kernels_on_the_design_grid(gpu, monkeypatch):
    """The same equality on the inverse-design grid (192x192x112 nodes, CPML 12, u32 set indices,
    every block touching a CPML slab) over 1000 steps, long enough for the adjoint field to fill
    the box. A block that read its LDS tables before they were filled gave a gradient 1e-4 off
    here while the smaller grids above still agreed."""
    args, _ = W.design(steps=1000)
    g_march, g_gather = _backward_both_ways(gpu, args, monkeypatch, every=100)
    assert np.max(np.abs(g_gather)) > 0
    assert np.array_equal(g_march, g_gather)


def _grids_of_every_class():
    big, _ = W.design(dims=(40, 64, 140), dx=1e-3, pml=5, box=((12, 20), (20, 30), (60, 80)), port=(9, 30),
                      probe=(30, 40), steps=130, sigma=0.02)
    wide, _ = W.design(dims=(48, 64, 140), dx=1e-3, pml=5, box=((8, 38), (12, 52), (40, 100)), port=(6, 30),
                       probe=(42, 40), steps=130, sigma=0.02)
    small, _ = W.design_small(steps=130)
    return [("40x64x140", big), ("48x64x140 u32 sets", wide), ("30x28x26", small)]


def _pairs_and_singles(gpu, args, monkeypatch, every):
    d = make(gpu, args, every=every)
    try:
        s = d.forward()
        _, dJ = W.dft_objective(s[:, 0], F.time_step(args["grid"]), freq_of(args))
        monkeypatch.setenv("FIELDGRID_ADJ_PAIRS", "1")
        g2 = d.backward(dJ)
        monkeypatch.setenv("FIELDGRID_ADJ_PAIRS", "0")
        g1 = d.backward(dJ)
        return g2, g1
    finally:
        d.close()


@pytest.mark.parametrize("seg", ["16", "5", "2", "1"])
def test_two_adjoint_steps_per_pass_equal_one(gpu, monkeypatch, seg):
    """k_adj_march2 (adjoint steps n and n-1 in one pass, lam^n in registers and LDS) against
    k_adj_march (one step per pass), bit for bit: two-step segments of 16 down to 1 plane (warm-up
    planes clamped at plane 0, the last segment ending at the grid's last plane), snapshot
    intervals of 40 and 7 steps (odd segments leave one single step at each segment's start), on
    grids with interior, z-slab and x/y-slab blocks, u16 and u32 coefficient-set indices."""
    monkeypatch.setenv("FIELDGRID_ADJ_SEG2", seg)
    for name, args in _grids_of_every_class():
        for every in (40, 7):
            print(f"two-step adjoint, {name}, planes per segment {seg}, snapshots every {every}", flush=True)
            g2, g1 = _pairs_and_singles(gpu, args, monkeypatch, every)
            assert np.max(np.abs(g1)) > 0
            assert np.array_equal(g2, g1), (name, every)


def test_two_adjoint_steps_per_pass_equal_one_on_the_design_grid(gpu, monkeypatch):
    """The same on the inverse-design grid (192x192x112 nodes, CPML 12, u32 set indices) over
    1000 steps."""
    args, _ = W.design(steps=1000)
    g2, g1 = _pairs_and_singles(gpu, args, monkeypatch, 100)
    assert np.max(np.abs(g1)) > 0
    assert np.array_equal(g2, g1)


# ---- Debye media, inductors and thin wires in the adjoint ----------------------------------------

def test_the_media_gradient_matches_the_fp64_adjoint(gpu):
    """The design problem with a Debye block (pole states are state), a lumped inductor (its
    current is state) and a thin wire between the port and the observation: the forward is the
    FP32 path bit for bit, and the gradient agrees with the FP64 oracle adjoint (whose media
    transposes are pinned by central differences in tests/test_adjoint_oracle.py)."""
    args, freq = W.design_small(steps=300, media=True)
    d = make(gpu, args, every=64)
    try:
        s = d.forward()
        ex = F.Extras(max_steps=args["steps"], decay_db=math.inf, probes=args["obs"])
        o32 = O.run_cpu(d.desc, args["driven"], args["pulse"], ex, 32, fields=False)
        assert np.array_equal(s[:, 0], np.asarray(o32.probes, np.float32))
        J, dJ = W.dft_objective(s[:, 0], F.time_step(args["grid"]), freq)
        g = d.backward(dJ)
        go, _ = O.design_gradient(d.desc, args["driven"], args["pulse"], args["steps"], args["obs"], args["lo"],
                                  args["hi"], dJ)
        rel = np.max(np.abs(g - go)) / np.max(np.abs(go))
        print(f"\nmedia: J = {J:.4e}; max |grad - fp64 adjoint| / max |fp64 adjoint| = {rel:.2e}")
        assert rel < 1e-3
    finally:
        d.close()


def test_the_media_gradient_matches_central_differences_of_the_f64_forward(gpu):
    args, freq = W.design_small(steps=300, sigma=0.02, media=True)
    d = make(gpu, args, every=50)
    try:
        s = d.forward()
        dt = F.time_step(args["grid"])
        _, dJ = W.dft_objective(s[:, 0], dt, freq)
        g = d.backward(dJ)
        box = box_of(args).astype(np.float64)
        rng = np.random.default_rng(4)
        h = 1e-3
        for _ in range(5):
            c = tuple(int(rng.integers(0, n)) for n in box.shape)
            up, dn = box.copy(), box.copy()
            up[c] += h
            dn[c] -= h
            fd = (W.dft_objective(d.forward64(up)[:, 0], dt, freq)[0] -
                  W.dft_objective(d.forward64(dn)[:, 0], dt, freq)[0]) / (2 * h)
            print(f"media cell {c}: adjoint {g[c]:.6e}  central difference {fd:.6e}")
            assert abs(g[c] - fd) < 1e-3 * max(abs(fd), 1e-3 * np.max(np.abs(g)))
    finally:
        d.close()


def test_the_media_state_survives_the_snapshots(gpu):
    """The pole states and inductor currents are part of every snapshot: the gradient does not
    depend on the snapshot interval."""
    args, freq = W.design_small(steps=200, media=True)
    grads = []
    for every in (7, 200):
        d = make(gpu, args, every=every)
        s = d.forward()
        _, dJ = W.dft_objective(s[:, 0], F.time_step(args["grid"]), freq)
        grads.append(d.backward(dJ))
        d.close()
    assert np.max(np.abs(grads[0])) > 0
    assert np.array_equal(grads[0], grads[1])


# ---- the whole tape: the forward tapes every step, the backward recomputes nothing ---------------

def test_the_segment_tuner_leaves_the_gradient_alone(gpu, monkeypatch):
    """A design run of 800 steps or more times the forward march's segment lengths on its first
    passes (the checkpointed forward through enqueue_steps, the whole-tape forward on its taped
    passes); the backward's recompute and adjoint sweep keep the untuned length. On the
    192x192x112 design grid (16-plane segments untuned, so 16, 12, 8, 6 and 4 are timed) over
    900 steps: series and gradient bit-identical, tuned or not (FIELDGRID_SEG2_TUNE=0),
    checkpointed or with the whole tape, and a second backward equal to the first (the tuned
    march restored after the recompute)."""
    args, freq = W.design(steps=900)
    dt = F.time_step(args["grid"])
    out = {}
    for tune in ("1", "0"):
        monkeypatch.setenv("FIELDGRID_SEG2_TUNE", tune)
        for whole in (False, True):
            d = make(gpu, args, every=100, whole_tape=whole)
            try:
                s1 = d.forward()
                _, dJ = W.dft_objective(s1[:, 0], dt, freq)
                g1 = d.backward(dJ)
                s2 = d.forward()
                g2 = d.backward(dJ)
                assert np.array_equal(s1, s2) and np.array_equal(g1, g2), (tune, whole)
                out[(tune, whole)] = (s1, g1)
            finally:
                d.close()
    s0, g0 = out[("0", False)]
    assert np.max(np.abs(g0)) > 0
    for key, (s1, g1) in out.items():
        assert np.array_equal(s1, s0), key
        assert np.array_equal(g1, g0), key


def test_the_whole_tape_gives_the_checkpointed_gradient(gpu):
    """fg_design.whole_tape against checkpoint-and-recompute, bit for bit (series and gradient):
    the tape holds the same design-edge values whichever pass writes it, and the adjoint adds
    them in the same order. On the small design problem with and without media (a Debye block,
    an inductor, a thin wire: one step per pass in the adjoint), with odd step counts, and on a
    grid with every block class."""
    small, freq_s = W.design_small(steps=157)
    media, _ = W.design_small(steps=131, media=True)
    big, _ = W.design(dims=(40, 64, 140), dx=1e-3, pml=5, box=((12, 20), (20, 30), (60, 80)), port=(9, 30),
                      probe=(30, 40), steps=130, sigma=0.02)
    for name, args in (("small", small), ("media", media), ("40x64x140", big)):
        ck, wt = make(gpu, args, every=40), make(gpu, args, every=40, whole_tape=True)
        try:
            s1, s2 = ck.forward(), wt.forward()
            assert np.array_equal(s1, s2), name
            _, dJ = W.dft_objective(s1[:, 0], F.time_step(args["grid"]), freq_of(args))
            g1, g2 = ck.backward(dJ), wt.backward(dJ)
            (_, _, rec, _), snaps = wt.timing()
            print(f"\n{name}: max |grad| {np.max(np.abs(g1)):.3e}, whole-tape snapshots {snaps}, recompute {rec:.1f} ms")
            assert np.max(np.abs(g1)) > 0
            assert np.array_equal(g1, g2), name
            assert snaps == 0 and rec == 0.0
        finally:
            ck.close()
            wt.close()
