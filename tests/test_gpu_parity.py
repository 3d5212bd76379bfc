"""GPU parity: the HIP engine through the C-ABI against the oracle on the same inputs.

Tolerances: the reference's own CPU-vs-GPU bar, worst = max|a-b|/max|a| < 1e-5 on port
series, plane DFT and NTFF (crates/sims-field/tests/fdtd.rs:127-174), and relative L2 < 1e-5
on the final E and H fields against the FP32 oracle. Errors against the FP64 oracle are
printed and bounded at 1e-4 (the FP32 path itself differs from FP64 by ~1e-6).
"""
import math

import numpy as np
import pytest

from fieldgrid_amd import abi
from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W
from oracle import oracle as O

from .fixtures import everything, everything_extras, rel_l2, worst

ENGINES = [abi.FG_ENGINE_FUSED, abi.FG_ENGINE_SPLIT]
NAMES = {abi.FG_ENGINE_FUSED: "fused", abi.FG_ENGINE_SPLIT: "split"}


def peak_relative(rec, ref):
    """||(dE, eta dH)|| over the run's peak ||(E, eta H)||. The peak comes from the reference's
    own bookkeeping: decay_db = 10 log10(peak energy / final energy) (run.rs:739-740)."""
    eta = 376.73
    final = np.sqrt(np.sum(ref.e.astype(np.float64) ** 2) + np.sum((eta * ref.h.astype(np.float64)) ** 2))
    scale = final * 10.0 ** (ref.decay_db / 20.0)
    diff = np.sqrt(np.sum((rec.e.astype(np.float64) - ref.e) ** 2) + np.sum((eta * (rec.h.astype(np.float64) - ref.h)) ** 2))
    return float(diff / scale)


def _compare(rec, o32, o64, what, limit=1e-5):
    """Integrated outputs (series, plane, ntff) must meet the reference's 1e-5 bar against the
    FP32 oracle. Pointwise probe series must be no further from the FP64 oracle than twice
    the FP32 oracle's own distance (plus 1e-6), and within 1e-4 of the FP32 oracle. Final
    fields must be within 1e-5 of the FP32 oracle relative to the run's peak field norm."""
    print(f"\n{what}: steps {rec.steps}")
    for name in ("series", "plane", "ntff", "probes"):
        a, b = getattr(o32, name), getattr(rec, name)
        if len(a) == 0:
            continue
        w = worst(a, b)
        w64 = worst(getattr(o64, name), b)
        ref64 = worst(getattr(o64, name), a)
        print(f"  {name}: worst vs fp32 oracle {w:.2e}, vs fp64 {w64:.2e} (fp32 oracle vs fp64 {ref64:.2e})")
        if name == "probes":
            assert w < 1e-4 and w64 <= 2.0 * ref64 + 1e-6, f"{what} {name}: {w}, {w64}"
        else:
            assert w < limit, f"{what} {name}: {w}"
            assert w64 < 1e-4
    if rec.e is not None and o32.e is not None:
        d32, d64, ref64 = peak_relative(rec, o32), peak_relative(rec, o64), peak_relative(o32, o64)
        print(f"  fields: peak-relative L2 vs fp32 oracle {d32:.2e}, vs fp64 {d64:.2e} "
              f"(fp32 oracle vs fp64 {ref64:.2e})")
        assert d32 < limit, f"{what} fields: {d32}"


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ENGINES, ids=lambda e: NAMES[e])
def test_plain_vacuum_matches_the_oracle(gpu, engine):
    """The 64^3 plain workload in full: 500 steps, Ez/Hy at 8 probe nodes, final fields."""
    sim, probes = W.plain()
    ex = F.Extras(max_steps=500, decay_db=math.inf, check_every=100, probes=probes, engine=engine)
    pulse = W.plain_pulse()
    rec = F.run(gpu, sim, 0, pulse, ex, fields=True)
    o32 = O.run_cpu(sim.desc, 0, pulse, ex, 32)
    o64 = O.run_cpu(sim.desc, 0, pulse, ex, 64)
    assert rec.steps == 500
    assert np.abs(rec.probes).max() > 0
    _compare(rec, o32, o64, f"plain {NAMES[engine]}")


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ENGINES, ids=lambda e: NAMES[e])
def test_plain_vacuum_fields_mid_run(gpu, engine):
    """At 250 steps the pulse is still inside the grid, so E and H are well conditioned:
    each within 1e-5 relative L2 of the FP32 oracle."""
    sim, probes = W.plain()
    ex = F.Extras(max_steps=250, decay_db=math.inf, check_every=50, probes=probes, engine=engine)
    rec = F.run(gpu, sim, 0, W.plain_pulse(), ex, fields=True)
    o32 = O.run_cpu(sim.desc, 0, W.plain_pulse(), ex, 32)
    o64 = O.run_cpu(sim.desc, 0, W.plain_pulse(), ex, 64)
    for name in ("e", "h"):
        l32 = rel_l2(getattr(o32, name), getattr(rec, name))
        l64 = rel_l2(getattr(o64, name), getattr(rec, name))
        print(f"{name}: rel L2 vs fp32 oracle {l32:.2e}, vs fp64 {l64:.2e}")
        assert l32 < 1e-5 and l64 < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("driven", [0, 1])
@pytest.mark.parametrize("engine", ENGINES, ids=lambda e: NAMES[e])
def test_every_device_feature_matches_the_oracle(gpu, engine, driven):
    """tests/fdtd.rs:147-174: Debye band poles, PEC, C/R/L lumped, two multi-column ports,
    resistive sheets, thin wires, plane DFT, NTFF; both driven ports."""
    sim = everything()
    ex = everything_extras(sim, 1500)
    ex.engine = engine
    pulse = F.Pulse(5e9, 5e9)
    rec = F.run(gpu, sim, driven, pulse, ex, fields=True)
    o32 = O.run_cpu(sim.desc, driven, pulse, ex, 32)
    o64 = O.run_cpu(sim.desc, driven, pulse, ex, 64)
    assert rec.steps == 1500 and o32.steps == 1500
    _compare(rec, o32, o64, f"everything {NAMES[engine]} driven {driven}")


@pytest.mark.gpu
def test_fused_and_split_engines_agree(gpu):
    """tests/fdtd.rs:164-167: separate against fused, within 1e-5."""
    sim = everything(sheets=False, wires=False)
    pulse = F.Pulse(5e9, 5e9)
    recs = []
    for engine in ENGINES:
        ex = everything_extras(sim, 1500)
        ex.engine = engine
        recs.append(F.run(gpu, sim, 0, pulse, ex))
    for name in ("series", "plane", "ntff"):
        assert worst(getattr(recs[1], name), getattr(recs[0], name)) < 1e-5


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ENGINES, ids=lambda e: NAMES[e])
def test_sheets_and_wires_run_on_the_device(gpu, engine):
    """tests/fdtd.rs:43-174 with every feature, resistive sheets (two adaptive, the reference's
    branch table) and thin wires included, both ports driven: bit-identical to the FP32 path.
    The fused engine adds the wire terms inside the march (every copy of the H node) and runs the
    sheets after it from the step's input E; the split engine runs them between its sweeps."""
    sim = everything()
    assert sim.sheets == 6 and sim.wires > 30
    for driven in (0, 1):
        ex = everything_extras(sim, 600)
        ex.engine = engine
        rec = F.run(gpu, sim, driven, F.Pulse(5e9, 5e9), ex, fields=True)
        _exact(rec, O.run_cpu(sim.desc, driven, F.Pulse(5e9, 5e9), ex, 32), f"everything with sheets, wires, driven {driven}")


@pytest.mark.gpu
@pytest.mark.parametrize("radius", [0.09e-3, 1.7e-3])
def test_thin_wires_at_either_end_of_their_radius_stay_stable_on_the_device(gpu, radius):
    """tests/fdtd.rs:238-270 on the device (40 000 steps), and its first 3000 steps bit-identical
    to the FP32 path."""
    from .test_oracle_sheets_wires import thin_wire_case
    sim, ex, pulse = thin_wire_case(radius)
    rec = F.run(gpu, sim, 0, pulse, ex)
    v = np.abs(np.asarray(rec.series).reshape(-1, 2)[:, 0])
    early, late = v[5000:10000].max(), v[35000:40000].max()
    print(f"\nradius {radius}: early {early:.3e} late {late:.3e}")
    assert late < 2.0 * early
    short, _, _ = thin_wire_case(radius, steps=3000)
    ex = F.Extras(max_steps=3000, decay_db=math.inf)
    _exact(F.run(gpu, short, 0, pulse, ex, fields=True), O.run_cpu(short.desc, 0, pulse, ex, 32), f"wire {radius}")


@pytest.mark.gpu
def test_a_run_stops_when_progress_says_so(gpu):
    """tests/fdtd.rs:176-187 on the device path."""
    sim = everything(sheets=False, wires=False)
    calls = []

    def progress(steps, decay):
        calls.append(steps)
        return len(calls) < 2

    with pytest.raises(abi.FieldgridError) as e:
        F.run(gpu, sim, 0, F.Pulse(5e9, 5e9), everything_extras(sim, 5000), progress)
    assert str(e.value) == "cancelled" and e.value.status == abi.FG_CANCELLED
    assert calls == [250, 500]


@pytest.mark.gpu
def test_a_run_past_the_courant_limit_is_reported_unstable(gpu):
    """run.rs:736-738: raising dt above the Courant limit makes the energy blow up."""
    sim, _ = W.plain(nodes=24, pml=4)
    p = sim.problem()
    p.dt = sim.dt * 1.5  # the reference's `pub dt` (sim.rs:176), edited after packing
    ex = F.Extras(max_steps=2000, decay_db=math.inf, check_every=200)
    keep = []
    x = ex.to_c(keep)
    rec, bufs = F._record_buffers(sim, ex, False)
    pl = abi.fg_pulse(10e9, 5e9)
    st = abi.lib().fg_run(gpu._h, p, 0, pl, x, abi.PROGRESS_FN(lambda u, s, d: 1), None, rec)
    assert st == abi.FG_UNSTABLE
    assert abi.lib().fg_last_error().decode() == "the FDTD run went unstable"


@pytest.mark.gpu
def test_decay_stops_a_run_early_like_the_oracle(gpu):
    """run.rs:745-747: a finite decay_db stops after pulse_end once the energy has decayed."""
    sim, _ = W.plain(nodes=32, pml=8)
    ex = F.Extras(max_steps=20000, decay_db=30.0, check_every=100)
    rec = F.run(gpu, sim, 0, W.plain_pulse(), ex)
    o = O.run_cpu(sim.desc, 0, W.plain_pulse(), ex, 32, fields=False)
    assert rec.steps == o.steps < 20000
    assert abs(rec.decay_db - o.decay_db) < 0.5


def _exact(rec, o32, what):
    """The device follows the FP32 path operation for operation (no contraction, the pulse and
    DFT phases from the same f32 expressions): every field value, probe sample and DFT
    accumulator is bit-identical. Port V/I are f64 sums of f32 products whose order differs
    (a wave reduction against a sequential loop), so they may differ in the last f32 bit."""
    for name in ("e", "h", "probes", "plane", "ntff"):
        a, b = getattr(rec, name), getattr(o32, name)
        if a is None or b is None or len(b) == 0:
            continue
        b = np.asarray(b, np.float32)
        diff = int(np.sum(a != b))
        print(f"  {what} {name}: {diff} of {a.size} values differ")
        assert diff == 0, f"{what} {name}"
    w = worst(o32.series, rec.series)
    print(f"  {what} series: worst {w:.2e}")
    assert w < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("engine", ENGINES, ids=lambda e: NAMES[e])
def test_the_device_reproduces_the_fp32_path_bit_for_bit(gpu, engine):
    sim, probes = W.plain(nodes=40, pml=6)
    ex = F.Extras(max_steps=300, decay_db=math.inf, probes=probes, engine=engine)
    rec = F.run(gpu, sim, 0, W.plain_pulse(), ex, fields=True)
    _exact(rec, O.run_cpu(sim.desc, 0, W.plain_pulse(), ex, 32), "plain")
    sim = everything(sheets=False, wires=False)
    for driven in (0, 1):
        ex = everything_extras(sim, 600)
        ex.engine = engine
        rec = F.run(gpu, sim, driven, F.Pulse(5e9, 5e9), ex, fields=True)
        _exact(rec, O.run_cpu(sim.desc, driven, F.Pulse(5e9, 5e9), ex, 32), f"everything driven {driven}")


@pytest.mark.gpu
@pytest.mark.parametrize("per_box_tau", [False, True], ids=["global-pole", "pole-per-box"])
def test_debye_boxes_match_the_oracle(gpu, per_box_tau):
    """The boxes recipe at 56^3 (CPML 6): dielectric and Debye boxes, one port, NTFF. With
    pole-per-box each Debye box keeps its own tau through per-edge pole sets (an extension the
    oracle restates the same way); either way the device is bit-identical to the FP32 path."""
    sim = W.boxes(nodes=56, pml=6, nboxes=8, ndebye=4, per_box_tau=per_box_tau)
    ex = W.boxes_extras(sim, steps=400, inset=2)
    rec = F.run(gpu, sim, 0, W.boxes_pulse(), ex, fields=True)
    o32 = O.run_cpu(sim.desc, 0, W.boxes_pulse(), ex, 32)
    assert sim.problem().n_debye > 0
    _exact(rec, o32, f"boxes {'per-box' if per_box_tau else 'global'}")


@pytest.mark.gpu
@pytest.mark.parametrize("per_box_tau", [False, True], ids=["global-pole", "pole-per-box"])
@pytest.mark.parametrize("seg2", ["32", "3"])
def test_two_steps_per_launch_with_debye_media_change_nothing(gpu, monkeypatch, seg2, per_box_tau):
    """march2 with the first step's Debye ADE inside the pass (every copy of a Debye edge applies
    it from the pole states; the owner writes the new states and E for the second step's sparse
    fixes): the boxes recipe at 56^3 without NTFF, against one step per pass and the FP32 oracle,
    bit for bit. An E probe on a Debye edge makes the run fall back to one step per pass (the
    first step's observers recompute without the ADE), which must change nothing either."""
    sim = W.boxes(nodes=56, pml=6, nboxes=8, ndebye=4, per_box_tau=per_box_tau)
    p = sim.problem()
    assert p.n_debye > 0
    pulse = W.boxes_pulse()
    n = sim.dims()
    nn = n[0] * n[1] * n[2]
    debye = {int(p.debye_edges[t].f) for t in range(p.n_debye)}
    c, rest = divmod(min(debye), nn)
    on_debye = (c, rest // (n[1] * n[2]), (rest // n[2]) % n[1], rest % n[2])
    # the first probe set and the port stay off Debye edges, so that run takes two steps per pass
    port_edges = {int(p.ports[q].comp) * nn + int(p.port_edges[p.ports[q].first_edge + e].id)
                  for q in range(p.n_ports) for e in range(p.ports[q].n_edges)}
    assert not port_edges & debye
    assert not {2 * nn + (30 * n[1] + 28) * n[2] + 20} & debye
    monkeypatch.setenv("FIELDGRID_SEG2", seg2)
    for probes in ([(2, 30, 28, 20), (3, 25, 30, 31)], [(2, 30, 28, 20), on_debye]):
        ex = F.Extras(max_steps=351, decay_db=math.inf, check_every=100, probes=probes)
        monkeypatch.delenv("FIELDGRID_FUSE2", raising=False)
        two = F.run(gpu, sim, 0, pulse, ex, fields=True)
        monkeypatch.setenv("FIELDGRID_FUSE2", "0")
        one = F.run(gpu, sim, 0, pulse, ex, fields=True)
        for name in ("series", "probes", "e", "h"):
            a, b = getattr(two, name), getattr(one, name)
            assert np.array_equal(a, b), f"{name}: {int(np.sum(a != b))} of {a.size} differ"
        _exact(two, O.run_cpu(sim.desc, 0, pulse, ex, 32), f"boxes seg2 {seg2} probes {probes}")


@pytest.mark.gpu
@pytest.mark.parametrize("seg2", ["32", "3"])
def test_two_steps_per_launch_with_ntff_and_plane_dft_change_nothing(gpu, monkeypatch, seg2):
    """The first step's plane-DFT and NTFF observers read values the pass captured at their points
    (Cap2: each point's owner writes step n's E/H): the boxes recipe at 56^3 with Debye boxes,
    NTFF at three frequencies and a plane DFT, against one step per pass and the FP32 oracle,
    bit for bit."""
    sim = W.boxes(nodes=56, pml=6, nboxes=8, ndebye=4, per_box_tau=True)
    pulse = W.boxes_pulse()
    ex = W.boxes_extras(sim, steps=351, inset=2, check_every=100, plane_k=23, probes=[(2, 30, 28, 20), (3, 25, 30, 31)])
    monkeypatch.setenv("FIELDGRID_SEG2", seg2)
    monkeypatch.delenv("FIELDGRID_FUSE2", raising=False)
    two = F.run(gpu, sim, 0, pulse, ex, fields=True)
    monkeypatch.setenv("FIELDGRID_FUSE2", "0")
    one = F.run(gpu, sim, 0, pulse, ex, fields=True)
    for name in ("series", "probes", "plane", "ntff", "e", "h"):
        a, b = getattr(two, name), getattr(one, name)
        assert a.size > 0 and np.array_equal(a, b), f"{name}: {int(np.sum(a != b))} of {a.size} differ"
    _exact(two, O.run_cpu(sim.desc, 0, pulse, ex, 32), f"boxes with NTFF and plane DFT, seg2 {seg2}")


@pytest.mark.gpu
@pytest.mark.parametrize("seg2", ["16", "5"])
def test_side_work_at_tile_edges_changes_nothing(gpu, monkeypatch, seg2):
    """A side-work kernel runs the lean march for every block whose tile holds no Debye edge in
    the planes, rows and lanes it computes and no captured point it owns. Debye slabs one and two
    cells thick at the k-tile edges (cells 57-58 and 60-61 around the first tile's 60 owned lanes
    and two-lane halo, 118-120 around the second's), across a row-tile edge (cells 11-12) and one
    plane thick at a segment edge, with NTFF patches and a plane DFT: two steps per pass against
    one step per pass and the FP32 oracle, bit for bit, at 16- and 5-plane segments."""
    grid = W._grid([41, 50, 135], 1e-3, 6)
    n = grid.dims()
    cells = (n[0] - 1, n[1] - 1, n[2] - 1)
    mats = F.Materials.new(grid)
    eps, inf, dd = (getattr(mats, a).reshape(cells) for a in ("eps", "inf", "debye"))
    for sl in ((slice(8, 30), slice(8, 40), slice(57, 59)), (slice(10, 13), slice(8, 40), slice(60, 62)),
               (slice(15, 33), slice(11, 13), slice(100, 121)), (slice(16, 17), slice(20, 30), slice(20, 40))):
        eps[sl], inf[sl], dd[sl] = 3.0, 3.0, 2.0
    media = F.Media(debye=F.Debye([(1.0 / 20e-12, 1.0)]))
    port = F.PortDef("p1", [[F.Edge(2, (20, 25, k)) for k in range(80, 84)]], 50.0, False)
    sim = F.Sim.try_new(grid, mats, None, [], [port], [], media)
    assert sim.problem().n_debye > 0
    pulse = W.boxes_pulse()
    ex = W.boxes_extras(sim, steps=351, inset=2, check_every=100, plane_k=70, probes=[(2, 30, 28, 90), (3, 25, 30, 31)])
    monkeypatch.setenv("FIELDGRID_SEG2", seg2)
    monkeypatch.delenv("FIELDGRID_FUSE2", raising=False)
    two = F.run(gpu, sim, 0, pulse, ex, fields=True)
    monkeypatch.setenv("FIELDGRID_FUSE2", "0")
    one = F.run(gpu, sim, 0, pulse, ex, fields=True)
    for name in ("series", "probes", "plane", "ntff", "e", "h"):
        a, b = getattr(two, name), getattr(one, name)
        assert a.size > 0 and np.array_equal(a, b), f"{name}: {int(np.sum(a != b))} of {a.size} differ"
    _exact(two, O.run_cpu(sim.desc, 0, pulse, ex, 32), f"side work at tile edges, seg2 {seg2}")


# the two-step pass's dispatch: one launch over every block class (mixed), one launch per class
# on the session stream (classes), or one per class on side streams forked from it (streams)
_DISPATCH = {"mixed": {"FIELDGRID_MIX2": "1", "FIELDGRID_STREAMS": "0"},
             "classes": {"FIELDGRID_MIX2": "0", "FIELDGRID_STREAMS": "0"},
             "streams": {"FIELDGRID_MIX2": "0", "FIELDGRID_STREAMS": "1"}}


@pytest.mark.gpu
@pytest.mark.parametrize("seg2,dispatch", [("32", "mixed"), ("5", "mixed"), ("1", "mixed"),
                                           ("5", "classes"), ("1", "streams")])
def test_two_steps_per_launch_change_nothing(gpu, monkeypatch, seg2, dispatch):
    """march2 (two steps per pass, the intermediate step in registers) against one step per
    pass and against the FP32 oracle: every output bit-identical. Odd grid sizes, an odd
    check_every (an odd step count per chunk ends on a single step), segments of 1, 5 and 32
    planes (segment starts at plane 0, 1 and further in), CPML 6, two ports, E and H probes,
    under each of the pass's three dispatches."""
    sim = W.layered(dims=(70, 90, 66), dx=1e-3, pml=6, port_i=(20, 41))
    pulse = W.layered_pulse()
    ex = F.Extras(max_steps=301, decay_db=math.inf, check_every=77,
                  probes=[(2, 30, 40, 33), (4, 50, 52, 20), (0, 35, 45, 10)])
    monkeypatch.setenv("FIELDGRID_SEG2", seg2)
    for k, v in _DISPATCH[dispatch].items():
        monkeypatch.setenv(k, v)
    two = F.run(gpu, sim, 0, pulse, ex, fields=True)
    monkeypatch.setenv("FIELDGRID_FUSE2", "0")
    one = F.run(gpu, sim, 0, pulse, ex, fields=True)
    for name in ("series", "probes", "e", "h"):
        a, b = getattr(two, name), getattr(one, name)
        print(f"  {name}: {int(np.sum(a != b))} of {a.size} differ")
        assert np.array_equal(a, b), name
    _exact(two, O.run_cpu(sim.desc, 0, pulse, ex, 32), f"layered seg2 {seg2} {dispatch}")


@pytest.mark.gpu
@pytest.mark.parametrize("check_every", [1000, 37])
def test_the_segment_tuner_changes_nothing(gpu, monkeypatch, check_every):
    """A run of 800 steps or more times a few two-step passes of several segment lengths before
    it goes on with the fastest (the timed passes are steps of the run). The boxes recipe at
    128^3 (8-plane segments by default, so 8, 6 and 4 are timed) with Debye boxes, NTFF and a
    plane DFT over 840 steps, tuned and untuned, the tuner's 20 steps taken from a first chunk of
    1000 or of 37 steps (an odd rest, so the chunk ends on a single step): every output
    bit-identical, and equal to the FP32 oracle."""
    sim = W.boxes(nodes=128, pml=8, nboxes=8, ndebye=4, per_box_tau=True)
    pulse = W.boxes_pulse()
    ex = W.boxes_extras(sim, steps=840, inset=2, check_every=check_every, plane_k=61,
                        probes=[(2, 60, 58, 40), (3, 55, 70, 71)])
    monkeypatch.delenv("FIELDGRID_SEG2", raising=False)
    monkeypatch.delenv("FIELDGRID_SEG2_TUNE", raising=False)
    tuned = F.run(gpu, sim, 0, pulse, ex, fields=True)
    monkeypatch.setenv("FIELDGRID_SEG2_TUNE", "0")
    plain = F.run(gpu, sim, 0, pulse, ex, fields=True)
    for name in ("series", "probes", "plane", "ntff", "e", "h"):
        a, b = getattr(tuned, name), getattr(plain, name)
        assert a.size > 0 and np.array_equal(a, b), f"{name}: {int(np.sum(a != b))} of {a.size} differ"
    if check_every == 1000:
        _exact(tuned, O.run_cpu(sim.desc, 0, pulse, ex, 32), "boxes 128^3 with the segment tuner")


# (n0, n1, n2) around the march tiles' edges: k extents just under, at and over one and two
# 60-lane tiles and 64 lanes, row counts at and just over multiples of the 12- and 14-row tiles,
# plane counts under one segment and over two
RAGGED = [(6, 7, 64), (9, 13, 61), (17, 25, 62), (33, 12, 121), (40, 14, 120), (12, 40, 59), (65, 30, 8),
          (20, 61, 125)]


@pytest.mark.gpu
@pytest.mark.parametrize("seg2", [None, "1"], ids=["seg-default", "seg1"])
@pytest.mark.parametrize("dims", RAGGED, ids=lambda d: "x".join(map(str, d)))
def test_ragged_grids_match_the_fp32_path(gpu, monkeypatch, dims, seg2):
    """Layered grids of ragged shapes with CPML 2 and two ports, 41 steps (an odd count: the
    last step takes a pass of its own), on the default engine with its own segment length and
    with one-plane segments: every field value bit-identical to the FP32 oracle, port series to
    its last bit."""
    if seg2:
        monkeypatch.setenv("FIELDGRID_SEG2", seg2)
    sim = W.layered(dims=dims, pml=2, port_i=(1, dims[0] // 2))
    pulse = W.layered_pulse()
    ex = F.Extras(max_steps=41, decay_db=math.inf, check_every=13)
    rec = F.run(gpu, sim, 0, pulse, ex, fields=True)
    _exact(rec, O.run_cpu(sim.desc, 0, pulse, ex, 32), "x".join(map(str, dims)))
