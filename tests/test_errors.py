"""fg_run refuses malformed problems with FG_INVALID and a message, before any kernel reads them.

The product builder never makes these tables; a caller that packs its own (the C-ABI takes plain
pointers) could. Each case copies a valid packed problem and breaks one thing: a
coefficient-set index past the set table (the kernels read sets[index] unchecked), a CPML slot
that is not -1 or below 2 pml (slots become psi buffer offsets), a table with entries but a NULL
pointer.
"""
import ctypes as C
import math

import numpy as np
import pytest

from fieldgrid_amd import abi
from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W

pytestmark = pytest.mark.gpu


def _copy(p):
    q = abi.fg_problem()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(p))
    return q


def _run_with(gpu, sim, problem):
    sim._problem = (0, problem)  # the run reads the packed problem from the Sim's cache
    try:
        F.run(gpu, sim, 0, W.plain_pulse(), F.Extras(max_steps=4, decay_db=math.inf))
    finally:
        sim._problem = None


def _expect_invalid(gpu, sim, problem, text):
    with pytest.raises(abi.FieldgridError) as e:
        _run_with(gpu, sim, problem)
    print(f"\n  refused: {e.value}")
    assert e.value.status == abi.FG_INVALID and text in str(e.value)


@pytest.mark.parametrize("wide", [False, True], ids=["u16-index", "u32-index"])
def test_a_coefficient_index_past_the_set_table_is_refused(gpu, wide):
    sim, _ = W.plain(nodes=24, pml=4)
    p = sim.problem(narrow=1 if wide else 0)  # a narrow limit of 1 set makes the index u32
    nn = int(np.prod(sim.dims()))
    assert bool(p.coef_wide) == wide
    words = 3 * nn if p.coef_wide else (3 * nn + 1) // 2
    idx = np.ctypeslib.as_array(p.coef_index, shape=(words,)).copy()
    bad = idx.copy()
    if p.coef_wide:
        bad[words // 2] = p.n_sets
    else:
        bad[words // 2] = (bad[words // 2] & 0xFFFF) | (p.n_sets << 16)
    q = _copy(p)
    q.coef_index = bad.ctypes.data_as(C.POINTER(C.c_uint32))
    _expect_invalid(gpu, sim, q, "is past the")
    # the untouched table still runs
    q.coef_index = idx.ctypes.data_as(C.POINTER(C.c_uint32))
    _run_with(gpu, sim, q)


def test_a_cpml_slot_out_of_range_is_refused(gpu):
    sim, _ = W.plain(nodes=24, pml=4)
    p = sim.problem()
    rows = sum(sim.dims())
    axes = np.ctypeslib.as_array(p.axes, shape=(rows * 10,)).copy()
    for bad_slot in (2.0 * sim.pml, 1.5, float("nan")):
        ax = axes.copy()
        ax[8] = bad_slot  # slot_e of node 0 on axis 0
        q = _copy(p)
        q.axes = ax.ctypes.data_as(C.POINTER(C.c_float))
        _expect_invalid(gpu, sim, q, "CPML slot")


def test_a_table_with_entries_and_no_pointer_is_refused(gpu):
    sim, _ = W.plain(nodes=24, pml=4)
    p = sim.problem()
    assert p.n_port_edges > 0
    q = _copy(p)
    q.port_edges = None
    _expect_invalid(gpu, sim, q, "is NULL")


def test_a_bad_run_in_a_batch_is_refused_and_the_batch_runs_once_fixed(gpu):
    """The batch builds its runs' device state on several host threads; an error in one of them
    (here a coefficient index past the set table in run 6 of 10) comes back as that run's
    FG_INVALID once every thread has stopped, and nothing of the failed batch is left behind."""
    runs = W.sweep_runs(count=10, lo=24, hi=40, pml=4, nboxes=2, ndebye=1, steps=(20, 40), nfreq=4)
    sim = runs[6].sim
    p = sim.problem()
    nn = int(np.prod(sim.dims()))
    words = 3 * nn if p.coef_wide else (3 * nn + 1) // 2
    bad = np.ctypeslib.as_array(p.coef_index, shape=(words,)).copy()
    if p.coef_wide:
        bad[words // 3] = p.n_sets
    else:
        bad[words // 3] = (bad[words // 3] & 0xFFFF) | (p.n_sets << 16)
    q = _copy(p)
    q.coef_index = bad.ctypes.data_as(C.POINTER(C.c_uint32))
    tune = F.Tune(tj=14, seg=16, chunk=1)
    sim._problem = (0, q)
    try:
        with pytest.raises(abi.FieldgridError) as e:
            F.run_batch(gpu, runs, tune=tune)
    finally:
        sim._problem = None
    print(f"\n  refused: {e.value}")
    assert e.value.status == abi.FG_INVALID and "is past the" in str(e.value)
    recs, _ = F.run_batch(gpu, runs, tune=tune)
    for r in (0, 6, 9):
        one = F.run(gpu, runs[r].sim, runs[r].driven, runs[r].pulse, runs[r].extras)
        assert np.array_equal(recs[r].series, one.series), f"run {r}"


def test_a_batch_larger_than_the_device_is_refused_before_it_allocates(gpu):
    """The runs' sessions are built concurrently, so the batch checks the sum of their device
    bytes against the free memory first: 700 runs of one 200^3 problem (~0.5 GB each) ask for
    far more than one GPU holds."""
    sim, _ = W.plain(nodes=200, pml=8)
    runs = [F.BatchRun(sim, 0, W.plain_pulse(), F.Extras(max_steps=10, decay_db=math.inf), [])] * 700
    with pytest.raises(abi.FieldgridError) as e:
        F.run_batch(gpu, runs, tune=F.Tune(tj=14, seg=16, chunk=1))
    print(f"\n  refused: {e.value}")
    assert e.value.status == abi.FG_TOO_LARGE and "the batch's 700 runs need" in str(e.value)
