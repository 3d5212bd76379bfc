"""Edge cases of the device path against the FP32 oracle, bit for bit.

Grids smaller than one tile and grids whose extents sit on either side of a tile boundary (60
lanes of k and 12 or 14 rows of j per block, 3-32-plane segments of i), with and without CPML,
random per-cell permittivity; runs of 0, 1 and an odd number of steps (the two-step march takes
an odd last step alone); the same runs through a batch.
"""
import math

import numpy as np
import pytest

from fieldgrid_amd import abi
from fieldgrid_amd import fdtd as F
from oracle import oracle as O

from .test_gpu_parity import _exact

pytestmark = pytest.mark.gpu

ENGINES = [abi.FG_ENGINE_FUSED, abi.FG_ENGINE_SPLIT]
NAMES = {abi.FG_ENGINE_FUSED: "fused", abi.FG_ENGINE_SPLIT: "split"}


def _sim(nodes, pml, seed=0, dx=1e-3):
    cells = [n - 1 for n in nodes]
    grid = F.Grid.uniform([0.0] * 3, [c * dx for c in cells], cells, pml)
    mats = F.Materials.new(grid)
    rng = np.random.default_rng(seed)
    mats.eps[:] = rng.uniform(1.0, 4.0, size=mats.eps.shape)
    n = grid.dims()
    c = (n[0] // 2, n[1] // 2, min(n[2] // 2, n[2] - 2))
    port = F.PortDef("src", [[F.Edge(2, c)]], 50.0, False)
    return F.Sim.try_new(grid, mats, None, [], [port], [], F.Media())


SHAPES = [((7, 13, 61), 2), ((33, 12, 60), 3), ((9, 25, 121), 0), ((40, 11, 5), 2), ((3, 3, 3), 0), ((66, 15, 64), 4)]


@pytest.mark.parametrize("engine", ENGINES, ids=lambda e: NAMES[e])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s[0])) + f"-pml{s[1]}")
def test_odd_shapes_follow_the_fp32_path_bit_for_bit(gpu, shape, engine):
    nodes, pml = shape
    sim = _sim(nodes, pml)
    pulse = F.Pulse(10e9, 5e9)
    ex = F.Extras(max_steps=61, decay_db=math.inf, engine=engine)
    rec = F.run(gpu, sim, 0, pulse, ex, fields=True)
    o32 = O.run_cpu(sim.desc, 0, pulse, ex, 32)
    assert rec.steps == 61 and np.abs(rec.e).max() > 0
    _exact(rec, o32, f"{nodes} pml {pml} {NAMES[engine]}")


@pytest.mark.parametrize("steps", [0, 1, 2, 3])
def test_runs_of_zero_to_three_steps(gpu, steps):
    sim = _sim((20, 18, 16), 3, seed=1)
    pulse = F.Pulse(10e9, 5e9)
    ex = F.Extras(max_steps=steps, decay_db=math.inf)
    rec = F.run(gpu, sim, 0, pulse, ex, fields=True)
    o32 = O.run_cpu(sim.desc, 0, pulse, ex, 32)
    assert rec.steps == steps and len(rec.series) == 2 * steps
    if steps == 0:  # nothing ran: the fields are the initial zeros
        assert not np.any(rec.e) and not np.any(rec.h) and not np.any(o32.e)
        return
    _exact(rec, o32, f"{steps} steps")


def test_a_batch_of_odd_shapes_and_lengths_equals_single_runs(gpu):
    pulse = F.Pulse(10e9, 5e9)
    runs = []
    for q, ((nodes, pml), steps) in enumerate(zip(SHAPES, [0, 1, 7, 61, 2, 30])):
        runs.append(F.BatchRun(_sim(nodes, pml, seed=q), 0, pulse, F.Extras(max_steps=steps, decay_db=math.inf), []))
    for tune in (F.Tune(tj=14, seg=16, chunk=1), F.Tune(seg=4, chunk=4, depth=2)):
        recs, _ = F.run_batch(gpu, runs, tune=tune)
        for r, (job, rec) in enumerate(zip(runs, recs)):
            one = F.run(gpu, job.sim, 0, pulse, job.extras)
            assert rec.steps == one.steps == job.extras.max_steps, f"run {r}"
            assert np.array_equal(rec.series, one.series), f"run {r}"
