"""The committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from
the FP32 oracle) against the oracle as built now (CPU) and against the HIP engine (GPU).

Fields are compared by the SHA-256 of their f32 bytes, traces value for value; the device's port
V/I are f64 sums of the same f32 products in another order, so they are held to worst < 1e-6
(everything else is bit-identical)."""
import os

import numpy as np
import pytest

from fieldgrid_amd import fdtd as F
from oracle import oracle as O

from .fixtures import worst
from .golden.make_golden import cases, record_arrays

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = cases()


def _load(name):
    with np.load(os.path.join(HERE, f"{name}.npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _compare(rec, n, gold, series_exact):
    arrays, digests = record_arrays(rec, n)
    assert int(arrays["steps"][0]) == int(gold["steps"][0])
    for key, value in digests.items():
        assert value == str(gold[key]), key
    for key, value in arrays.items():
        if key == "steps":
            continue
        g = gold[key]
        assert value.shape == g.shape, key
        if key == "series" and not series_exact:
            assert worst(g, value) < 1e-6, key
        else:
            assert np.array_equal(value, g), f"{key}: {int(np.sum(value != g))} of {g.size} differ"


@pytest.mark.parametrize("name", sorted(CASES))
def test_the_oracle_reproduces_its_golden_fixture(name):
    sim, driven, pulse, ex = CASES[name]
    rec = O.run_cpu(sim.desc, driven, pulse, ex, 32)
    _compare(rec, sim.dims(), _load(name), series_exact=True)


@pytest.mark.gpu
@pytest.mark.parametrize("engine", [F.abi.FG_ENGINE_FUSED, F.abi.FG_ENGINE_SPLIT], ids=["fused", "split"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_the_device_reproduces_the_golden_fixture(gpu, name, engine):
    sim, driven, pulse, ex = CASES[name]
    ex.engine = engine
    rec = F.run(gpu, sim, driven, pulse, ex, fields=True)
    _compare(rec, sim.dims(), _load(name), series_exact=False)
