"""Extras.profile: fg_run prints the mean kernel times of its first chunk's steps from step 200 on,
in the reference's wording (crates/sims-field/src/fdtd/run.rs:152-176, 686-732), and the run's
record is the one an unprofiled run gives."""
import math
import re

import numpy as np
import pytest

from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("check_every,timed", [(400, True), (150, False)])
def test_profile_prints_kernel_means(gpu, capfd, check_every, timed):
    sim, probes = W.plain(nodes=40, pml=6)
    base = dict(max_steps=600, decay_db=math.inf, check_every=check_every, probes=probes)
    plain = F.run(gpu, sim, 0, W.plain_pulse(), F.Extras(**base))
    capfd.readouterr()
    prof = F.run(gpu, sim, 0, W.plain_pulse(), F.Extras(profile=True, **base))
    err = capfd.readouterr().err
    print(err)
    assert np.array_equal(plain.series, prof.series) and np.array_equal(plain.probes, prof.probes)
    lines = re.findall(r"fdtd kernel (.+) \(every step\): ([0-9.]+) us over (\d+) steps", err)
    if not timed:  # a first chunk of 200 steps or fewer has nothing to time, as in the reference
        assert not lines
        return
    assert [name.split()[0] for name, _, _ in lines] == ["k_march2", "k_fix", "k_observe"]
    assert all(int(n) == check_every - 200 for _, _, n in lines)
    total = float(re.search(r"fdtd kernel time per step: ([0-9.]+) us", err).group(1))
    assert total > 0 and abs(total - sum(float(us) for _, us, _ in lines)) < 0.2
