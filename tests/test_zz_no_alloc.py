"""fg_port_spectra_static, the port-spectra entry point that allocates no device memory, against
fg_port_spectra (the same sums; bit-identical) on two calls in a row, the second smaller, and
its refusal of a call larger than its static scratch."""
import ctypes as C

import numpy as np
import pytest

from fieldgrid_amd import abi

pytestmark = pytest.mark.gpu


def _call(fn, gpu, series, ports, dt, freqs):
    s = np.ascontiguousarray(series, np.float32)
    steps = s.size // (2 * ports)
    f = np.ascontiguousarray(freqs, np.float64)
    v = np.full(ports * len(f) * 2, np.nan)
    i = np.full_like(v, np.nan)
    dp = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    st = fn(gpu._h, dp(s, C.c_float), steps, ports, float(dt), len(f), dp(f, C.c_double), dp(v, C.c_double),
            dp(i, C.c_double))
    return st, v, i


@pytest.mark.parametrize("order", ["static first", "allocating first"])
def test_the_static_spectra_equal_the_allocating_ones_on_two_calls(gpu, order):
    L = abi.lib()
    rng = np.random.default_rng(11)
    dt = 9.5329e-13
    cases = [(rng.standard_normal(3000 * 3 * 2), 3, np.linspace(1e9, 20e9, 64)),
             (rng.standard_normal(500 * 1 * 2), 1, np.linspace(2e9, 9e9, 8))]  # the second call smaller
    for series, ports, freqs in cases:
        print(f"no-alloc spectra: {series.size // (2 * ports)} steps, {ports} ports, {len(freqs)} freqs, {order}",
              flush=True)
        fns = [L.fg_port_spectra_static, L.fg_port_spectra]
        if order != "static first":
            fns.reverse()
        outs = {}
        for fn in fns:
            st, v, i = _call(fn, gpu, series, ports, dt, freqs)
            abi.check(st)
            outs[fn.__name__] = (v, i)
        (va, ia), (vb, ib) = outs["fg_port_spectra_static"], outs["fg_port_spectra"]
        assert np.all(np.isfinite(va)) and np.all(np.isfinite(ia))
        assert np.array_equal(va, vb) and np.array_equal(ia, ib)


def test_a_call_beyond_the_static_scratch_is_refused(gpu):
    L = abi.lib()
    ports, steps = 4, (32 << 20) // 32 + 1  # 8 * steps * ports bytes of series alone exceed 32 MiB
    series = np.zeros(steps * ports * 2, np.float32)
    st, _, _ = _call(L.fg_port_spectra_static, gpu, series, ports, 1e-12, [1e9])
    assert st == abi.FG_INVALID
    assert "static scratch" in L.fg_last_error().decode()
