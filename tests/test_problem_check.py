"""fg_problem_check: the host-side checks fg_run makes before it allocates, with no device.

The march kernels address each plane of a field, coefficient-index or psi array with a 32-bit
byte offset from a plane-uniform base, and send lanes that must not touch memory to an offset
(0x7FFFFFF0) past every plane. A grid whose n1*n2 plane holds more than 0x7FFFFFF0 / 4 =
536 870 908 nodes is refused with FG_TOO_LARGE instead of running with wrapped offsets.
These checks read only the problem's sizes and small tables, so they run here on the CPU.
"""
import ctypes as C
import math

from fieldgrid_amd import abi
from fieldgrid_amd import fdtd as F
from fieldgrid_amd import workloads as W

LIMIT = 0x7FFFFFF0 // 4


def _check(problem, extras=None):
    keep = []
    x = (extras or F.Extras(max_steps=4, decay_db=math.inf)).to_c(keep)
    st = abi.lib().fg_problem_check(C.byref(problem), 0, C.byref(x))
    return st, abi.lib().fg_last_error().decode()


def _resized(p, n):
    q = abi.fg_problem()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(p))
    for a in range(3):
        q.n[a] = n[a]
    return q


def test_a_packed_problem_passes_the_check():
    sim, _ = W.plain(nodes=24, pml=4)
    st, msg = _check(sim.problem())
    assert st == abi.FG_OK, msg


def test_a_plane_past_the_32_bit_offsets_is_refused():
    sim, _ = W.plain(nodes=24, pml=4)
    p = sim.problem()
    # the largest plane the kernels address, and one node more (the check reads no table that
    # scales with the grid: ids of the small tables are compared against the new sizes only)
    at = _resized(p, (24, 16384, LIMIT // 16384))
    assert at.n[1] * at.n[2] <= LIMIT
    st, msg = _check(at)
    assert st == abi.FG_OK, msg
    past = _resized(p, (24, 2, LIMIT // 2 + 1))
    st, msg = _check(past)
    assert st == abi.FG_TOO_LARGE, msg
    assert "more than this engine addresses" in msg
    # a flat grid of 536 M-node planes with only three of them is refused too
    st, msg = _check(_resized(p, (3, 32768, 32768)))
    assert st == abi.FG_TOO_LARGE, msg


def test_a_bad_driven_port_is_refused_without_a_device():
    grid = F.Grid.uniform([0.0] * 3, [15e-3] * 3, [15] * 3, 2)
    port = F.PortDef("p", [[F.Edge(2, (8, 8, 8))]], 50.0, False)
    sim = F.Sim.try_new(grid, F.Materials.new(grid), None, [], [port], [], F.Media())
    keep = []
    x = F.Extras(max_steps=4, decay_db=math.inf).to_c(keep)
    st = abi.lib().fg_problem_check(C.byref(sim.problem()), 1, C.byref(x))
    assert st == abi.FG_INVALID
    assert "driven port 1 does not exist" in abi.lib().fg_last_error().decode()
