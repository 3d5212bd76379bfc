"""The C-ABI library loads without a GPU and exports every symbol include/*.h declares."""
import ctypes as C
import os
import re

import pytest

from fieldgrid_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[a-z_][\w\s\*]*?\b(fg_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_engine_and_builder_entry_points():
    names = declared_functions()
    for must in ("fg_open", "fg_close", "fg_name", "fg_max_cells", "fg_run", "fg_last_error",
                 "fg_sim_new", "fg_sim_pack", "fg_patches", "fg_run_batch"):
        assert must in names


@pytest.mark.parametrize("name", declared_functions())
def test_every_declared_symbol_is_exported(name):
    lib = C.CDLL(abi.LIB_PATH)
    assert hasattr(lib, name), name
    assert name in abi.SIGNATURES, f"{name} has no ctypes signature"


def test_errors_carry_the_reference_text_without_a_gpu():
    """A bad problem fails in the builder with the library's message (no device needed)."""
    from fieldgrid_amd import fdtd as F
    import numpy as np
    grid = F.Grid(*(np.arange(11) * 1e-3 for _ in range(3)), 0)
    with pytest.raises(abi.FieldgridError) as e:
        F.Sim.try_new(grid, F.Materials.new(grid), None, [], [], [],
                      F.Media(wires=[(F.Edge(2, (5, 5, 4)), 0.6e-3)]))
    assert e.value.status == abi.FG_INVALID
    assert "from a 55th to half of its cell" in str(e.value)
