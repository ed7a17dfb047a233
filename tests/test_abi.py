"""CPU tests of the drop-in boundary: the C-ABI library loads, exports exactly what
include/blpg/blpg.h declares, and rejects bad calls without a device (no compute here)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

import simplexprojects_amd as sa

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "blpg" / "blpg.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(blpg_[A-Za-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = sa.load_library()
    decl = declared_symbols()
    assert decl == sorted(sa.ABI_SYMBOLS)
    for name in decl:
        assert hasattr(lib, name), name


def test_struct_layouts_match_header():
    # blpg_lp: 2 int32 + 4 pointers + int32 + pointer; blpg_result: 4 int32 + 3 double + 2 ptr
    assert C.sizeof(sa._LP) == 8 + 4 * 8 + 8 + 8
    assert C.sizeof(sa._Result) == 16 + 24 + 16
    assert C.sizeof(sa._Options) == 4 * 4 + 4 * 8 + 8 + 4 * 4


def test_default_options():
    lib = sa.load_library()
    o = sa._Options()
    lib.blpg_default_options(C.byref(o))
    assert o.precision == sa.FP64 and o.verify == 1 and o.kernel == sa.KERNEL_AUTO
    assert o.eps_d == 0 and o.max_iter == 0


def test_version_and_null_context():
    lib = sa.load_library()
    assert b"gfx950" in lib.blpg_version()
    assert lib.blpg_solve_batch(None, None, 0, None) == sa.E_INVALID
    assert lib.blpg_batch_solve(None, None) == sa.E_INVALID


def test_gen_config_rejects_bad_sizes():
    lib = sa.load_library()
    buf = np.zeros(4)
    p = buf.ctypes.data_as(C.POINTER(C.c_double))
    assert lib.blpg_gen_config(2, 0, 0, 0, 1, 1, 0, 2, p, p, p, None, None) == sa.E_INVALID
    assert lib.blpg_gen_config(9, 0, 0, 0, 1, 1, 1, 1, p, p, p, None, None) == sa.E_INVALID


def test_lp_validation():
    import pytest
    with pytest.raises(ValueError):
        sa.LP(np.zeros((2, 3)), np.zeros(3), np.zeros(3))
    with pytest.raises(ValueError):
        sa.LP(np.zeros((2, 3)), np.zeros(2), np.zeros(3), sense=np.zeros(3))


def test_oracle_header_is_test_only():
    # the product package never loads the oracle
    src = "\n".join(p.read_text() for p in (ROOT / "simplexprojects_amd").rglob("*.py"))
    src += "\n".join(p.read_text() for p in (ROOT / "simplexprojects_amd" / "csrc").glob("*"))
    assert "import oracle" not in src and "liboracle" not in src
