"""The C ABI library loads and exports every entry point declared in include/cbc.h
(no compute calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "cbc.h"
LIB = ROOT / "coupled-bubble-control_amd" / "lib" / "libcbc.so"


def declared_functions():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(cbc_[a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    for n in ("cbc_abi_version", "cbc_last_error", "cbc_plan_create", "cbc_plan_destroy", "cbc_launch",
              "cbc_solve", "cbc_solve_periodic", "cbc_plan_query"):
        assert n in names


@pytest.mark.skipif(not LIB.exists(), reason="libcbc.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(LIB))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"missing exports: {missing}"


@pytest.mark.skipif(not LIB.exists(), reason="libcbc.so not built")
def test_abi_version_and_python_binding_cover_header():
    from coupledbubble_control.backends.hip import _abi

    lib = _abi.lib()
    assert lib.cbc_abi_version() == _abi.CBC_ABI_VERSION
    assert set(declared_functions()) == set(_abi.SIGNATURES), "ctypes signatures out of sync with cbc.h"
    assert lib.cbc_last_error() in (b"",) or isinstance(lib.cbc_last_error(), bytes)


@pytest.mark.skipif(not LIB.exists(), reason="libcbc.so not built")
def test_plan_create_rejects_bad_specs_without_gpu_work():
    """Spec validation happens before any device call, so it runs on a CPU host."""
    from coupledbubble_control.backends.hip import _abi

    lib = _abi.lib()
    m = _abi.ModelSpecC(4, 13, 5, 3, 2, 1, 4, 8, _abi.CBC_FIELD_SW_N)
    e = _abi.ExecSpecC(16, 1, 2, 0, 3, _abi.CBC_MAP_AUTO, 0)   # tile 3 is not next_pow2(2)
    s = _abi.SolverSpecC(1e-9, 1e-9, 1e-16, 0.1, 1e-5, 1e-9, 1e-9, 2 / 3, 100, _abi.CBC_LIN_DIRECT, 0.1)
    h = ctypes.c_void_p()
    rc = lib.cbc_plan_create(ctypes.byref(m), ctypes.byref(e), ctypes.byref(s), ctypes.byref(h))
    assert rc == _abi.CBC_ERR_INVALID
    assert b"tile" in lib.cbc_last_error()
    e.tile = 2
    s.atol = 0.0
    rc = lib.cbc_plan_create(ctypes.byref(m), ctypes.byref(e), ctypes.byref(s), ctypes.byref(h))
    assert rc == _abi.CBC_ERR_INVALID
    assert lib.cbc_launch(None, None, 1, None) == _abi.CBC_ERR_INVALID


def test_plan_create_rejects_bad_event_modes_without_gpu_work():
    """An unknown event mode is invalid."""
    from coupledbubble_control.backends.hip import _abi

    lib = _abi.lib()
    m = _abi.ModelSpecC(4, 13, 5, 3, 2, 1, 4, 8, _abi.CBC_FIELD_SW_N)
    e = _abi.ExecSpecC(16, 1, 2, 0, 2, _abi.CBC_MAP_AUTO, 0)
    s = _abi.SolverSpecC(1e-9, 1e-9, 1e-16, 0.1, 1e-12, 1e-9, 1e-9, 2 / 3, 100, _abi.CBC_LIN_DIRECT, 0.1, 7)
    h = ctypes.c_void_p()
    assert lib.cbc_plan_create(ctypes.byref(m), ctypes.byref(e), ctypes.byref(s), ctypes.byref(h)) == \
        _abi.CBC_ERR_INVALID
    assert b"event mode" in lib.cbc_last_error()


@pytest.mark.skipif(not LIB.exists(), reason="libcbc.so not built")
def test_fused_kernel_lists_without_gpu_work():
    """The fused persistent kernels cover the bubble-count lists {8, 4, 3, 2} and {16, 2}
    (any subset of one list, in any order) at 1 or 2 harmonics and every field type;
    malformed lists are errors, not answers.  Host-side selection only."""
    from coupledbubble_control.backends.hip import _abi

    lib = _abi.lib()

    def sup(ups, nk=2, field=_abi.CBC_FIELD_SW_N):
        arr = (ctypes.c_int32 * max(1, len(ups)))(*ups)
        return lib.cbc_fused_supported(arr, len(ups), nk, field)

    for ups in ([8, 4, 3, 2], [2, 3, 4, 8], [2, 4], [8], [16, 2], [2, 16], [16], [3, 2]):
        for nk in (1, 2):
            for f in (_abi.CBC_FIELD_CONST, _abi.CBC_FIELD_SW_N, _abi.CBC_FIELD_SW_A):
                assert sup(ups, nk, f) == 1, (ups, nk, f)
    for ups in ([2, 5], [16, 4], [16, 8, 2], [6]):
        assert sup(ups) == 0, ups
    assert sup([2, 4], nk=3) == 0
    assert sup([2, 4], field=9) == 0
    assert sup([2, 2]) == _abi.CBC_ERR_INVALID
    assert b"two bins" in lib.cbc_last_error()
    assert sup([]) == _abi.CBC_ERR_INVALID
    assert sup([2, 3, 4, 8, 16]) == _abi.CBC_ERR_INVALID
    assert lib.cbc_launch_fused(None, None, None, None, None, 1, 0, 1, None) == _abi.CBC_ERR_INVALID
