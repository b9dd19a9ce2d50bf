"""The C-ABI library loads on a machine without a GPU and exports exactly
what include/vipe_hip.h declares; the ctypes binding covers every symbol."""
import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "vipe_hip.h"


def declared():
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(vh_[a-z0-9_A-Z]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_path():
    names = declared()
    for must in ["vh_ba_run", "vh_ba_plan_create", "vh_corr_index_forward", "vh_corr_index_backward",
                 "vh_corr_build", "vh_altcorr_forward", "vh_altcorr_backward", "vh_altcorr_index_forward",
                 "vh_lie_expm", "vh_lie_act4_backward", "vh_lie_Jinv"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from vipe_amd import _lib

    if not _lib.LIB_PATH.exists():
        pytest.fail(f"{_lib.LIB_PATH} is not built (python -m vipe_amd.build)")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared()) == set(_lib.exported_symbols())
    loaded = _lib.load()
    assert loaded.vh_abi_version() == 1


def test_errors_are_reported_without_a_device():
    from vipe_amd import _lib

    lib = _lib.load()
    rc = lib.vh_corr_index_forward(99, 1, 1, 1, 1, 1, None, None, 3, None, None)
    assert rc == _lib.VH_ERR_UNSUPPORTED
    assert "dtype" in _lib.last_error()
    rc = lib.vh_lie_expm(9, _lib.VH_F32, 1, None, None, _lib.VH_HOST, None)
    assert rc == _lib.VH_ERR_UNDEFINED


def test_update_op_argument_struct_matches_the_header():
    from vipe_amd import _lib

    lib = _lib.load()
    assert lib.vh_uo_conv_args_size() == ctypes.sizeof(_lib.UoConvArgs)
    a = _lib.UoConvArgs()
    a.nseg = 0
    assert lib.vh_uo_conv(ctypes.byref(a), None) == _lib.VH_ERR_ARG
    assert "segments" in _lib.last_error()
    a.nseg, a.seg_c[0], a.seg_stride[0], a.KH, a.KW = 1, 48, 64, 3, 3
    a.seg[0] = 16
    a.Cout, a.Cout_pad = 128, 128
    assert lib.vh_uo_conv(ctypes.byref(a), None) == _lib.VH_ERR_ARG
    assert "segment 0" in _lib.last_error()
    assert lib.vh_uo_conv_small(None, 128, 0, 128, None, None, 1, 4, 4, 3, 3, 3, 0, None, 3, None) == _lib.VH_ERR_ARG
