"""The C-ABI library loads and exports every entry point include/tct_hip.h
declares (no device calls; runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, 'include', 'tct_hip.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(tct_[a-z0-9_]+)\s*\(', src)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ('tct_far_p2p', 'tct_pairs_integrate', 'tct_bsr9_spmv', 'tct_last_error'):
        assert s in syms


def test_library_exports_all_declared_symbols():
    from tectosaur_amd import _native as N
    L = N.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    assert L.tct_abi_version() == 1


def test_python_binding_covers_header():
    from tectosaur_amd import _native as N
    assert sorted(N.SIGNATURES) == declared_symbols()


def test_argument_errors_are_reported_without_a_device():
    # argument validation runs on the host before any launch
    from tectosaur_amd import _native as N
    L = N.lib()
    rc = L.tct_far_p2p(9, 0, 64, None, None, None, 64, None, None, None, 0.0, 0.0, None, None,
                       None)
    assert rc < 0
    assert b'kernel' in L.tct_last_error()
    # unknown launch-shape flag bits are refused before anything runs
    bad = N.TileCfg(256, 1, 1, 8)
    buf = ctypes.c_double(0.0)
    rc = L.tct_far_p2p(3, 0, 64, ctypes.byref(buf), ctypes.byref(buf), None, 64,
                       ctypes.byref(buf), ctypes.byref(buf), ctypes.byref(buf), 0.0, 0.0,
                       ctypes.byref(bad), ctypes.byref(buf), None)
    assert rc < 0 and b'flags' in L.tct_last_error()
    # a NaN extent or a non-finite eps2 is refused
    ok = N.TileCfg(256, 1, 1, 0)
    for eps2, ext in ((float('inf'), 0.0), (1e-12, float('nan'))):
        rc = L.tct_far_p2p(3, 0, 64, ctypes.byref(buf), ctypes.byref(buf), None, 64,
                           ctypes.byref(buf), ctypes.byref(buf), ctypes.byref(buf), eps2, ext,
                           ctypes.byref(ok), ctypes.byref(buf), None)
        assert rc < 0 and (b'eps2' in L.tct_last_error() or b'extent' in L.tct_last_error())
    # RH3 rule chunks: at most 11, a multiple of 4 up to 44 or of 8 up to 88,
    # and at most nq
    ip = ctypes.c_int32(0)
    for chunks, nq in ((14, 400), (52, 400), (96, 400), (13, 400), (16, 8)):
        rc = L.tct_pairs_integrate(6, nq, ctypes.byref(buf), ctypes.byref(buf), ctypes.byref(buf),
                                   ctypes.byref(ip), 1, ctypes.byref(ip), 2, 0,
                                   ctypes.byref(buf), -chunks, ctypes.byref(buf), None)
        assert rc < 0 and b'chunks' in L.tct_last_error(), chunks
    rc = L.tct_calib_valu(9, 1, 1, ctypes.byref(buf), None)
    assert rc < 0 and b'mode' in L.tct_last_error()
    n = ctypes.c_int64(-1)
    rc = L.tct_find_near_pairs(0, None, None, 0, None, None, 1.0, None, 0, ctypes.byref(n))
    assert rc == 0 and n.value == 0
