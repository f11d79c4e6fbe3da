"""ctypes binding of libtct_hip.so (the C ABI declared in include/tct_hip.h).

torch is imported before the library is opened: torch ships its own HIP
runtime whose soname (libamdhip64.so.7) is the one libtct_hip.so links
against, so the dynamic loader binds both to the same runtime and torch's
device pointers and streams are valid here.

There is no fallback: if the library is missing, every call raises.
"""
import ctypes
import os

import torch  # noqa: F401  (must be loaded before the HIP library)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libtct_hip.so')

KERNEL_IDS = {
    'elasticU3': 0, 'elasticT3': 1, 'elasticA3': 2, 'elasticH3': 3,
    'elasticRT3': 4, 'elasticRA3': 5, 'elasticRH3': 6,
}
F32, F64 = 0, 1
GROUP = 64
REC = 12


class TctError(RuntimeError):
    pass


class FarSegment(ctypes.Structure):
    """tct_far_segment: one problem of a batched N-body launch."""
    _fields_ = [('n_obs_pad', ctypes.c_int64), ('loc32', ctypes.c_void_p),
                ('obs_org', ctypes.c_void_p), ('n_src_pad', ctypes.c_int64),
                ('rec', ctypes.c_void_p), ('src_org', ctypes.c_void_p),
                ('eps2', ctypes.c_double), ('group_extent', ctypes.c_double),
                ('src_chunks', ctypes.c_int32), ('flags', ctypes.c_int32),
                ('out_partial', ctypes.c_void_p)]


class TileCfg(ctypes.Structure):
    _fields_ = [('block', ctypes.c_int32), ('obs_per_lane', ctypes.c_int32),
                ('src_chunks', ctypes.c_int32), ('flags', ctypes.c_int32)]


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    'tct_last_error': [],
    'tct_abi_version': [],
    'tct_far_pack_obs': [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp],
    'tct_far_pack_src': [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp],
    'tct_far_pack_src_ex': [_i64, _vp, _vp, _i64, _int, _vp, _vp, _vp, _vp],
    'tct_far_gather_curl': [_int, _int, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    'tct_far_scatter_curl': [_int, _i64, _int, _vp, _vp, _vp, _vp, _i64, _int, _vp, _vp, _vp,
                             _vp],
    'tct_far_strength_pts': [_int, _int, _i64, _vp, _vp, _vp, _vp],
    'tct_far_gather_tris': [_int, _int, _i64, _int, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    'tct_far_p2p': [_int, _int, _i64, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _dbl, _dbl, _vp, _vp,
                    _vp],
    'tct_far_p2p_batched': [_int, _vp, _int, _int, _int, _vp, _vp],
    'tct_far_reduce_pts': [_i64, _i64, _int, _vp, _vp, _vp],
    'tct_tree_p2m': [_int, _int, _int, _i64, _i64, _vp, _vp, _vp, _vp, _vp],
    'tct_tree_m2m': [_int, _int, _int, _int, _int, _int, _int, _vp, _vp, _vp],
    'tct_tree_eval': [_int, _int, _int, _int, _vp, _vp, _vp, _vp, _vp, _vp, _int, _vp, _vp,
                      _vp, _dbl, _vp, _vp],
    'tct_tree_reduce': [_int, _int, _int, _vp, _vp, _vp, _vp],
    'tct_far_scatter_tris': [_i64, _int, _vp, _vp, _vp, _i64, _int, _vp, _vp, _vp],
    'tct_pairs_integrate': [_int, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _int, _int, _vp, _int,
                            _vp, _vp],
    'tct_sub_f64': [_i64, _vp, _vp, _vp, _vp],
    'tct_bsr9_spmv': [_i64, _vp, _vp, _vp, _vp, _vp, _int, _vp],
    'tct_interior_pairs': [_int, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _int, _int, _vp,
                           _int, _vp, _vp],
    'tct_interior_corners': [_i64, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp],
    'tct_bsr39_spmv': [_i64, _vp, _vp, _vp, _vp, _vp, _int, _vp],
    'tct_limit_combine': [_int, _i64, _int, _vp, _vp, _vp, _vp],
    'tct_krylov_scratch': [_int, _i64, _vp],
    'tct_cgs2': [_int, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    'tct_krylov_combine': [_int, _i64, _i64, _vp, _vp, _dbl, _dbl, _vp, _vp],
    'tct_cgs2_noalloc': [_int, _i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp],
    'tct_calib_valu': [_int, _int, _int, _vp, _vp],
    'tct_arnoldi_step': [_vp, _int, _i64, _i64, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    'tct_tick_gather': [_int, _vp, _vp, _int, _vp, _vp, _vp, _i64, _vp, _vp, _vp],
    'tct_interior_pairs_sinh': [_int, _int, _vp, _vp, _int, _vp, _vp, _vp, _vp, _vp, _vp, _i64,
                                _vp, _vp, _vp, _vp],
    'tct_near_grid_params': [_i64, _vp, _dbl, _vp, _vp, _vp],
    'tct_near_pairs_dev_scratch': [_i64, _i64, _vp, _vp],
    'tct_near_pairs_dev_count': [_i64, _vp, _vp, _i64, _vp, _vp, _dbl, _dbl, _vp, _dbl, _vp, _vp,
                                 ctypes.c_size_t, _vp, _vp],
    'tct_near_pairs_dev_fill': [_i64, _vp, _vp, _i64, _vp, _vp, _dbl, _dbl, _vp, _dbl, _vp, _vp,
                                ctypes.c_size_t, _vp, _vp, _vp],
    'tct_split_adjacent_dev_scratch': [_i64, _vp],
    'tct_split_adjacent_close_dev': [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     ctypes.c_size_t, _vp],
    'tct_bsr9_spmv_f32': [_i64, _vp, _vp, _vp, _vp, _vp, _int, _vp],
    'tct_csr_spmv': [_i64, _vp, _vp, _vp, _vp, _vp, _int, _vp],
    'tct_find_near_pairs': [_i64, _vp, _vp, _i64, _vp, _vp, _dbl, _vp, _i64, _vp],
    'tct_set_host_threads': [_int],
    'tct_split_adjacent_close': [_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    'tct_build_constraint_matrix': [_i64, _vp, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _i64,
                                    _vp, _vp, _vp, _vp, _vp],
}
_RESTYPE = {'tct_last_error': ctypes.c_char_p}

_LIB = None


def lib():
    """The loaded library; raises TctError if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise TctError('libtct_hip.so not built (run python -m tectosaur_amd.build)')
        L = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _RESTYPE.get(name, ctypes.c_int)
        _LIB = L
    return _LIB


def call(name, *args):
    """Call an entry point and raise TctError with tct_last_error() on failure."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != 0:
        msg = L.tct_last_error().decode(errors='replace')
        raise TctError('%s failed (%d): %s' % (name, rc, msg))
    return rc


def ptr(t):
    """Device (or host) address of a tensor / numpy array, or None."""
    if t is None:
        return None
    if isinstance(t, torch.Tensor):
        return t.data_ptr()
    return t.ctypes.data


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream


def require_device():
    """The HIP path only: raise when no GPU is visible."""
    if not torch.cuda.is_available():
        raise TctError('no HIP device visible: tectosaur_amd runs on the GPU only')
    lib()
    return torch.device('cuda', torch.cuda.current_device())


def upload(a, dtype=None, device=None):
    """Host array -> new device tensor without waiting for the device.

    The data is copied into pinned host memory (a new pinned tensor from
    torch's caching host allocator) and from there to the device
    asynchronously on the current stream; the pinned block is reused only
    after that copy has completed.  A pageable ``torch.as_tensor(a,
    device=...)`` instead synchronises the stream and spins the calling
    thread until all work queued before it is done.  Device tensors are
    returned as they are (converted to ``dtype`` if given)."""
    import numpy as np
    if isinstance(a, torch.Tensor) and a.device.type != 'cpu':
        return a if dtype is None or a.dtype == dtype else a.to(dtype)
    t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None and t.dtype != dtype:
        t = t.to(dtype)
    if device is None:
        device = torch.device('cuda', torch.cuda.current_device())
    if t.numel() == 0:
        return torch.empty(t.shape, dtype=t.dtype, device=device)
    p = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    p.copy_(t)
    return p.to(device, non_blocking=True)


def params_array(params):
    import numpy as np
    return np.ascontiguousarray(np.array(params, dtype=np.float64))
