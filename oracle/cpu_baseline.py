"""CPU baseline of the farfield N-body -- BENCH / TEST INFRASTRUCTURE.

Loads libcpubase.so (cpu_farfield.c: the H-kernel point N-body of
tectosaur/farfield_direct.cl:15-113 for many-core CPUs, FP32 and FP64,
-O3 -march=native -ffast-math -fopenmp) and times it.  Only bench.py's
cpu_baseline leg and tests/ use this module.

``build()`` compiles the library for the CPU it runs on (-march=native)
into a private directory, so a prebuilt copy for another CPU is never
loaded by the baseline.
"""
import ctypes
import os
import platform
import subprocess
import tempfile
import time

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))


def cpu_model():
    """The CPU model name (from /proc/cpuinfo) and logical CPU count."""
    name = platform.processor() or 'unknown'
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    name = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return name, os.cpu_count()


def build(march='native'):
    """Compile cpu_farfield.c for this CPU; returns the loaded library."""
    out_dir = tempfile.mkdtemp(prefix='tct_cpubase_')
    out = os.path.join(out_dir, 'libcpubase.so')
    subprocess.run(['make', '-s', '-B', '-C', _HERE, 'libcpubase.so',
                    'CPUBASE_MARCH=%s' % march, 'CPUBASE_OUT=%s' % out], check=True)
    L = ctypes.CDLL(out)
    for nm, t in (('cpu_farfield_pts_H_f32', ctypes.c_float),
                  ('cpu_farfield_pts_H_f64', ctypes.c_double)):
        p = ctypes.POINTER(t)
        fn = getattr(L, nm)
        fn.argtypes = [ctypes.c_int64, p, p, ctypes.c_int64] + [p] * 9 + \
            [ctypes.POINTER(ctypes.c_double), p]
        fn.restype = None
    return L


class HFarfieldCPU:
    """Source set prepared once (structure of arrays) in FP32 and FP64."""

    def __init__(self, lib, src_pts, src_ns, vec, params):
        self.lib = lib
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        self.n_src = src_pts.shape[0]
        self.soa = {}
        for dt in (np.float32, np.float64):
            self.soa[dt] = [np.ascontiguousarray(a, dtype=dt) for a in
                            (src_pts[:, 0], src_pts[:, 1], src_pts[:, 2], src_ns[:, 0],
                             src_ns[:, 1], src_ns[:, 2], vec[:, 0], vec[:, 1], vec[:, 2])]

    def run(self, obs_pts, obs_ns, dtype):
        dt = np.dtype(dtype).type
        ct = ctypes.c_float if dt == np.float32 else ctypes.c_double
        p = lambda a: a.ctypes.data_as(ctypes.POINTER(ct))
        op = np.ascontiguousarray(obs_pts, dtype=dt)
        on = np.ascontiguousarray(obs_ns, dtype=dt)
        out = np.empty((op.shape[0], 3), dtype=dt)
        fn = self.lib.cpu_farfield_pts_H_f32 if dt == np.float32 else \
            self.lib.cpu_farfield_pts_H_f64
        fn(op.shape[0], p(op), p(on), self.n_src, *[p(a) for a in self.soa[dt]],
           self.params.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), p(out))
        return out

    def rate(self, obs_pts, obs_ns, dtype, seconds):
        """Run growing prefixes of the observation sample until one takes at
        least ``seconds``; returns (pair-evals/s, observation points, s)."""
        n = min(64, obs_pts.shape[0])
        while True:
            t0 = time.perf_counter()
            self.run(obs_pts[:n], obs_ns[:n], dtype)
            dt = time.perf_counter() - t0
            if dt >= seconds or n == obs_pts.shape[0]:
                return n * self.n_src / dt, n, dt
            n = min(obs_pts.shape[0], max(2 * n, int(n * 1.2 * seconds / max(dt, 1e-6))))
