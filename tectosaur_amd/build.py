"""Build libtct_hip.so in-tree (hipcc for gfx950, g++ for the host C++).

    python -m tectosaur_amd.build

All device code is compiled for gfx950 only.  Packed-FP32 instruction
selection is disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``):
the FP32 kernels are written for one FP32 lane op per VALU slot.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libtct_hip.so')
HIP_SOURCES = ['abi.hip', 'far_p2p.hip', 'far_tree.hip', 'far_tree_eval_u.hip',
               'far_tree_eval_rt.hip', 'far_tree_eval_ra.hip', 'far_tree_eval_rh.hip', 'pairs.hip',
               'interior.hip', 'near_find.hip', 'krylov.hip', 'sparse.hip', 'calib.hip']
HOST_SOURCES = ['host_mesh.cpp', 'host_constraints.cpp']
ARCH = 'gfx950'


def _run(cmd):
    print('+', ' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _stale(obj, src):
    """True when obj is missing or older than its source or any header."""
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    deps = [src, os.path.join(HERE, '..', 'include', 'tct_hip.h')]
    deps += [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.h')]
    return any(os.path.getmtime(d) > t for d in deps)


def build(verbose=False, jobs=None, incremental=False):
    """Compile every source (``incremental``: only objects older than their
    source or a header) and link libtct_hip.so."""
    hipcc = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    bdir = os.path.join(HERE, '_build')
    os.makedirs(bdir, exist_ok=True)
    objs = []
    procs = []
    hip_flags = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH, '-mcode-object-version=5',
                 '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
                 '-Wno-unused-command-line-argument']
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, src.replace('.hip', '.o'))
        objs.append(obj)
        if incremental and not _stale(obj, os.path.join(CSRC, src)):
            continue
        cmd = [hipcc] + hip_flags + ['-c', os.path.join(CSRC, src), '-o', obj]
        print('+', ' '.join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    for src in HOST_SOURCES:
        obj = os.path.join(bdir, src.replace('.cpp', '.o'))
        objs.append(obj)
        if incremental and not _stale(obj, os.path.join(CSRC, src)):
            continue
        cmd = ['g++', '-O3', '-std=c++17', '-fPIC', '-fopenmp', '-ffp-contract=off', '-Wall',
               '-c', os.path.join(CSRC, src), '-o', obj]
        print('+', ' '.join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    bad = [p.args for p in procs if p.wait() != 0]
    if bad:
        raise RuntimeError('compile failed: %s' % bad)
    gomp = subprocess.run(['g++', '-print-file-name=libgomp.so'], check=True,
                          capture_output=True, text=True).stdout.strip()
    tmp = LIB + '.tmp'
    _run([hipcc, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs +
         [gomp, '-Wl,-rpath,' + os.path.dirname(gomp)])
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
