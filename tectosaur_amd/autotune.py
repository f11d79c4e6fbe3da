"""Launch-shape autotuner for the farfield N-body kernel.

The tunable shape is ``TileCfg(block, obs_per_lane, src_chunks, flags)``
with flags 0 or 2 (the XCD chunk order, src_chunks a multiple of 8).  A tuning
run times every candidate on the device for one (kernel, precision, size
class) and records the fastest in a JSON table (``autotune_table.json``
next to this file, or the path in TCT_AUTOTUNE_TABLE).  ``lookup`` reads the
table and falls back to a heuristic shape when no entry matches.

Size classes are powers of two of the padded observation and source counts,
so one entry serves a range of problem sizes.
"""
import json
import math
import os
import threading
import time

import torch

from . import _native as N

_LOCK = threading.Lock()
_TABLE = None
DEFAULT_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'autotune_table.json')


def table_path():
    return os.environ.get('TCT_AUTOTUNE_TABLE', DEFAULT_PATH)


def _size_class(n):
    return int(math.ceil(math.log2(max(n, 1))))


_ARCH = {}
_ARCH_LOCK = threading.Lock()   # its own lock: _key is called with _LOCK held


def device_arch(device):
    """gfx architecture name of a device, looked up once per device under a
    lock (torch's device-property query is not safe to run for the first
    time from several threads at once)."""
    idx = torch.device(device).index
    if idx is None:
        idx = torch.cuda.current_device()
    with _ARCH_LOCK:
        arch = _ARCH.get(idx)
        if arch is None:
            arch = torch.cuda.get_device_properties(idx).gcnArchName.split(':')[0]
            _ARCH[idx] = arch
    return arch


def _key(K_name, prec, n_obs_pad, n_src_pad, device):
    arch = 'gfx950'
    if device is not None and torch.cuda.is_available():
        arch = device_arch(device)
    return '%s|%s|%s|o%d|s%d' % (arch, K_name, prec, _size_class(n_obs_pad),
                                 _size_class(n_src_pad))


def _load():
    global _TABLE
    with _LOCK:
        if _TABLE is None:
            try:
                with open(table_path()) as f:
                    _TABLE = json.load(f)
            except (OSError, ValueError):
                _TABLE = {}
        return _TABLE


def lookup(K_name, prec, n_obs_pad, n_src_pad, device=None):
    from .farfield import default_tile_cfg, GROUP
    t = _load()
    e = t.get(_key(K_name, prec, n_obs_pad, n_src_pad, device))
    if e is not None:
        n_tiles = n_src_pad // GROUP
        chunks = max(1, min(int(e['src_chunks']), n_tiles))
        flags = 2 if (e.get('xcd') and chunks % 8 == 0) else 0
        if e.get('lds') and prec == 'f32':
            flags |= 4
        return N.TileCfg(int(e['block']), int(e['obs_per_lane']), chunks, flags)
    return default_tile_cfg(n_obs_pad, n_src_pad, prec, device)


def candidates(n_obs_pad, n_src_pad, n_cu=256, lds=False):
    """Launch shapes to time.  ``lds`` adds every shape again with the
    source tiles staged in LDS (flags bit 2; FP32 only)."""
    from .farfield import GROUP
    n_tiles = n_src_pad // GROUP
    out = []
    for block in (64, 128, 256):
        for per_lane in (1, 2, 4):
            obs_blocks = -(-n_obs_pad // (block * per_lane))
            for waves_target in (4, 8, 16, 32):
                chunks = max(1, min(n_tiles, -(-waves_target * n_cu // obs_blocks)))
                cfg = (block, per_lane, chunks, 0)
                if cfg not in out:
                    out.append(cfg)
            # the XCD chunk order: 8, 16, 24 or 32 chunks
            for chunks in (8, 16, 24, 32):
                if chunks <= n_tiles:
                    out.append((block, per_lane, chunks, 2))
    if lds:
        out += [(b, p, c, f | 4) for b, p, c, f in out]
    return out


def _run_and_reduce(engine, prec, cfg):
    """The N-body launch and what a matvec then pays for its source chunks:
    one read of every chunk's partial sums (each chunk writes n_out x
    n_obs_pad doubles, so many chunks cost traffic at small sizes)."""
    out = engine.run(prec, cfg)
    if cfg.src_chunks > 1:
        out[:cfg.src_chunks * engine.n_out * engine.n_obs_pad].view(cfg.src_chunks, -1).sum(0)
    return out


def tune(engine, prec, reps=3, save=True, verbose=False):
    """Time every candidate launch shape on ``engine`` (a FarfieldEngine with
    strengths already set), the N-body and the reduction of its chunk
    partials together, and store the fastest.  Returns (cfg, seconds)."""
    # candidates come from the fixed CU count of default_tile_cfg (MI355X:
    # 256), not the device's: the chunk count sets the summation order, so
    # a table tuned on any box gives the same results on every box
    best = None
    for block, per_lane, chunks, flags in candidates(engine.n_obs_pad, engine.n_src_pad,
                                                     lds=(prec == 'f32')):
        cfg = N.TileCfg(block, per_lane, chunks, flags)
        _run_and_reduce(engine, prec, cfg)  # warm-up / allocation
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            _run_and_reduce(engine, prec, cfg)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        if verbose:
            print('tune %s %s block=%d P=%d chunks=%d flags=%d: %.3f ms' % (
                engine.K.name, prec, block, per_lane, chunks, flags, dt * 1e3), flush=True)
        if best is None or dt < best[1]:
            best = (cfg, dt)
    engine.tile_cfg[prec] = best[0]
    if save:
        t = _load()
        key = _key(engine.K.name, prec, engine.n_obs_pad, engine.n_src_pad, engine.device)
        with _LOCK:
            t[key] = {
                'block': best[0].block, 'obs_per_lane': best[0].obs_per_lane,
                'src_chunks': best[0].src_chunks, 'xcd': bool(best[0].flags & 2),
                'lds': bool(best[0].flags & 4), 'seconds': best[1]}
            try:
                with open(table_path(), 'w') as f:
                    json.dump(t, f, indent=1, sort_keys=True)
            except OSError:
                pass
    return best
