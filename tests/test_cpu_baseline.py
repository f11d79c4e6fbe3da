"""The CPU baseline of the farfield N-body (oracle/cpu_farfield.c, built for
this CPU) against the oracle's FP64 point N-body: the baseline that bench.py
times must compute the same sums (FP64 to 1e-12; FP32 with fast math to the
FP32 rounding of absolute coordinates)."""
import numpy as np


def test_cpu_baseline_h_matches_oracle(oracle):
    from oracle import cpu_baseline as CB
    rng = np.random.default_rng(41)
    n_obs, n_src = 150, 1200
    src = rng.uniform(-1, 1, size=(n_src, 3))
    sn = rng.normal(size=(n_src, 3))
    sn /= np.linalg.norm(sn, axis=1)[:, None]
    vec = rng.normal(size=(n_src, 3))
    obs = np.concatenate([rng.uniform(-1, 1, size=(n_obs - 10, 3)), src[:10]])  # 10 r2 == 0 pairs
    on = rng.normal(size=(n_obs, 3))
    on /= np.linalg.norm(on, axis=1)[:, None]
    params = [3e10, 0.25]
    want = oracle.farfield_pts('elasticH3', obs, on, src, sn, vec, params).reshape((-1, 3))
    eng = CB.HFarfieldCPU(CB.build(), src, sn, vec, params)
    got64 = eng.run(obs, on, np.float64)
    got32 = eng.run(obs, on, np.float32).astype(np.float64)
    rel = lambda a: np.linalg.norm(a - want) / np.linalg.norm(want)
    assert rel(got64) <= 1e-12, rel(got64)
    assert rel(got32) <= 1e-4, rel(got32)
    rate, n, dt = eng.rate(obs, on, np.float32, 0.01)
    assert rate > 0 and 0 < n <= n_obs
    name, ncpu = CB.cpu_model()
    assert name and ncpu >= 1
