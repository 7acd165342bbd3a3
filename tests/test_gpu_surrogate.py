"""Surrogate MLP kernels (rapt_amd/csrc/rapt_mlp.hip) on the GPU.

- Gradients of one training step against the float64 oracle with the same
  bfloat16 rounding points (oracle/mlp_oracle.py, itself pinned by finite
  differences): relative error of every parameter tensor below 1e-2 (float32
  accumulation and rare bf16 rounding flips against float64).
- The fused inference kernel against the training-path forward (GEMM
  kernels) on the same weights.
- End to end: traced labels -> training -> inference with fallback; the
  trained ensemble must beat its own untrained start by a wide margin, and
  the fallback bounds the error of flagged queries by the traced labels.
"""
import time

import numpy as np
import pytest

import rapt_amd as rapt
from rapt_amd import constants as K
from rapt_amd import fields as F
from rapt_amd import surrogate as S

pytestmark = pytest.mark.gpu

TENSORS = [("W1", 0, 768), ("b1", 768, 1024), ("W2", 1024, 66560), ("b2", 66560, 66816),
           ("W3", 66816, 132352), ("b3", 132352, 132608), ("W4", 132608, 133376),
           ("b4", 133376, 133379)]


def _queries(n, seed):
    rng = np.random.default_rng(seed)
    return np.c_[rng.uniform(2, 8, n), rng.uniform(10, 89, n), rng.uniform(4, 7, n)]


def test_mlp_gradients_match_oracle():
    from oracle import mlp_oracle as O
    rng = np.random.default_rng(11)
    th = S.init_params(3, 2)
    th[:, 768:1024] = rng.normal(0, 0.1, (2, 256))
    n = 256
    X = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    T = rng.normal(0, 1, (n, 3)).astype(np.float32)
    s = S.Surrogate(F.EarthDipole(), members=2, theta0=th)
    try:
        loss, g = s.train_steps(X, T, np.arange(n), n, 0.0, 0.0, want_grad=True)
    finally:
        s.close()
    for k in range(2):
        L, gr = O.loss_and_grad(th[k], X.astype(np.float64), T.astype(np.float64), True)
        assert abs(loss[k] - L) <= 1e-4 * L
        for name, lo, hi in TENSORS:
            e = np.linalg.norm(g[k, lo:hi] - gr[lo:hi]) / np.linalg.norm(gr[lo:hi])
            print("member %d %s: rel err %.2e" % (k, name, e))
            assert e < 1e-3, (k, name, e)


def test_fused_inference_matches_training_forward():
    rng = np.random.default_rng(12)
    th = S.init_params(4, 3)
    th[:, 768:1024] = rng.normal(0, 0.1, (3, 256))
    s = S.Surrogate(F.EarthDipole(), members=3, theta0=th, split=False)
    try:
        q = _queries(1000, 13)
        from oracle import mlp_oracle as O
        X = O.features(q).astype(np.float32)
        Y = s.forward(X)
        _, flag, raw, _ = s.infer(q, raw=True)
    finally:
        s.close()
    d = np.abs(raw - Y)
    sc = np.abs(Y).max()
    print("fused vs GEMM path: median %.3e, max %.3e, scale %.3e" % (np.median(d), d.max(), sc))
    # The two paths accumulate in FP32 in different orders (and layer 1 runs
    # on MFMA with a hi/lo split in the fused kernel), so a bf16 rounding of a
    # hidden activation can land one ulp (2^-8) apart: typical differences
    # stay at FP32 level, the largest at a few such flips.
    assert np.median(d) <= 1e-4 * sc
    assert d.max() <= 5e-3 * sc


def test_regime_split_features_and_flags():
    """The drift label's two regimes (the reference's gradI falls back to a
    one-sided difference where the inner line L Re - eyegradientstep has no
    mirror point, rapt/flutils.py:196-227): the pitch-angle feature moves by
    2.5 exactly for queries whose inner equatorial |B| reaches the mirror
    field (host double-precision restatement of the test), the fused
    inference uses the same features as the training path, and queries within
    1e-5 of the curve are flagged for the traced fallback."""
    import rapt_amd as rapt
    f = F.DipolePlusUniform(Bu=(0.0, 0.0, -10e-9))
    rng = np.random.default_rng(21)
    n = 4000
    q = np.c_[rng.uniform(2, 8, n), rng.uniform(60, 89, n), rng.uniform(4, 7, n)]
    # a few queries on the curve itself: alpha from B(L Re - d) = B(L Re) / sin^2 alpha
    d = rapt.params["eyegradientstep"]
    for i in range(8):
        L = 3.0 + 0.5 * i
        Bq, Bin = f.magB([0, L * K.Re, 0, 0]), f.magB([0, L * K.Re - d, 0, 0])
        q[i] = [L, np.degrees(np.arcsin(np.sqrt(Bq / Bin))), 5.0]
    q32 = q.astype(np.float32)
    Bq = np.array([f.magB([0, float(L) * K.Re, 0, 0]) for L in q32[:, 0].astype(np.float64)])
    Bin = np.array([f.magB([0, float(L) * K.Re - d, 0, 0])
                    for L in q32[:, 0].astype(np.float64)])
    sa = np.sin(np.radians(q32[:, 1].astype(np.float64)))
    Bm = Bq / sa ** 2
    side = (Bin > Bm) | (np.abs(Bin - Bm) / Bm < 1e-12)
    near = np.abs(Bin - Bm) < 1e-5 * Bm
    assert 0 < side.mean() < 1 and near[:8].all()
    s = S.Surrogate(f, members=2, seed=3)
    try:
        X, _ = s.targets(q32, np.ones((n, 3)))
        base = (q32[:, 1] - 49.5) * np.float32(1 / 39.5)
        far = ~near
        np.testing.assert_array_equal(X[far, 1] - base[far] > 1.0, side[far])
        Y = s.forward(X)
        _, flag, raw, _ = s.infer(q32, thresh=1e30, raw=True)
    finally:
        s.close()
    assert np.all(flag[near] == 1) and np.all(flag[far] == 0)
    d = np.abs(raw - Y)
    assert np.median(d) <= 1e-4 * np.abs(Y).max() and d.max() <= 5e-3 * np.abs(Y).max()


def test_surrogate_training_and_fallback():
    f = F.DipolePlusUniform(Bu=(0.0, 0.0, -10e-9))
    qtr = _queries(1 << 14, 1)
    t0 = time.perf_counter()
    lab, st = S.traced_labels(f, qtr)
    t_trace = time.perf_counter() - t0
    assert np.mean(st == 1) > 0.99
    s = S.Surrogate(f, members=4, seed=0)
    try:
        qte = _queries(4096, 2)
        truth, st2 = S.traced_labels(f, qte)
        good = st2 == 1
        v0, _, _, _ = s.infer(qte)  # untrained (default normalisation)
        hist = s.fit(qtr[st == 1], lab[st == 1], epochs=30, batch=1024)
        vals, flag, _, ms = s.infer(qte, thresh=1e-3)
        err = np.abs(vals - truth) / np.abs(truth)
        err0 = np.abs(v0 - truth) / np.abs(truth)
        med = np.median(err[good], axis=0)
        print("loss first/last", hist[0], hist[-1])
        print("median rel err per output", med, "untrained", np.median(err0[good], axis=0))
        print("flagged %.3f; err<1e-3 among unflagged %.3f" % (
            flag.mean(), np.mean(np.all(err[good & (flag == 0)] < 1e-3, axis=1))))
        print("trace %.3f s for %d, infer %.3f ms for %d" % (t_trace, len(qtr), ms, len(qte)))
        assert np.all(hist[-1] < 0.05 * hist[0])
        assert np.all(med < 0.1 * np.median(err0[good], axis=0))
        pv, pflag, info = s.predict(qte, thresh=1e-3)
        fl = (pflag == 1) & good
        assert np.allclose(pv[fl], truth[fl], rtol=1e-12)
        # what the network gains over tracing: the same 65,536 queries by
        # both methods (inference kernel time against the traced labels)
        qg = _queries(1 << 16, 3)
        s.infer(qg[:4096])
        _, _, _, ms_inf = s.infer(qg)
        t0 = time.perf_counter()
        S.traced_labels(f, qg)
        t_tr = time.perf_counter() - t0
        print("65,536 queries: surrogate %.3f ms, traced %.1f ms, gain %.0fx" % (
            ms_inf, t_tr * 1e3, t_tr * 1e3 / ms_inf))
        # measured: 0.115 ms against 67 ms (582x); the bar leaves half of it
        # for run-to-run variation of the traced timing
        assert t_tr * 1e3 / ms_inf > 300
    finally:
        s.close()
