"""The MLP oracle (oracle/mlp_oracle.py) pinned on the CPU.

- Unrounded gradients against central finite differences of the unrounded
  loss for a sample of parameters in every tensor (a wrong transpose, sign or
  missing SiLU' factor fails this).
- bfloat16 rounding against exactly representable values and ties.
- The rounded forward pass stays within bf16 precision of the unrounded one.
- The base-2 parametrisation is the SiLU network: act2 at exact points, and
  the unrounded forward against a plain SiLU network built from converted
  parameters (a factor of ln 2 dropped or misplaced anywhere fails this).
"""
import numpy as np

from oracle import mlp_oracle as O


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 1e-3], np.float32)
    got = O.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1); 1 + 3*2^-8 ties up to 1 + 2^-6
    assert got[0] == 1.0 and got[1] == 1.0 and got[2] == 1.0 + 2 ** -7
    assert got[3] == 1.0 + 2 ** -6 and got[4] == -2.5
    assert abs(got[5] - 1e-3) / 1e-3 < 2 ** -8


def test_unrounded_gradients_match_finite_differences():
    rng = np.random.default_rng(0)
    theta = O.init(3)[0].astype(np.float64)
    theta[O.OFF_B1:O.OFF_W2] = rng.normal(0, 0.1, O.H)
    theta[O.OFF_B4:] = rng.normal(0, 0.1, 3)
    X = rng.uniform(-1, 1, (16, 3))
    T = rng.normal(0, 1, (16, 3))
    _, g = O.loss_and_grad(theta, X, T, quantize=False)
    idx = []
    for lo, hi in ((O.OFF_W1, O.OFF_B1), (O.OFF_B1, O.OFF_W2), (O.OFF_W2, O.OFF_B2),
                   (O.OFF_B2, O.OFF_W3), (O.OFF_W3, O.OFF_B3), (O.OFF_B3, O.OFF_W4),
                   (O.OFF_W4, O.OFF_B4), (O.OFF_B4, O.NPARAM)):
        idx += list(rng.choice(np.arange(lo, hi), size=min(6, hi - lo), replace=False))
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (O.loss_and_grad(tp, X, T, False)[0] - O.loss_and_grad(tm, X, T, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_rounded_forward_close_to_unrounded():
    rng = np.random.default_rng(1)
    theta = O.init(5)[0]
    X = rng.uniform(-1, 1, (64, 3))
    y0, _ = O.forward(theta, X, quantize=False)
    y1, _ = O.forward(theta, X, quantize=True)
    assert np.max(np.abs(y1 - y0)) < 0.05 * np.max(np.abs(y0)) + 1e-3


def test_base2_activation_known_values():
    a = np.array([0.0, 1.0, -1.0, 2.0, -2.0])
    # a / (1 + 2^-a): 0, 2/3, -1/3, 8/5, -2/5 exactly
    np.testing.assert_allclose(O.act(a), [0.0, 2 / 3, -1 / 3, 8 / 5, -2 / 5], rtol=1e-15, atol=0)
    z = np.linspace(-8, 8, 33)
    np.testing.assert_allclose(O.act(z / np.log(2.0)), O.silu(z) / np.log(2.0), rtol=1e-14,
                               atol=1e-300)
    h = 1e-6
    fd = (O.act(z + h) - O.act(z - h)) / (2 * h)
    np.testing.assert_allclose(O.dact(z), fd, rtol=1e-8, atol=1e-10)


def test_base2_network_is_the_silu_network():
    rng = np.random.default_rng(4)
    theta = O.init(7)[0].astype(np.float64)
    for lo, hi in ((O.OFF_B1, O.OFF_W2), (O.OFF_B2, O.OFF_W3), (O.OFF_B3, O.OFF_W4),
                   (O.OFF_B4, O.NPARAM)):
        theta[lo:hi] = rng.normal(0, 0.3, hi - lo)
    X = rng.uniform(-1.5, 1.5, (128, 3))
    y2, _ = O.forward(theta, X, quantize=False)
    ys = O.silu_forward(O.to_silu_params(theta), X)
    assert np.max(np.abs(y2 - ys)) <= 1e-12 * np.max(np.abs(ys))
