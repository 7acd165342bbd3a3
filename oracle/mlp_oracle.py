"""Oracle (test infrastructure, not product code): float64 numpy restatement
of the surrogate MLP in librapt_hip.so (rapt_amd/csrc/rapt_mlp.hip), used only
by tests/ to check the device kernels.

Network 3 -> 256 -> 256 -> 256 -> 3 with SiLU; parameter layout as
RAPT_MLP_NPARAM in include/rapt_hip.h.  The parameters are kept in base-2
units, as the device keeps them: the hidden layers compute a = log2(e) z and
g = a / (1 + 2^-a) = log2(e) silu(z); to_silu_params converts them to the
plain SiLU network's (tests/test_oracle_mlp.py checks the two forwards
agree).  With quantize=True the forward and
backward passes round to bfloat16 (round to nearest even) exactly where the
device does: every weight, the activations H1 and H2 that feed the MFMA
layers, and the back-propagated dZ3 and dZ2 that feed the gradient GEMMs;
H3 and dZ1 stay unrounded.  With quantize=False the same code is the plain
float64 network, whose gradients tests/test_oracle_mlp.py pins by central
finite differences.  The rounded gradients are those of the straight-through
convention (rounding treated as identity in the backward pass), which is what
the device computes.

Loss: mean squared error over the batch and the three outputs,
L = sum (Y - T)^2 / (3 B).
"""
import numpy as np

H = 256
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * 3
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H * H
OFF_W4 = OFF_B3 + H
OFF_B4 = OFF_W4 + 3 * H
NPARAM = OFF_B4 + 3


def bf16(x):
    """Round to bfloat16 (nearest, ties to even); returns float64 values."""
    a = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)
    r = ((a >> 16) & 1) + np.uint32(0x7FFF)
    return ((a + r) & np.uint32(0xFFFF0000)).view(np.float32).astype(np.float64)


def unpack(theta):
    t = np.asarray(theta, dtype=np.float64)
    return dict(W1=t[OFF_W1:OFF_B1].reshape(H, 3), b1=t[OFF_B1:OFF_W2],
                W2=t[OFF_W2:OFF_B2].reshape(H, H), b2=t[OFF_B2:OFF_W3],
                W3=t[OFF_W3:OFF_B3].reshape(H, H), b3=t[OFF_B3:OFF_W4],
                W4=t[OFF_W4:OFF_B4].reshape(3, H), b4=t[OFF_B4:NPARAM])


def pack(p):
    return np.concatenate([p["W1"].ravel(), p["b1"], p["W2"].ravel(), p["b2"],
                           p["W3"].ravel(), p["b3"], p["W4"].ravel(), p["b4"]])


def init(seed, members=1):
    """He-normal hidden weights, output layer scaled by 0.1, zero biases."""
    out = np.zeros((members, NPARAM), np.float32)
    for k in range(members):
        rng = np.random.default_rng([seed, k])
        p = dict(W1=rng.normal(0, np.sqrt(2 / 3), (H, 3)), b1=np.zeros(H),
                 W2=rng.normal(0, np.sqrt(2 / H), (H, H)), b2=np.zeros(H),
                 W3=rng.normal(0, np.sqrt(2 / H), (H, H)), b3=np.zeros(H),
                 W4=0.1 * rng.normal(0, np.sqrt(1 / H), (3, H)), b4=np.zeros(3))
        out[k] = pack(p)
    return out


def _sig(a):
    """Logistic factor in base-2 units: sigma(a ln 2) = 1 / (1 + 2^-a)."""
    return 1.0 / (1.0 + np.exp2(-a))


def act(a):
    """Hidden activation in base-2 units, a sigma(a ln 2) = log2(e) silu(a ln 2)."""
    return a * _sig(a)


def dact(a):
    s = _sig(a)
    return s * (1 + np.log(2.0) * a * (1 - s))


def silu(z):
    """The plain SiLU, z sigma(z) (for the equivalence check)."""
    return z / (1.0 + np.exp(-z))


def to_silu_params(theta):
    """The stored (base-2) parameters as the plain SiLU network's:
    W1 = ln2 W1', b1 = ln2 b1', W2 = W2', b2 = ln2 b2', W3 = W3', b3 = ln2 b3',
    W4 = log2(e) W4', b4 = b4'."""
    p = unpack(theta)
    ln2 = np.log(2.0)
    return dict(W1=ln2 * p["W1"], b1=ln2 * p["b1"], W2=p["W2"], b2=ln2 * p["b2"],
                W3=p["W3"], b3=ln2 * p["b3"], W4=p["W4"] / ln2, b4=p["b4"])


def silu_forward(p, X):
    """The plain float64 SiLU network with parameters p (to_silu_params)."""
    H1 = silu(np.asarray(X, np.float64) @ p["W1"].T + p["b1"])
    H2 = silu(H1 @ p["W2"].T + p["b2"])
    H3 = silu(H2 @ p["W3"].T + p["b3"])
    return H3 @ p["W4"].T + p["b4"]


def forward(theta, X, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    p = {k: q(v) for k, v in unpack(theta).items()}
    X = np.asarray(X, np.float64)
    Z1 = X @ p["W1"].T + p["b1"]
    H1 = q(act(Z1))
    Z2 = H1 @ p["W2"].T + p["b2"]
    H2 = q(act(Z2))
    Z3 = H2 @ p["W3"].T + p["b3"]
    H3 = act(Z3)
    Y = H3 @ p["W4"].T + p["b4"]
    return Y, dict(p=p, X=X, Z1=Z1, H1=H1, Z2=Z2, H2=H2, Z3=Z3, H3=H3)


def loss_and_grad(theta, X, T, quantize=True):
    q = bf16 if quantize else (lambda a: np.asarray(a, np.float64))
    Y, c = forward(theta, X, quantize)
    p = c["p"]
    B = X.shape[0]
    D = Y - T
    L = np.sum(D * D) / (3 * B)
    dY = 2 * D / (3 * B)
    dH3 = dY @ p["W4"]
    dZ3 = q(dH3 * dact(c["Z3"]))
    dH2 = dZ3 @ p["W3"]
    dZ2 = q(dH2 * dact(c["Z2"]))
    dH1 = dZ2 @ p["W2"]
    dZ1 = dH1 * dact(c["Z1"])
    g = dict(W4=dY.T @ c["H3"], b4=dY.sum(0), W3=dZ3.T @ c["H2"], b3=dZ3.sum(0),
             W2=dZ2.T @ c["H1"], b2=dZ2.sum(0), W1=dZ1.T @ c["X"], b1=dZ1.sum(0))
    return L, pack(g)


def features(query):
    q = np.asarray(query, np.float64)
    return np.stack([(q[:, 0] - 5) / 3, (q[:, 1] - 49.5) / 39.5, (q[:, 2] - 5.5) / 1.5], axis=1)
