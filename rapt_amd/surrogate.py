"""Learned bounce-drift surrogate with a traced fallback.

An ensemble of MLPs (3 -> 256 -> 256 -> 256 -> 3, SiLU, BF16 weights, FP32
accumulation; include/rapt_hip.h, rapt_mlp_*) maps a query
(L [Re], equatorial pitch angle [deg], log10 K [eV]) to
(bounce period [s], bounce-averaged drift angular velocity [rad/s], mirror
latitude [deg]).  Labels come from the traced bounce quantities on the GPU
(rapt_bounce_labels: BounceCenter's drift, rapt/BounceCenter.py:235-244, and
flutils.bounceperiod, rapt/flutils.py:232-252).  At inference, queries on
which the members' relative spread exceeds a threshold are traced instead,
so the traced method is both the fallback and the yardstick of the surrogate.
Parameters are held in base-2 units (hidden units compute a / (1 + 2^-a) =
log2(e) silu(z)); the plain SiLU network's are
(ln2 W1, ln2 b1, W2, ln2 b2, W3, ln2 b3, log2(e) W4, b4).

Training and inference run on the device (rapt_amd/csrc/rapt_mlp.hip); the
host only shuffles sample indices and holds the arrays.
"""
import ctypes as C

import numpy as np

from . import _native as N
from . import params as _params
from .constants import e as _e, m_pr as _m_pr

NPARAM = 133379


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def init_params(seed, members):
    """He-normal hidden weights, output layer x0.1, zero biases; [members][NPARAM]."""
    H = 256
    out = np.zeros((members, NPARAM), np.float32)
    for k in range(members):
        rng = np.random.default_rng([seed, k])
        parts = [rng.normal(0, np.sqrt(2 / 3), (H, 3)).ravel(), np.zeros(H),
                 rng.normal(0, np.sqrt(2 / H), (H, H)).ravel(), np.zeros(H),
                 rng.normal(0, np.sqrt(2 / H), (H, H)).ravel(), np.zeros(H),
                 0.1 * rng.normal(0, np.sqrt(1 / H), (3, H)).ravel(), np.zeros(3)]
        out[k] = np.concatenate(parts)
    return out


def traced_labels(field, query, mass=_m_pr, charge=_e, device=0, params=None):
    """(T_b, omega_d, mirror latitude) of each query by field-line tracing on
    the GPU; returns (labels (n, 3), status (n,))."""
    P = params or _params
    return N.context(device).bounce_labels(field.descriptor(), np.asarray(query, np.float64),
                                           mass, charge, P["eyegradientstep"],
                                           P["fieldlineresolution"])


class Surrogate:
    """Ensemble surrogate on one device.  theta0: initial parameters
    [members][NPARAM] (init_params by default)."""

    def __init__(self, field, mass=_m_pr, charge=_e, members=4, seed=0, theta0=None,
                 device=0, split=True, params=None):
        from .constants import e as eV
        self.field = field
        self.mass, self.charge = float(mass), float(charge)
        self.members = members
        self.ctx = N.context(device)
        self.ctx.set_field(field.descriptor())
        th = init_params(seed, members) if theta0 is None else np.ascontiguousarray(theta0,
                                                                                    np.float32)
        h = C.c_void_p()
        self.ctx.check(N.lib().rapt_mlp_create(self.ctx.h, members, _fp(th), self.mass, eV,
                                               C.byref(h)), "rapt_mlp_create")
        self.h = h
        self.norm = np.array([0, 0, 0, 1, 1, 1], np.float32)
        # the two regimes of the drift label (rapt_mlp_set_split): on for the
        # dipole-type fields whose labels the device traces
        from .constants import Re
        P = params or _params
        self.split = bool(split) and field.descriptor()["model"] in (1, 2, 5, 8)
        self.ctx.check(N.lib().rapt_mlp_set_split(
            self.h, float(P["eyegradientstep"]) if self.split else 0.0, float(Re)),
            "rapt_mlp_set_split")

    def close(self):
        if self.h:
            N.lib().rapt_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_norm(self, shift, scale):
        self.norm = np.r_[shift, scale].astype(np.float32)
        self.ctx.check(N.lib().rapt_mlp_set_norm(self.h, _fp(self.norm)), "rapt_mlp_set_norm")

    def targets(self, query, labels, normalised=True):
        q = np.ascontiguousarray(query, np.float32)
        lab = np.ascontiguousarray(labels, np.float64)
        n = q.shape[0]
        X = np.zeros((n, 3), np.float32)
        T = np.zeros((n, 3), np.float32)
        self.ctx.check(N.lib().rapt_mlp_targets(self.h, n, _fp(q), N._dp(lab),
                                                1 if normalised else 0, _fp(X), _fp(T)),
                       "rapt_mlp_targets")
        return X, T

    def train_steps(self, X, T, order, batch, lr0, lr1, want_grad=False):
        """Adam steps over the rows order[k*batch:(k+1)*batch]; returns
        (loss per member of the last step, gradients of the last step or None)."""
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        order = np.ascontiguousarray(order, np.int64)
        steps = order.size // batch
        loss = np.zeros(self.members)
        grad = np.zeros((self.members, NPARAM), np.float32) if want_grad else None
        self.ctx.check(N.lib().rapt_mlp_train(
            self.h, X.shape[0], _fp(X), _fp(T), batch, steps, N._ip64(order), lr0, lr1,
            N._dp(loss), _fp(grad) if want_grad else None), "rapt_mlp_train")
        return loss, grad

    def fit(self, query, labels, epochs=40, batch=4096, lr0=2e-3, lr1=1e-5, seed=0,
            log=None):
        """Train all members on (query, traced labels); returns the loss history."""
        lab = np.asarray(labels, np.float64)
        ok = np.all(np.isfinite(lab), axis=1) & np.all(lab[:, [0, 2]] > 0, axis=1) & \
            (lab[:, 1] != 0)
        q = np.asarray(query, np.float32)[ok]
        lab = lab[ok]
        _, R = self.targets(q, lab, normalised=False)
        self.set_norm(R.mean(0), np.maximum(R.std(0), 1e-3))
        X, T = self.targets(q, lab, normalised=True)
        n = X.shape[0]
        per = n // batch
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            # cosine schedule across epochs, constant within an epoch
            f0, f1 = ep / epochs, (ep + 1) / epochs
            a = lr1 + 0.5 * (lr0 - lr1) * (1 + np.cos(np.pi * f0))
            b = lr1 + 0.5 * (lr0 - lr1) * (1 + np.cos(np.pi * f1))
            order = rng.permutation(n)[:per * batch]
            loss, _ = self.train_steps(X, T, order, batch, a, b)
            hist.append(loss.copy())
            if log:
                log("epoch %d: loss %s" % (ep, np.array2string(loss, precision=3)))
        return np.array(hist)

    def params(self):
        th = np.zeros((self.members, NPARAM), np.float32)
        self.ctx.check(N.lib().rapt_mlp_get_params(self.h, _fp(th)), "rapt_mlp_get_params")
        return th

    def forward(self, X):
        """Training-path forward: normalised outputs [members][n][3]."""
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((self.members, X.shape[0], 3), np.float32)
        self.ctx.check(N.lib().rapt_mlp_forward(self.h, X.shape[0], _fp(X), _fp(Y)),
                       "rapt_mlp_forward")
        return Y

    def infer(self, query, thresh=1e-3, raw=False):
        """Fused ensemble inference.  Returns (values (n, 3) with the drift
        angular velocity signed as the species drifts, flag (n,), raw member
        outputs or None, kernel ms)."""
        q = np.ascontiguousarray(query, np.float32)
        n = q.shape[0]
        out = np.zeros((n, 3), np.float32)
        flag = np.zeros(n, np.int32)
        rw = np.zeros((self.members, n, 3), np.float32) if raw else None
        ms = C.c_double(0.0)
        self.ctx.check(N.lib().rapt_mlp_infer(self.h, n, _fp(q), thresh, _fp(out),
                                              N._ip32(flag), _fp(rw) if raw else None,
                                              C.byref(ms)), "rapt_mlp_infer")
        out = out.astype(np.float64)
        out[:, 1] *= -np.sign(self.charge)   # positive charges drift westward
        return out, flag, rw, ms.value

    def predict(self, query, thresh=1e-3, params=None):
        """Surrogate values, with flagged queries replaced by traced labels.
        Returns (values (n, 3), flag (n,), info)."""
        vals, flag, _, ms = self.infer(query, thresh)
        idx = np.nonzero(flag)[0]
        if idx.size:
            lab, st = traced_labels(self.field, np.asarray(query, np.float64)[idx], self.mass,
                                    self.charge, params=params)
            vals[idx] = lab
        return vals, flag, dict(kernel_ms=ms, fallback=int(idx.size))
