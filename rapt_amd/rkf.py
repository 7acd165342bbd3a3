"""Runge-Kutta-Fehlberg 4(5) with the interface of rapt/rkf.py:13-143.

Host-side solver used by the field-line tracer (rapt_amd.fieldline).  The
device kernels carry the same method (rapt_amd/csrc/rapt_steppers.hpp).
"""
import numpy as np

# Fehlberg tableau as rapt/rkf.py:48-87 writes it (16-digit decimals)
_A = (2.500000000000000e-01, 3.750000000000000e-01, 9.230769230769231e-01,
      1.000000000000000e+00, 5.000000000000000e-01)
_B = ((2.500000000000000e-01,),
      (9.375000000000000e-02, 2.812500000000000e-01),
      (8.793809740555303e-01, -3.277196176604461e+00, 3.320892125625853e+00),
      (2.032407407407407e+00, -8.000000000000000e+00, 7.173489278752436e+00,
       -2.058966861598441e-01),
      (-2.962962962962963e-01, 2.000000000000000e+00, -1.381676413255361e+00,
       4.529727095516569e-01, -2.750000000000000e-01))
_R = (2.777777777777778e-03, 0.0, -2.994152046783626e-02, -2.919989367357789e-02,
      2.000000000000000e-02, 3.636363636363636e-02)
_C = (1.157407407407407e-01, 0.0, 5.489278752436647e-01, 5.353313840155945e-01,
      -2.000000000000000e-01)


def rkf(f, a, b, x0, tol, hmax, hmin):
    """Solve x' = f(x, t), x(a) = x0 on [a, b]; returns (T, X) of accepted points.

    Steps start at hmax; a step is accepted when max|err|/h <= tol; the step
    factor is min(max(0.84 (tol/r)^(1/4), 0.1), 4); RuntimeError below hmin.
    """
    t = a
    x = np.array(x0, dtype=float)
    h = hmax
    T = [t]
    X = [x]
    while t < b:
        if t + h > b:
            h = b - t
        k = []
        for s in range(6):
            xs = x.copy()
            for j, bj in enumerate(_B[s - 1] if s else ()):
                xs = xs + bj * k[j]
            k.append(h * f(xs, t + (_A[s - 1] * h if s else 0.0)))
        r = abs(_R[0] * k[0] + _R[2] * k[2] + _R[3] * k[3] + _R[4] * k[4] + _R[5] * k[5]) / h
        r = np.max(r) if np.ndim(r) else r
        if r <= tol:
            t = t + h
            x = x + _C[0] * k[0] + _C[2] * k[2] + _C[3] * k[3] + _C[4] * k[4]
            T.append(t)
            X.append(x)
        with np.errstate(divide="ignore"):
            h = h * min(max(0.84 * (tol / r) ** 0.25, 0.1), 4.0)
        if h > hmax:
            h = hmax
        elif h < hmin:
            raise RuntimeError("Error: Could not converge to the required tolerance %e with "
                               "minimum stepsize  %e." % (tol, hmin))
    return np.array(T), np.array(X)
