"""Closed forms of the tanh current sheet B = (B0 tanh(z/lam), 0, Bn) used by
the switching tests (oracle and device): rho_c / L of Particle and
GuidingCenter rows, and the splitting of Adaptive rows into segments."""
import numpy as np

from rapt_amd import constants as K

C = float(K.c)


def sheet_ratio(row, mode, mu, mass, charge, B0=20e-9, lam=0.5 * K.Re, Bn=1e-9):
    """rho_c / L of a row in B = (B0 tanh(z/lam), 0, Bn) in closed form:
    L = |B| / max|dB_i/dx_j| = |B| / (B0/lam sech^2(z/lam)); rho_c as
    utils.cyclotron_radius (Particle rows) or cyclotron_radius2 with the
    GuidingCenter.cycrad speeds (guiding-centre rows)."""
    z = row[3]
    B = np.array([B0 * np.tanh(z / lam), 0.0, Bn])
    Bm = np.sqrt(B @ B)
    L = Bm / (B0 / lam / np.cosh(z / lam) ** 2)
    if mode == 0:
        p = np.asarray(row[4:7])
        gm = np.sqrt(mass ** 2 + p @ p / C ** 2)
        v = p / gm
        g = 1 / np.sqrt(1 - (v @ v) / C ** 2)
        vpar = v @ B / Bm
        vperp = np.sqrt(v @ v - vpar ** 2)
    else:
        pp = row[4]
        g = np.sqrt(1 + 2 * mu * Bm / (mass * C * C) + (pp / mass / C) ** 2)
        if g - 1 < 1e-6:
            vp = pp / mass
            sp = np.sqrt(2 * mu * Bm / mass + vp ** 2)
        else:
            vp = pp / mass / g
            sp = C * np.sqrt(1 - 1 / g ** 2)
        g = 1 / np.sqrt(1 - (sp / C) ** 2)
        vperp = np.sqrt((sp - vp) * (sp + vp))
    return g * mass * vperp / (abs(charge) * Bm) / L


def segments(rows):
    cut = np.r_[0, np.nonzero(np.diff(rows[:, 0]))[0] + 1, len(rows)]
    return [rows[a:b] for a, b in zip(cut[:-1], cut[1:])]
