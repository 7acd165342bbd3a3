"""Batched tracing on the GPU: many tracers, one launch.

The reference advances one Python object at a time (rapt/Particle.py:230-309,
rapt/GuidingCenter.py:397-458, rapt/Adaptive.py:187-222).  Here a whole batch
of particles in one field is packed into the slot-major state of
include/rapt_hip.h and advanced by one persistent kernel.  Particles whose row
buffer fills up are continued in further launches with exactly the state they
stopped in, so trajectories of any length come back complete.

Entry points:
  advance_particles(field, last7, mass, charge, delta)        Particle.advance
  advance_gcs(field, last5, mass, charge, mu, v, dt, delta)    GuidingCenter.advance
  advance_adaptive(field, pos, vel, t0, mass, charge, delta)  Adaptive
  advance(tracers, delta)   the tracer objects themselves, grouped by field
"""
import numpy as np

from . import _native as N
from .constants import c


class TraceResult:
    """Per-particle trajectory rows (list of arrays) plus the final state."""

    def __init__(self, rows, state, flags, status, naccpt, counters, launches):
        self.rows = rows
        self.state = state
        self.flags = flags
        self.status = status
        self.naccpt = naccpt
        self.counters = counters
        self.launches = launches

    @property
    def tcur(self):
        return self.state[N.S_TCUR]


def _params(params):
    if params is None:
        from . import params as P
        return P
    return params


def _run(field, mode, state, flags, row_cap, desc, device=0, max_launches=64, grow=2.0):
    ctx = N.context(device)
    ctx.set_field(field.descriptor())
    n = state.shape[1]
    width = N.ROW_WIDTH[mode]
    rows = [[] for _ in range(n)]
    idx = np.arange(n)
    cur_state, cur_flags, cur_cap = state.copy(), flags.copy(), np.asarray(row_cap, np.int64)
    final_state = state.copy()
    final_flags = flags.copy()
    status = np.zeros(n, np.int32)
    naccpt = np.zeros(n, np.int64)
    tot = dict(naccpt=0, nrejct=0, nstep=0, nfcn=0, nswitch=0, kernel_ms=0.0)
    launches = 0
    while idx.size:
        b = N.Batch(ctx, idx.size)
        try:
            b.upload(cur_state, cur_flags)
            b.set_rows(width, cur_cap)
            cnt = b.run(desc)
            out = b.download()
        finally:
            b.close()
        launches += 1
        for k in ("naccpt", "nrejct", "nstep", "nfcn", "nswitch"):
            tot[k] += getattr(cnt, k)
        tot["kernel_ms"] += cnt.kernel_ms
        off = b.row_off
        for j, i in enumerate(idx):
            m = out["nrows"][j]
            if m:
                rows[i].append(out["rows"][off[j]:off[j] + m])
        final_state[:, idx] = out["state"]
        final_flags[idx] = out["flags"]
        status[idx] = out["status"]
        naccpt[idx] += out["naccpt"]
        again = out["status"] == N.ST_CAP
        if not again.any():
            break
        if launches >= max_launches:
            break
        idx = idx[again]
        cur_state = out["state"][:, again]
        cur_flags = out["flags"][again]
        cur_cap = np.maximum(16, (cur_cap[again] * grow).astype(np.int64))
    rows = [np.concatenate(r) if r else np.zeros((0, width)) for r in rows]
    return TraceResult(rows, final_state, final_flags, status, naccpt, tot, launches)


def _rows_needed(delta, dt):
    with np.errstate(divide="ignore", invalid="ignore"):
        k = np.ceil(np.asarray(delta) / np.asarray(dt))
    k = np.where(np.isfinite(k), k, 0)
    return np.clip(k, 0, 1 << 40).astype(np.int64) + 3


def particle_state(last7, mass, charge):
    """Slot-major state for Particle rows (t, x, y, z, px, py, pz)."""
    last7 = np.atleast_2d(np.asarray(last7, dtype=np.float64))
    n = last7.shape[0]
    S = np.zeros((N.NSLOT, n))
    S[N.S_T] = last7[:, 0]
    S[N.S_Y0:N.S_Y0 + 6] = last7[:, 1:7].T
    S[N.S_MASS] = np.broadcast_to(mass, (n,))
    S[N.S_CHARGE] = np.broadcast_to(charge, (n,))
    return S


def gc_state(last5, mass, charge, mu, v, dt):
    """Slot-major state for GuidingCenter rows (t, x, y, z, ppar)."""
    last5 = np.atleast_2d(np.asarray(last5, dtype=np.float64))
    n = last5.shape[0]
    S = np.zeros((N.NSLOT, n))
    S[N.S_T] = last5[:, 0]
    S[N.S_Y0:N.S_Y0 + 4] = last5[:, 1:5].T
    S[N.S_MASS] = np.broadcast_to(mass, (n,))
    S[N.S_CHARGE] = np.broadcast_to(charge, (n,))
    S[N.S_MU] = np.broadcast_to(mu, (n,))
    S[N.S_V] = np.broadcast_to(v, (n,))
    S[N.S_DT] = np.broadcast_to(dt, (n,))
    return S


def particle_dt(field, last7, mass, charge, cyclotronresolution, device=0):
    """Output step of Particle.advance, cyclotron_period/resolution
    (rapt/Particle.py:282), with |B| evaluated on the device."""
    last7 = np.atleast_2d(last7)
    B = N.context(device).field_eval(field.descriptor(), "magB", last7[:, :4])[:, 0]
    mom = last7[:, 4:7]
    m = np.broadcast_to(mass, (last7.shape[0],))
    gm = np.sqrt(m ** 2 + np.sum(mom * mom, axis=1) / c ** 2)
    vsq = np.sum((mom / gm[:, None]) ** 2, axis=1)
    g = 1.0 / np.sqrt(1 - vsq / c ** 2)
    return 2 * np.pi * g * m / B / np.abs(charge) / cyclotronresolution


def gc_mirror_speed(field, last5, mass, mu, device=0):
    """Mirror field and speed of guiding centres, as GuidingCenter.bounceperiod
    computes them (rapt/GuidingCenter.py:593-606), with |B| from the device."""
    last5 = np.atleast_2d(last5)
    n = last5.shape[0]
    m = np.broadcast_to(np.asarray(mass, float), (n,))
    mu = np.broadcast_to(np.asarray(mu, float), (n,))
    B = N.context(device).field_eval(field.descriptor(), "magB", last5[:, :4])[:, 0]
    pp = last5[:, 4]
    g = np.sqrt(1 + 2 * mu * B / (m * c * c) + (pp / (m * c)) ** 2)
    nr = g - 1 < 1e-6
    p_nr = np.sqrt(2 * m * mu * B + pp ** 2)
    p_r = m * c * np.sqrt((g + 1) * (g - 1))
    p = np.where(nr, p_nr, p_r)
    with np.errstate(divide="ignore", invalid="ignore"):
        Bmir = np.where(nr, p_nr ** 2 / (2 * m * mu), p_r ** 2 / ((p_r - pp) * (p_r + pp)) * B)
    v = np.where(nr, p / m, p / m / g)
    return Bmir, v


def gc_bounce_dt(field, last5, mass, mu, params=None, device=0):
    """Output step bounceperiod()/bounceresolution for a batch of guiding
    centres, with the half-bounce path computed on the device."""
    P = _params(params)
    last5 = np.atleast_2d(np.asarray(last5, dtype=np.float64))
    Bmir, v = gc_mirror_speed(field, last5, mass, mu, device)
    sb, st = N.context(device).halfbounce(field.descriptor(), last5[:, :4], Bmir,
                                          P["fieldlineresolution"])
    if np.any(st != 1):
        bad = np.flatnonzero(st != 1)
        raise RuntimeError("bounce period failed for %d particle(s), first %d (status %d)"
                           % (bad.size, bad[0], st[bad[0]]))
    return 2.0 / v * sb / P["bounceresolution"], sb


def advance_particles(field, last7, mass, charge, delta, params=None, check=False,
                      stepper=None, row_cap=None, lanes=0, device=0, **kw):
    """Particle.advance for a batch.  last7: (n, 7) rows (t, x, y, z, px, py, pz)."""
    P = _params(params)
    last7 = np.atleast_2d(np.asarray(last7, dtype=np.float64))
    n = last7.shape[0]
    S = particle_state(last7, mass, charge)
    flags = np.zeros(n, np.int32)
    if row_cap is None:
        dt = particle_dt(field, last7, mass, charge, P["cyclotronresolution"], device)
        row_cap = _rows_needed(delta, dt)
    desc = N.make_desc(N.MODE_PARTICLE, P, delta, stepper=stepper, check=check, lanes=lanes,
                       **kw)
    return _run(field, N.MODE_PARTICLE, S, flags, row_cap, desc, device)


def advance_gcs(field, last5, mass, charge, mu, v, dt, delta, params=None,
                eom="TaoChanBrizardEOM", check=False, stepper=None, row_cap=None, lanes=0,
                device=0, **kw):
    """GuidingCenter.advance for a batch with per-particle output step dt."""
    P = _params(params)
    last5 = np.atleast_2d(np.asarray(last5, dtype=np.float64))
    n = last5.shape[0]
    S = gc_state(last5, mass, charge, mu, v, dt)
    flags = np.zeros(n, np.int32)
    if row_cap is None:
        row_cap = _rows_needed(delta, np.broadcast_to(dt, (n,)))
    desc = N.make_desc(N.MODE_GC, P, delta, eom=eom, stepper=stepper, check=check,
                       lanes=lanes, **kw)
    return _run(field, N.MODE_GC, S, flags, row_cap, desc, device)


def advance_adaptive(field, pos, vel, t0, mass, charge, delta, params=None, row_cap=None,
                     lanes=0, device=0, **kw):
    """Adaptive(pos, vel, t0, ...).advance(delta) for a batch.

    Rows are 8 wide: (mode, t, x, y, z, a, b, c); mode 0 = Particle with
    (a, b, c) = momentum, mode 1 = GuidingCenter with a = ppar, b = mu.  Each
    segment starts with its initial row, like the reference's trajlist.
    """
    P = _params(params)
    pos = np.atleast_2d(np.asarray(pos, dtype=np.float64))
    vel = np.atleast_2d(np.asarray(vel, dtype=np.float64))
    n = pos.shape[0]
    m = np.broadcast_to(mass, (n,)).astype(float)
    vsq = np.sum(vel * vel, axis=1)
    g = 1 / np.sqrt(1 - vsq / c ** 2)
    last7 = np.zeros((n, 7))
    last7[:, 0] = t0
    last7[:, 1:4] = pos
    last7[:, 4:7] = (m * g)[:, None] * vel
    S = particle_state(last7, mass, charge)
    flags = np.full(n, N.FLAG_AINIT, np.int32)
    if row_cap is None:
        dtp = particle_dt(field, last7, mass, charge, P["cyclotronresolution"], device)
        gcdt = float(P.get("GCtimestep", 0.0)) or 1.0
        row_cap = _rows_needed(delta, np.minimum(dtp, gcdt)) + 8
    desc = N.make_desc(N.MODE_ADAPTIVE, P, delta, lanes=lanes, **kw)
    return _run(field, N.MODE_ADAPTIVE, S, flags, row_cap, desc, device)


def continue_adaptive(field, mode, last, mass, charge, mu, v, delta, params=None,
                      row_cap=None, lanes=0, device=0, **kw):
    """A further Adaptive.advance(delta) of a batch from each tracer's last
    segment (rapt/Adaptive.py:187-222): mode[i] 0 continues a Particle whose
    last row is last[i] = (t, x, y, z, px, py, pz), 1 a GuidingCenter with
    last row (t, x, y, z, ppar), moment mu[i] and speed v[i] (last: an array
    of rows of one width, or a list of rows of either width).  Rows as
    advance_adaptive; the continued segment gets no start row."""
    P = _params(params)
    rows = [np.asarray(r, dtype=np.float64).ravel() for r in np.atleast_1d(last)] \
        if isinstance(last, (list, tuple)) else list(np.atleast_2d(np.asarray(last, float)))
    n = len(rows)
    last = np.zeros((n, 7))
    for i, r in enumerate(rows):
        last[i, :min(7, r.size)] = r[:7]
    mode = np.broadcast_to(np.asarray(mode, np.int32), (n,))
    S = np.zeros((N.NSLOT, n))
    S[N.S_T] = last[:, 0]
    gc = mode == 1
    S[N.S_Y0:N.S_Y0 + 6, ~gc] = last[~gc, 1:7].T
    S[N.S_Y0:N.S_Y0 + 4, gc] = last[gc, 1:5].T
    S[N.S_MASS] = np.broadcast_to(mass, (n,))
    S[N.S_CHARGE] = np.broadcast_to(charge, (n,))
    S[N.S_MU] = np.where(gc, np.broadcast_to(mu, (n,)), 0.0)
    S[N.S_V] = np.where(gc, np.broadcast_to(v, (n,)), 0.0)
    flags = np.where(gc, N.FLAG_GCMODE, 0).astype(np.int32)
    if row_cap is None:
        gcdt = float(P.get("GCtimestep", 0.0)) or 1.0
        dtp = np.full(n, gcdt)
        if np.any(~gc):
            dtp[~gc] = particle_dt(field, last[~gc, :7], np.broadcast_to(mass, (n,))[~gc],
                                   np.broadcast_to(charge, (n,))[~gc], P["cyclotronresolution"],
                                   device)
        row_cap = _rows_needed(delta, np.minimum(dtp, gcdt)) + 8
    desc = N.make_desc(N.MODE_ADAPTIVE, P, delta, lanes=lanes, **kw)
    return _run(field, N.MODE_ADAPTIVE, S, flags, row_cap, desc, device)


def advance(tracers, delta, device_dt=True, **kw):
    """Advance a list of Particle / GuidingCenter objects together (grouped by
    type and field; tracers that check adiabaticity launch apart from those
    that do not); equivalent to calling ``t.advance(delta)`` on each.

    device_dt: a group of guiding centres takes its output steps
    bounceperiod()/bounceresolution from the device half-bounce kernel (within
    the reference's 1e-4 quadrature tolerance) rather than from one host
    quadrature per tracer, which is what a single ``advance`` uses."""
    from .GuidingCenter import GuidingCenter
    from .Particle import Particle
    groups = {}
    for t in tracers:
        groups.setdefault((type(t), id(t.field)), []).append(t)
    for (tp, _), ts in groups.items():
        if tp is Particle:
            Particle._advance_many(ts, delta, **kw)
        elif tp is GuidingCenter:
            GuidingCenter._advance_many(ts, delta, device_dt=device_dt, **kw)
        else:
            for t in ts:
                t.advance(delta)
