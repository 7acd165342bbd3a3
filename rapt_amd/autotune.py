"""Launch-shape autotuner for the batched tracers.

The best launch shape depends on the batch: at 65,536 guiding centres a
grouped kernel (two, four or eight lanes per guiding centre, or a mix) wins
-- one lane per particle leaves one wave per SIMD and the slowest particle's
chain sets the time -- and at 1,048,576 the one-lane kernel wins (enough
particles to fill the chip, and it spends the fewest issue slots per step).
The tuner times trial launches of the real batch -- the same particles, the
same work order, a fraction of the advance -- for each candidate (block
size, lanes per particle, resident lanes, i.e. the work refill threshold)
and keeps the one with the highest accepted-steps rate (best of two
launches each).
Results are cached per (mode, stepper, field model, precision, equations,
batch size rounded to a power of two).

Trials reuse one uploaded batch and restore its state on the device between
runs (rapt_batch_reset), so the batch is traced from the same start each
time and the caller's run is unaffected.
"""
import math

import numpy as np

from . import _native as N

_cache = {}


def candidates(mode, stepper, precision, n):
    """Candidate launch shapes: dicts of rapt_trace_desc fields."""
    out = []
    grouped_ok = (mode == N.MODE_GC and stepper == N.STEP_DOPRI5 and precision == N.PREC_FP64)
    for block in (256, 128):
        if grouped_ok:
            out.append(dict(lanes_per_particle=4, block=block, lanes=0))
            out.append(dict(lanes_per_particle=2, block=block, lanes=0))
        out.append(dict(lanes_per_particle=1, block=block, lanes=0))
    # eight lanes per particle: the shortest step chain, for batches whose time
    # is the longest trace's
    if grouped_ok:
        out.append(dict(lanes_per_particle=8, block=256, lanes=0))
    # mixed: the most expensive particles (the front of the cost order) on
    # quads -- the very front on octets with lanes_per_particle 5 -- and the
    # rest on pairs
    if grouped_ok:
        for frac in (8, 16):
            out.append(dict(lanes_per_particle=3, block=256, lanes=0,
                            mixed_quads=max(1, n // frac)))
        out.append(dict(lanes_per_particle=5, block=256, lanes=0,
                        mixed_octets=max(1, n // 128), mixed_quads=max(1, n // 8)))
    # fewer resident lanes than particles: refill from the queue sooner
    if n >= 8 * 65536:
        out.append(dict(lanes_per_particle=1, block=256, lanes=4 * 65536))
    return out


def key(desc, model, n, ordered=False):
    return (desc.mode, desc.stepper, model, desc.precision, desc.eom,
            int(math.log2(max(n, 1))), ordered)


def tune(field, S, flags, row_cap, desc, trial_fraction=0.1, device=0, force=False, log=None,
         cost=None, repeats=2):
    """Choose a launch shape for this batch.  Returns (best dict, table of
    (shape, accepted steps/s)).  desc is not modified.  cost: the work order
    the run will use (Batch.set_cost), applied to the trials too, so that
    they see the run's schedule."""
    n = S.shape[1]
    fd = field.descriptor()
    k = key(desc, fd["model"], n, cost is not None)
    if not force and k in _cache:
        return _cache[k], []
    ctx = N.context(device)
    ctx.set_field(fd)
    width = N.ROW_WIDTH[desc.mode]
    b = N.Batch(ctx, n)
    table = []
    try:
        b.upload(S, flags)
        if cost is not None:
            b.set_cost(cost)
        b.set_rows(width, row_cap)
        trial = N.TraceDesc.from_buffer_copy(desc)
        trial.delta = desc.delta * trial_fraction
        for cand in candidates(desc.mode, desc.stepper, desc.precision, n):
            for f, v in cand.items():
                setattr(trial, f, v)
            # best of `repeats` launches: the first launch of a kernel also
            # loads its code, and single trials vary by ~10% between runs
            best_ms = float("inf")
            for _ in range(repeats):
                b.reset()
                c = b.run(trial)
                best_ms = min(best_ms, c.kernel_ms)
            rate = c.naccpt / max(best_ms, 1e-6) * 1e3
            table.append((dict(cand), rate))
            if log:
                log("autotune %s: %.3g accepted steps/s (%.1f ms)" % (cand, rate, best_ms))
    finally:
        b.close()
    best = max(table, key=lambda t: t[1])[0]
    _cache[k] = best
    return best, table


def trial_cost(field, S, flags, row_cap, desc, trial_fraction=0.1, device=0):
    """Per-particle cost estimate for the library's work order
    (Batch.set_cost): accepted steps of each particle in a trial launch of the
    same batch over a fraction of the advance.  Step counts per unit of
    simulated time change slowly along a trace, so the trial ranks the
    particles nearly as the full advance does."""
    n = S.shape[1]
    ctx = N.context(device)
    ctx.set_field(field.descriptor())
    b = N.Batch(ctx, n)
    try:
        b.upload(S, flags)
        b.set_rows(N.ROW_WIDTH[desc.mode], row_cap)
        trial = N.TraceDesc.from_buffer_copy(desc)
        trial.delta = desc.delta * trial_fraction
        b.run(trial)
        return b.download(rows=False)["naccpt"].astype(np.float64)
    finally:
        b.close()


def apply(desc, shape):
    """Copy a tuned shape into a rapt_trace_desc."""
    for f, v in shape.items():
        setattr(desc, f, v)
    return desc
