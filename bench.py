"""Throughput benchmark: FP64 guiding-centre particle-steps/s on one MI355X.

Workload (the batched GuidingCenter configuration): 65,536 electrons in the
DoubleDipole field, log10 K[eV] ~ U[4,7], L ~ U[3,8], azimuth ~ U[0,2pi),
equatorial pitch angle ~ U[15,85] deg (seed 1); GuidingCenter.advance(60)
with the Tao-Chan-Brizard equations, DOPRI5 at rtol = atol = 1e-8, output
step bounceperiod()/bounceresolution per particle.  One bench step is one
full advance(60) of the whole batch; the state is restored on the device
between steps, so inputs stay resident in HBM.  A particle-step is one
accepted integrator step of one particle.

The same line carries, for the driver's one run: full-length parity of the
timed rows against the oracle (the first --cpu-sample particles, the sample
the CPU baseline integrates anyway), the rate with the library's default
work order, the time to deliver the rows to the host, and summaries of the
other configurations (Particle dipole batch with its own parity and first-
invariant check, Adaptive current sheet, mixed precision, surrogate) under
"configs" (--headline-only skips them).

Prints one JSON line.  With --gpus > 1 it exits: this build measures one GPU.

Other workloads (--workload, one JSON line each; not the default line):
  particle_dipole   1024 protons, 1 MeV, EarthDipole, Particle.advance(10), DOP853
  adaptive_sheet    262,144 protons, current sheet + uniform Ey, Adaptive.advance(120)
  mixed_precision   1,048,576 protons and electrons, EarthDipole - 10 nT z, advance(30):
                    FP64 against FP32-field, errors per species
  surrogate         bounce-drift MLP ensemble (4 x 3-256-256-256-3, BF16 MFMA): trained on
                    2^20 traced samples, queried 2^24 times with traced fallback
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic FP64 FLOP of the executed formulation (counted from
# rapt_amd/csrc/rapt_fields.hpp and rapt_steppers.hpp: +,-,* = 1, fma = 2,
# div/sqrt/transcendental = 1)
FLOP_RHS_TCB_DOUBLEDIPOLE = 484      # one Tao-Chan-Brizard RHS, 7 field evaluations
FLOP_DOPRI5_STEP_N4 = 282            # stage combinations, error norm, control
FLOP_RHS_TCB_REFERENCE = 795         # same RHS as the reference writes it (13 field evals)
FLOP_RHS_LORENTZ_DIPOLE = 34         # one Lorentz RHS in EarthDipole (static gamma m)
FLOP_RHS_LORENTZ_NOFIELD = 15        # the same without the field: v = p/(gamma m), p x B, q/(gamma m)
FLOP_DOP853_STEP_N6 = 988            # DOP853 attempt without its RHS: stage sums 666,
                                     # stage times 23, solution + two error norms 282,
                                     # norm combination 7, step-size control 10
PEAK_FP64 = 78.6e12                  # MI355X FP64 vector (and matrix) peak, FLOP/s
PEAK_HBM = 8.0e12                    # bytes/s
ROW_BYTES_GC = 40                    # one (t, x, y, z, ppar) output row
# HBM bytes per launch of the default batch's kernel (k_trace_gcmix with 8,192
# quads, the autotuner's usual choice), from rocprofv3 FETCH_SIZE and
# WRITE_SIZE (KiB) scaled by a 256 MiB copy probe (read x2.00, write x1.00);
# profiles/k_trace_gcmix_doubledipole_65536.md has the commands
PROFILED_TRAFFIC = {65536: (44261.8125 * 2.0 + 3032515.625 * 1.0) * 1024}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(n, device=0):
    import rapt_amd as rapt
    from rapt_amd import batch, workloads as W
    wl = W.doubledipole_electrons_gc(n=n)
    f = rapt.fields.DoubleDipole()
    m, q = rapt.m_el, -rapt.e
    v = wl["v"]
    g = 1 / np.sqrt(1 - (v / rapt.c) ** 2)
    pa = np.radians(wl["pa"])
    ppar = g * m * v * np.cos(pa)
    # GuidingCenter.__init__: mu = gamma^2 m (v - vpar)(v + vpar) / (2B)
    from rapt_amd import _native as N
    B = N.context(device).field_eval(f.descriptor(), "magB",
                                     np.c_[np.zeros(n), wl["pos"]])[:, 0]
    vpar = ppar / (m * g)
    mu = g ** 2 * m * (v - vpar) * (v + vpar) / (2 * B)
    last5 = np.c_[np.zeros(n), wl["pos"], ppar]
    params = dict(rapt.params, solvertolerances=(1e-8, 1e-8))
    t0 = time.time()
    dt, _ = batch.gc_bounce_dt(f, last5, m, mu, params=params, device=device)
    log("bounce-period output steps for %d particles: %.2f s" % (n, time.time() - t0))
    return dict(field=f, last5=last5, m=m, q=q, mu=mu, v=v, dt=dt, params=params,
                delta=wl["delta"])


def cpu_baseline(w, sample, threads):
    """The oracle (C restatement of the reference, OpenMP over particles) on
    the first `sample` particles of the same workload.  Returns (accepted
    steps, seconds, (rows, row offsets, row counts, status)) -- the rows are
    the parity check's reference."""
    from oracle import oracle as O
    f = O.field_from_desc(w["field"].descriptor())
    prm = O.gc_params(rtol=1e-8, atol=1e-8)
    s = slice(0, sample)
    cap = (np.ceil(w["delta"] / w["dt"][s]) + 3).astype(np.int64)
    t0 = time.perf_counter()
    rows, off, nrows, status, cnt = O.batch_gc(f, prm, 0, w["last5"][s], w["m"], w["q"],
                                               w["mu"][s], w["v"][s], w["dt"][s], w["delta"],
                                               cap, nthreads=threads)
    el = time.perf_counter() - t0
    return cnt.naccpt, el, (rows, off, nrows, status)


def _rel_pos_err(a, b):
    if a.shape != b.shape:
        return np.inf
    if len(a) == 0:
        return 0.0
    d = np.linalg.norm(a[:, 1:4] - b[:, 1:4], axis=1)
    return float(np.max(d / np.linalg.norm(b[:, 1:4], axis=1)))


def parity_summary(gpu_rows, ref_rows, truth_fn=None, truth_max=64, bar=1e-7):
    """Position parity of per-particle row arrays at the output times: the
    largest relative error over each particle's rows, the share within `bar`,
    and -- for up to truth_max particles beyond it, worst first -- how close
    the device and the reference come to a tight-tolerance trajectory
    (truth_fn(indices) -> list of row arrays).  An adaptive trace that is not
    bit-identical to the reference's leaves it where a step is accepted on one
    side and rejected on the other, and from there the two differ by the
    integrator's own error; the ratio device/reference distance to the tight
    trajectory then stays near 1."""
    err = np.array([_rel_pos_err(g, r) for g, r in zip(gpu_rows, ref_rows)])
    fin = err[np.isfinite(err)]
    out = {"particles": int(len(err)), "rows": int(sum(len(r) for r in ref_rows)),
           "row_count_mismatches": int(np.sum(~np.isfinite(err))),
           "max_pos_rel_err": float(fin.max()) if fin.size else None,
           "median_pos_rel_err": float(np.median(fin)) if fin.size else None,
           "p99_pos_rel_err": float(np.percentile(fin, 99)) if fin.size else None,
           "frac_within_%g" % bar: float(np.mean(err <= bar)), "bar": bar}
    far = np.nonzero(err > bar)[0]
    out["beyond_bar"] = int(far.size)
    if far.size and truth_fn is not None:
        sel = far[np.argsort(-err[far])][:truth_max]
        truth = truth_fn(sel)
        dg = np.array([_rel_pos_err(gpu_rows[i], t) for i, t in zip(sel, truth)])
        do = np.array([_rel_pos_err(ref_rows[i], t) for i, t in zip(sel, truth)])
        # a tight run that stopped early (its rows do not line up) checks nothing
        ok = np.isfinite(dg) & np.isfinite(do) & (do > 0)
        r = dg[ok] / do[ok]
        out["beyond_bar_vs_tight_tolerance"] = {
            "checked": int(ok.sum()), "tight_rtol": 1e-12,
            "tight_run_incomplete": int((~ok).sum()),
            "reference_dist_median": float(np.median(do[ok])) if ok.any() else None,
            "device_dist_median": float(np.median(dg[ok])) if ok.any() else None,
            "device_over_reference_median": float(np.median(r)) if ok.any() else None,
            "device_over_reference_max": float(np.max(r)) if ok.any() else None,
            "worst": [{"particle": int(i), "pos_rel_err_vs_reference": float(err[i]),
                       "device_dist": float(a), "reference_dist": float(b)}
                      for i, a, b in list(zip(sel, dg, do))[:4]]}
    return out

def timed_run(field, mode, S, flags, cap, desc, steps, warmup, width, rows=False, cost=None):
    """Upload once, then warmup + timed runs of the same batch (state reset on
    the device between runs).  Returns (totals, seconds, last download)."""
    import ctypes
    from rapt_amd import _native as N
    ctx = N.context(0)
    ctx.set_field(field.descriptor())
    b = N.Batch(ctx, S.shape[1])
    b.upload(S, flags)
    if cost is not None:
        b.set_cost(cost)
    b.set_rows(width, cap)
    for i in range(warmup):
        b.reset()
        c = b.run(desc)
        log("warmup %d: %.1f ms" % (i, c.kernel_ms))
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()
    tot = dict(naccpt=0, nrejct=0, nstep=0, nfcn=0, nswitch=0, kernel_ms=0.0)
    t0 = time.perf_counter()
    for i in range(steps):
        b.reset()
        c = b.run(desc)
        for k in ("naccpt", "nrejct", "nstep", "nfcn", "nswitch"):
            tot[k] += getattr(c, k)
        tot["kernel_ms"] += c.kernel_ms
        log("step %d: %.1f ms, %d accepted steps" % (i, c.kernel_ms, c.naccpt))
    hip.hipDeviceSynchronize()
    el = time.perf_counter() - t0
    out = b.download(rows=rows)
    out["row_off"] = np.asarray(b.row_off, np.int64).copy()
    b.close()
    return tot, el, out


def flat_rows(out, idx):
    """Rows of particles idx from a flat download: (rows, owner, group starts)."""
    nr = out["nrows"][idx]
    owner = np.repeat(idx, nr)
    first = np.r_[0, np.cumsum(nr)[:-1]]
    pos = np.arange(int(nr.sum())) - np.repeat(first, nr)
    g = out["row_off"][owner] + pos
    return out["rows"][g], owner, first


def base_line(metric, value, unit, args, el, dtype, config):
    return {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
            "data": "synthetic (seeded)", "config": config}


def run_particle_dipole(args, parity=None):
    """1024 protons at 1 MeV in EarthDipole, Particle.advance(10), DOP853.
    With parity: every proton's rows against the oracle, and the drift of the
    first invariant (mu of each row: the guiding centre by the fixed-point
    iteration, then gamma^2 m v_perp^2 / 2B, rapt_particle_gc on the device
    for both trajectories) against the oracle's trajectory."""
    import rapt_amd as rapt
    from rapt_amd import _native as N, batch as BT, workloads as W
    if parity is None:
        parity = not args.no_cpu
    n = args.n or 1024
    wl = W.dipole_protons_1mev(n=n)
    f = rapt.fields.EarthDipole()
    g = 1 / np.sqrt(1 - np.sum(wl["vel"] ** 2, axis=1) / rapt.c ** 2)
    last7 = np.zeros((n, 7))
    last7[:, 1:4] = wl["pos"]
    last7[:, 4:7] = (rapt.m_pr * g)[:, None] * wl["vel"]
    S = BT.particle_state(last7, rapt.m_pr, rapt.e)
    dt = BT.particle_dt(f, last7, rapt.m_pr, rapt.e, 20)
    cap = BT._rows_needed(10.0, dt)
    desc = N.make_desc(N.MODE_PARTICLE, dict(rapt.params, cyclotronresolution=20), 10.0,
                       lanes=args.lanes, block=args.block)
    tot, el, out = timed_run(f, N.MODE_PARTICLE, S, np.zeros(n, np.int32), cap, desc,
                             args.steps, args.warmup, 7, rows=parity)
    value = tot["naccpt"] / el
    res = base_line("FP64 particle-steps/s (1 GPU)", value, "particle-steps/s", args, el, "f64",
                    {"workload": "EarthDipole Particle batch: %d protons, 1 MeV, r U[4,7] Re, "
                                 "pitch U[10,90] deg; DOP853 default tolerances, advance(10), "
                                 "dt=T_c/20" % n, "particles": n})
    launches = args.steps
    ksec = tot["kernel_ms"] / launches / 1e3
    flop = (tot["nfcn"] * FLOP_RHS_LORENTZ_DIPOLE + tot["nstep"] * FLOP_DOP853_STEP_N6) / launches
    res["accepted_steps_per_step"] = tot["naccpt"] / launches
    res["rhs_evals_per_step"] = tot["nfcn"] / launches
    res["kernel_ms_avg"] = tot["kernel_ms"] / launches
    res["particles_done"] = int(np.sum(out["status"] == 1))
    res["roofline"] = {"bound": "valu_fp64", "achieved": flop / ksec / 1e12,
                       "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                       "frac": flop / ksec / PEAK_FP64, "traffic": None,
                       "flop_per_launch": flop,
                       "flop_per_particle_step": flop / (tot["naccpt"] / launches),
                       "note": "%d lanes are %.3f waves per SIMD: the launch is the longest "
                               "trace's dependent chain" % (n, n / 64 / 1024)}
    if parity:
        from oracle import oracle as O
        fo = O.field_from_desc(f.descriptor())
        threads = min(16, os.cpu_count() or 1)
        t0 = time.perf_counter()
        rows_o, off_o, nrows_o, st_o, cnt_o = O.batch_particle(
            fo, O.particle_params(cyclotronresolution=20), last7, rapt.m_pr, rapt.e, 10.0, cap,
            nthreads=threads)
        t_cpu = time.perf_counter() - t0
        ref = [rows_o[off_o[i]:off_o[i] + nrows_o[i]] for i in range(n)]
        gpu = [out["rows"][out["row_off"][i]:out["row_off"][i] + out["nrows"][i]]
               for i in range(n)]

        def truth(sel):
            r, o, k, st, _ = O.batch_particle(
                fo, O.particle_params(cyclotronresolution=20, rtol=1e-12, atol=1e-30,
                                      nsteps=10 ** 7), last7[sel], rapt.m_pr, rapt.e, 10.0,
                cap[sel], nthreads=threads)
            return [r[o[j]:o[j] + k[j]] for j in range(len(sel))]
        res["parity"] = parity_summary(gpu, ref, truth)
        # first invariant along both trajectories, start row included
        ctx = N.context(0)
        allg = np.vstack([np.vstack([last7[i:i + 1], gpu[i]]) for i in range(n)])
        allo = np.vstack([np.vstack([last7[i:i + 1], ref[i]]) for i in range(n)])
        cg = np.r_[0, np.cumsum([len(x) + 1 for x in gpu])]
        co = np.r_[0, np.cumsum([len(x) + 1 for x in ref])]
        mg, sg = ctx.particle_gc(f.descriptor(), allg, rapt.m_pr, rapt.e)
        mo, so = ctx.particle_gc(f.descriptor(), allo, rapt.m_pr, rapt.e)
        dg = np.array([np.max(np.abs(mg[cg[i]:cg[i + 1], 5] / mg[cg[i], 5] - 1))
                       for i in range(n)])
        do = np.array([np.max(np.abs(mo[co[i]:co[i + 1], 5] / mo[co[i], 5] - 1))
                       for i in range(n)])
        ok = np.isfinite(dg) & np.isfinite(do)
        rr = dg[ok] / do[ok]
        res["mu_drift"] = {"particles": int(ok.sum()),
                           "device_median": float(np.median(dg[ok])),
                           "reference_median": float(np.median(do[ok])),
                           "device_over_reference_median": float(np.median(rr)),
                           "device_over_reference_max": float(np.max(rr)),
                           "frac_not_larger": float(np.mean(dg[ok] <= do[ok] * (1 + 1e-9))),
                           "gc_iteration_failures": int(np.sum(sg < 0) + np.sum(so < 0))}
        res["cpu_baseline"] = {"value": cnt_o.naccpt / t_cpu, "unit": "particle-steps/s",
                               "cores": threads, "kind": "port",
                               "sample": "all %d protons, full advance(10), C oracle with "
                                         "OpenMP (%.1f s)" % (n, t_cpu)}
        res["gpu_vs_cpu"] = value / (cnt_o.naccpt / t_cpu)
    return res


def run_grid(args):
    """Particle.advance in a gridded field: EarthDipole sampled on a uniform
    257 x 257 x 193 grid over |x|, |y| <= 8 Re, |z| <= 6 Re (B only, 408 MB
    in HBM), the config-1 proton population at n = 65,536, advance(10),
    DOP853; the same batch in the analytic dipole for comparison."""
    import rapt_amd as rapt
    from rapt_amd import _native as N, batch as BT, workloads as W
    n = args.n or 65536
    Re = rapt.Re
    ax = (np.linspace(-8, 8, 257) * Re, np.linspace(-8, 8, 257) * Re,
          np.linspace(-6, 6, 193) * Re)

    class SampledDipole(rapt.fields.Grid):
        def parsefile(self, filename):
            x, y, z = ax
            X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
            r2 = X * X + Y * Y + Z * Z
            with np.errstate(divide="ignore", invalid="ignore"):
                s = rapt.fields.EarthDipole()._coeff / r2 ** 2.5
            s[~np.isfinite(s)] = 0.0
            zero = np.zeros_like(X)
            return dict(time=0.0, x=x, y=y, z=z, Bx=s * X * Z, By=s * Y * Z,
                        Bz=s * (Z * Z - r2 / 3), Ex=zero, Ey=zero, Ez=zero)

    t0 = time.perf_counter()
    grid = SampledDipole(["dipole"])
    log("grid sampled in %.1f s" % (time.perf_counter() - t0))
    ana = rapt.fields.EarthDipole()
    wl = W.dipole_protons_1mev(n=n)
    g = 1 / np.sqrt(1 - np.sum(wl["vel"] ** 2, axis=1) / rapt.c ** 2)
    last7 = np.zeros((n, 7))
    last7[:, 1:4] = wl["pos"]
    last7[:, 4:7] = (rapt.m_pr * g)[:, None] * wl["vel"]
    S = BT.particle_state(last7, rapt.m_pr, rapt.e)
    dt = BT.particle_dt(ana, last7, rapt.m_pr, rapt.e, 20)
    S[N.S_DT] = dt
    cap = BT._rows_needed(10.0, dt)
    desc = N.make_desc(N.MODE_PARTICLE, dict(rapt.params), 10.0, lanes=args.lanes)
    # work order for the gathers: particles whose start cells are close in
    # Z order share wavefronts, so a wave's corner reads share cache lines
    cells = [np.clip(((last7[:, 1 + k] - a[0]) / (a[1] - a[0])).astype(np.int64), 0, 1023)
             for k, a in enumerate(ax)]
    code = np.zeros(n, np.int64)
    for bit in range(10):
        for k in range(3):
            code |= ((cells[k] >> bit) & 1) << (3 * bit + k)
    cost = None if args.order != "cost" else -code.astype(np.float64)
    runs = {}
    for name, f in (("analytic", ana), ("grid", grid)):
        tot, el, out = timed_run(f, N.MODE_PARTICLE, S, np.zeros(n, np.int32), cap, desc,
                                 args.steps, args.warmup, 7, cost=cost if name == "grid" else None)
        runs[name] = (tot, el, out)
    tot, el, out = runs["grid"]
    value = tot["naccpt"] / el
    evals = tot["nfcn"] / args.steps
    gather = evals * 8 * 24.0          # 8 cell corners x (Bx, By, Bz) per evaluation
    ksec = tot["kernel_ms"] / args.steps / 1e3
    fa, fg = runs["analytic"][2]["state"], out["state"]
    pos_a, pos_g = fa[1:4], fg[1:4]
    rel = np.linalg.norm(pos_g - pos_a, axis=0) / np.linalg.norm(pos_a, axis=0)
    res = base_line("FP64 particle-steps/s in a gridded field (1 GPU)", value,
                    "particle-steps/s", args, el, "f64",
                    {"workload": "EarthDipole sampled on a uniform 257 x 257 x 193 grid "
                                 "(|x|,|y| <= 8 Re, |z| <= 6 Re, 408 MB of B in HBM); %d protons "
                                 "1 MeV, r U[4,7] Re, pitch U[10,90] deg; DOP853 default "
                                 "tolerances, advance(10), dt=T_c/20" % n, "particles": n})
    res.update({
        "kernel_ms_avg": tot["kernel_ms"] / args.steps,
        "accepted_steps_per_step": tot["naccpt"] / args.steps,
        "field_evaluations_per_step": evals,
        "analytic_dipole_steps_per_s": runs["analytic"][0]["naccpt"] / runs["analytic"][1],
        "analytic_dipole_kernel_ms_avg": runs["analytic"][0]["kernel_ms"] / args.steps,
        "final_position_rel_diff_grid_vs_analytic": _pct(rel),
        "particles_done": int(np.sum(out["status"] == 1)),
        "roofline": {"bound": "hbm", "achieved": gather / ksec / 1e9, "peak": PEAK_HBM / 1e9,
                     "unit": "GB/s", "frac": gather / ksec / PEAK_HBM, "traffic": None,
                     "algorithmic_bytes_per_launch": gather,
                     "note": "gather bytes if every corner read came from HBM; L2/MALL "
                             "reuse between nearby evaluations lowers the real traffic"}})
    return res


def run_user_field(args):
    """The reference's plugin contract at speed: EarthDipole + (0, 0, -10 nT)
    written as a user _Field subclass (B overridden in Python, calling the
    parent's B), config 1's protons at n = 65,536, Particle.advance(1).  The
    program compiled into the tracer kernel at run time against the same
    program interpreted (bit-identical rows) and the built-in model of the
    same field (same formula, different rounding)."""
    import rapt_amd as rapt
    from rapt_amd import _native as N, batch as BT, workloads as W

    class DipoleSouthward(rapt.fields.EarthDipole):
        def __init__(self, Bu=(0.0, 0.0, -10e-9)):
            rapt.fields.EarthDipole.__init__(self)
            self._Bu = np.array(Bu, dtype=float)

        def B(self, tpos):
            return rapt.fields.EarthDipole.B(self, tpos) + self._Bu

    n = args.n or 65536
    wl = W.dipole_protons_1mev(n=n)
    g = 1 / np.sqrt(1 - np.sum(wl["vel"] ** 2, axis=1) / rapt.c ** 2)
    last7 = np.zeros((n, 7))
    last7[:, 1:4] = wl["pos"]
    last7[:, 4:7] = (rapt.m_pr * g)[:, None] * wl["vel"]
    user, builtin = DipoleSouthward(), rapt.fields.DipolePlusUniform((0.0, 0.0, -10e-9))
    prog = user.descriptor()["expr"]
    S = BT.particle_state(last7, rapt.m_pr, rapt.e)
    dt = BT.particle_dt(builtin, last7, rapt.m_pr, rapt.e, 20)
    S[N.S_DT] = dt
    cap = BT._rows_needed(1.0, dt)
    desc = N.make_desc(N.MODE_PARTICLE, dict(rapt.params, cyclotronresolution=20), 1.0)
    ctx = N.context(0)
    ctx.set_expr_mode(True)
    t0 = time.perf_counter()  # the first compiled run compiles and loads its kernel
    BT.advance_particles(user, last7[:1], rapt.m_pr, rapt.e, 1e-3)
    compile_s = time.perf_counter() - t0
    runs = {}
    try:
        for name, f, compiled in (("compiled", user, True), ("interpreted", user, False),
                                  ("builtin", builtin, True)):
            ctx.set_expr_mode(compiled)
            runs[name] = timed_run(f, N.MODE_PARTICLE, S, np.zeros(n, np.int32), cap, desc,
                                   args.steps, args.warmup, 7, rows=name != "builtin")
    finally:
        ctx.set_expr_mode(True)
    tot, el, out = runs["compiled"]
    oi = runs["interpreted"][2]
    same = bool(np.array_equal(out["state"], oi["state"]) and
                np.array_equal(out["nrows"], oi["nrows"]) and
                np.array_equal(out["rows"], oi["rows"]))
    pa, pb = out["state"][1:4], runs["builtin"][2]["state"][1:4]
    rel = np.linalg.norm(pa - pb, axis=0) / np.linalg.norm(pb, axis=0)
    value = tot["naccpt"] / el
    ops = int(prog.nb)
    ksec = tot["kernel_ms"] / args.steps / 1e3
    flop = (tot["nfcn"] * (FLOP_RHS_LORENTZ_NOFIELD + ops) + tot["nstep"] * FLOP_DOP853_STEP_N6) \
        / args.steps
    res = base_line("FP64 particle-steps/s in a user-written field (1 GPU)", value,
                    "particle-steps/s", args, el, "f64",
                    {"workload": "EarthDipole + (0,0,-10 nT) written as a _Field subclass "
                                 "(B in Python, %d traced operations); %d protons 1 MeV, "
                                 "r U[4,7] Re, pitch U[10,90] deg; DOP853 default tolerances, "
                                 "advance(1), dt=T_c/20" % (ops, n), "particles": n})
    res.update({
        "kernel_ms_avg": {k: r[0]["kernel_ms"] / args.steps for k, r in runs.items()},
        "particle_steps_per_s": {k: r[0]["naccpt"] / r[1] for k, r in runs.items()},
        "gc_doubledipole": user_field_gc(args),
        "first_compiled_run_s": compile_s,
        "compiled_equals_interpreted": same,
        "final_position_rel_diff_vs_builtin": _pct(rel),
        "particles_done": int(np.sum(out["status"] == 1)),
        "roofline": {"bound": "valu_fp64", "achieved": flop / ksec / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                     "frac": flop / ksec / PEAK_FP64, "traffic": None,
                     "flop_per_launch": flop,
                     "note": "%d lanes: %.2f waves per SIMD" % (n, n / 64 / 1024)}})
    return res


def user_field_gc(args):
    """The headline's guiding-centre batch (65,536 electrons, advance(60),
    lane quads) in the two-dipole field written as a user _Field subclass
    (the image dipole shifted in place, pow(r2, 5/2)), compiled at run time,
    against the built-in DoubleDipole in the same layout and work order."""
    import rapt_amd as rapt
    from rapt_amd import _native as N, batch as BT

    class TwinDipole(rapt.fields._Field):
        def __init__(self, B0=3.07e-5, d=20 * rapt.Re, k=1):
            rapt.fields._Field.__init__(self)
            self.gradientstepsize = rapt.Re / 1000
            self._d, self._k, self._c = d, k, -B0 * rapt.Re ** 3

        def B(self, tpos):
            t, x, y, z = tpos
            B1 = np.array([3 * x * z, 3 * y * z, (2 * z * z - x * x - y * y)]) / \
                pow(x * x + y * y + z * z, 5.0 / 2.0)
            x -= self._d
            B2 = self._k * np.array([3 * x * z, 3 * y * z, (2 * z * z - x * x - y * y)]) / \
                pow(x * x + y * y + z * z, 5.0 / 2.0)
            return self._c * (B1 + B2)

    n = 65536
    w = build_workload(n)
    S = BT.gc_state(w["last5"], w["m"], w["q"], w["mu"], w["v"], w["dt"])
    cap = BT._rows_needed(w["delta"], w["dt"])
    desc = N.make_desc(N.MODE_GC, w["params"], w["delta"], lanes_per_particle=4)
    out = {}
    for name, f in (("compiled", TwinDipole()), ("builtin", w["field"])):
        tot, el, o = timed_run(f, N.MODE_GC, S, np.zeros(n, np.int32), cap, desc,
                               min(args.steps, 2), 1, 5)
        out[name] = {"particle_steps_per_s": tot["naccpt"] / el,
                     "kernel_ms_avg": tot["kernel_ms"] / min(args.steps, 2),
                     "state": o["state"][1:4]}
    a, b = out["compiled"].pop("state"), out["builtin"].pop("state")
    rel = np.linalg.norm(a - b, axis=0) / np.linalg.norm(b, axis=0)
    out["final_position_rel_diff_vs_builtin"] = _pct(rel)
    out["workload"] = ("%d DoubleDipole electrons of the headline batch, advance(60), lane "
                       "quads, output-step order" % n)
    return out


def run_adaptive_sheet(args):
    """262,144 protons, B = (20 nT tanh(z/0.5 Re), 0, 1 nT), E = 0.3 mV/m y,
    Adaptive.advance(120) with GCtimestep 0.5 s."""
    import rapt_amd as rapt
    from rapt_amd import _native as N, batch as BT, workloads as W
    n = args.n or 262144
    wl = W.current_sheet_protons(n=n)
    f = rapt.fields.CurrentSheet()
    prm = dict(rapt.params, GCtimestep=0.5)
    m = rapt.m_pr
    g = 1 / np.sqrt(1 - np.sum(wl["vel"] ** 2, axis=1) / rapt.c ** 2)
    last7 = np.zeros((n, 7))
    last7[:, 1:4] = wl["pos"]
    last7[:, 4:7] = (m * g)[:, None] * wl["vel"]
    S = BT.particle_state(last7, m, rapt.e)
    flags = np.full(n, N.FLAG_AINIT, np.int32)
    # rows: Particle segments may reach the strongest field of the sheet
    # (sqrt(B0^2 + Bn^2)), and every switch adds a segment start row
    bmax = np.hypot(f.B0, f.Bn)
    dt_min = 2 * np.pi * m / (rapt.e * bmax) / prm["cyclotronresolution"]
    cap = np.full(n, int(np.ceil(120.0 / min(dt_min, 0.5))) + 200, np.int64)
    desc = N.make_desc(N.MODE_ADAPTIVE, prm, 120.0, lanes=args.lanes)
    tot, el, out = timed_run(f, N.MODE_ADAPTIVE, S, flags, cap, desc, args.steps, args.warmup,
                             8)
    value = tot["naccpt"] / el
    res = base_line("FP64 particle-steps/s (1 GPU)", value, "particle-steps/s", args, el, "f64",
                    {"workload": "current sheet Adaptive batch: %d protons, log10 K[eV] U[3,5], "
                                 "isotropic; B0 20 nT, lam 0.5 Re, Bn 1 nT, Ey 0.3 mV/m; "
                                 "advance(120), GCtimestep 0.5 s" % n, "particles": n})
    res["accepted_steps_per_step"] = tot["naccpt"] / args.steps
    res["switches_per_step"] = tot["nswitch"] / args.steps
    res["kernel_ms_avg"] = tot["kernel_ms"] / args.steps
    st = out["status"]
    res["status_counts"] = {str(k): int(np.sum(st == k)) for k in np.unique(st)}
    if not args.no_cpu:
        res.update(adaptive_oracle_leg(f, wl, prm, m, min(1024, n)))
    return res


def adaptive_oracle_leg(f, wl, prm, m, k):
    """CPU baseline and full-length parity of the Adaptive batch on its first
    k particles: the oracle's Adaptive.advance(120) of each, timed on one
    host core, against the device's rows of the same particles (traced in a
    launch of their own: a particle's rows do not depend on the batch around
    it).  Orbits after a switch are chaotic (the reference says so), so the
    bar is the initial mode, the first switch's row and time, and the rows up
    to it (1e-7)."""
    import rapt_amd as rapt
    from rapt_amd import batch as BT
    from oracle import oracle as O
    gpu = BT.advance_adaptive(f, wl["pos"][:k], wl["vel"][:k], 0.0, m, rapt.e, 120.0,
                              params=prm)
    fo = O.field_from_desc(f.descriptor())
    po = O.particle_params(gctimestep=0.5)

    def first_switch(rows):
        j = np.nonzero(np.diff(rows[:, 0]) != 0)[0]
        return int(j[0]) + 1 if j.size else None

    t0 = time.perf_counter()
    ref = [O.adaptive_run(fo, po, wl["pos"][i], wl["vel"][i], 0.0, m, rapt.e, 120.0)
           for i in range(k)]
    t_cpu = time.perf_counter() - t0
    steps = sum(info["naccpt"] for _, info in ref)
    same_mode = same_first = same_count = 0
    err = []
    for i, (ro, info) in enumerate(ref):
        got = gpu.rows[i]
        same_mode += int(got[0, 0] == ro[0, 0])
        kg, ko = first_switch(got), first_switch(ro)
        same_first += int(kg == ko and (ko is None or
                                        abs(got[ko, 1] - ro[ko, 1]) <= 1e-12 * abs(ro[ko, 1])))
        same_count += int(np.sum(np.diff(got[:, 0]) != 0) == info["nswitch"])
        j = min(len(ro) if ko is None else ko + 1, len(got))
        d = np.linalg.norm(got[:j, 2:5] - ro[:j, 2:5], axis=1) / \
            np.linalg.norm(ro[:j, 2:5], axis=1)
        err.append(float(np.max(d)))
    err = np.array(err)
    return {
        "parity": {"particles": k, "same_initial_mode": same_mode,
                   "same_first_switch_row_and_time_1e-12": same_first,
                   "same_switch_count": same_count,
                   "max_pos_rel_err_to_first_switch": float(err.max()),
                   "frac_within_1e-07_to_first_switch": float(np.mean(err <= 1e-7)),
                   "what": "first %d particles of the timed batch, full advance(120), against "
                           "the oracle" % k},
        "cpu_baseline": {"value": steps / t_cpu, "unit": "particle-steps/s", "cores": 1,
                         "kind": "port",
                         "sample": "first %d particles, full advance(120), C oracle on one "
                                   "core (%.1f s)" % (k, t_cpu)},
    }


def _pct(a):
    a = np.asarray(a, float)
    if a.size == 0:
        return None
    return {"median": float(np.median(a)), "p99": float(np.percentile(a, 99)),
            "max": float(np.max(a))}


def run_mixed_precision(args):
    """Mixed precision: 2^20 particles, half protons, half
    electrons, EarthDipole + (0, 0, -10 nT), advance(30).  The whole
    population runs as guiding centres (output every 1 s) in FP64 and in
    FP32-field; a Particle-mode sub-batch (256 of each species; protons 30 s,
    electrons 0.02 s) gives the first-invariant comparison.  FP32-field runs
    use the I-controller (beta = 0); the comparisons are against FP64 with
    the same controller, and the FP64 controller change is reported apart."""
    import rapt_amd as rapt
    from rapt_amd import _native as N, batch as BT, workloads as W
    n = args.n or 1048576
    wl = W.mixed_dipole_uniform(n=n)
    f = rapt.fields.DipolePlusUniform(Bu=(0.0, 0.0, wl["dBz"]))
    ctx = N.context(0)
    m, q, v, pa = wl["mass"], wl["charge"], wl["v"], np.radians(wl["pa"])
    g = 1 / np.sqrt(1 - (v / rapt.c) ** 2)
    B0 = ctx.field_eval(f.descriptor(), "magB", np.c_[np.zeros(n), wl["pos"]])[:, 0]
    ppar = g * m * v * np.cos(pa)
    mu = g ** 2 * m * (v * np.sin(pa)) ** 2 / (2 * B0)
    last5 = np.c_[np.zeros(n), wl["pos"], ppar]
    dt_out = 1.0
    S = BT.gc_state(last5, m, q, mu, v, dt_out)
    cap = np.full(n, 33, np.int64)
    prm = dict(rapt.params)
    res_gc = {}
    outs = {}
    # fp32_field runs the I-controller (beta = 0); fp64_beta0 isolates the
    # precision effect, fp64 (dopri5 default beta) is the reference setting
    for name, kw in (("fp64", {}), ("fp64_beta0", {"beta": -1.0}),
                     ("fp32_field", {"precision": "fp32_field"})):
        desc = N.make_desc(N.MODE_GC, prm, 30.0, lanes_per_particle=1, **kw)
        tot, el, out = timed_run(f, N.MODE_GC, S, np.zeros(n, np.int32), cap, desc,
                                 args.steps, args.warmup, 5, rows=True)
        outs[name] = out
        res_gc[name] = {"particle_steps_per_s": tot["naccpt"] / el,
                        "kernel_ms_avg": tot["kernel_ms"] / args.steps,
                        "accepted_steps": tot["naccpt"] / args.steps,
                        "done": int(np.sum(out["status"] == 1))}
        log("GC %s: %.3g particle-steps/s" % (name, tot["naccpt"] / el))
    # FP64 on the four-lane kernel (the production FP64 path) for the speed line
    desc = N.make_desc(N.MODE_GC, prm, 30.0)
    tot, el, _ = timed_run(f, N.MODE_GC, S, np.zeros(n, np.int32), cap, desc, args.steps,
                           args.warmup, 5)
    res_gc["fp64_quad"] = {"particle_steps_per_s": tot["naccpt"] / el,
                           "kernel_ms_avg": tot["kernel_ms"] / args.steps}
    # per-species errors at the output rows
    a, b, a0 = outs["fp64_beta0"], outs["fp32_field"], outs["fp64"]
    errs = {}
    for sp, name in ((0, "proton"), (1, "electron")):
        idx = np.nonzero((wl["species"] == sp) & (a["status"] == 1) & (b["status"] == 1) &
                         (a0["status"] == 1) & (a["nrows"] == b["nrows"]) &
                         (a["nrows"] == a0["nrows"]) & (a["nrows"] > 0))[0]
        ra, owner, first = flat_rows(a, idx)
        rb, _, _ = flat_rows(b, idx)
        r0, _, _ = flat_rows(a0, idx)
        d = np.linalg.norm(ra[:, 1:4] - rb[:, 1:4], axis=1) / np.linalg.norm(ra[:, 1:4], axis=1)
        dc = np.linalg.norm(ra[:, 1:4] - r0[:, 1:4], axis=1) / np.linalg.norm(r0[:, 1:4], axis=1)
        mc = m[owner] * rapt.c
        Ba = ctx.field_eval(f.descriptor(), "magB", ra[:, :4])[:, 0]
        Bb = ctx.field_eval(f.descriptor(), "magB", rb[:, :4])[:, 0]
        ga = np.sqrt(1 + 2 * mu[owner] * Ba / (mc * rapt.c) + (ra[:, 4] / mc) ** 2)
        gb = np.sqrt(1 + 2 * mu[owner] * Bb / (mc * rapt.c) + (rb[:, 4] / mc) ** 2)
        Ka, Kb = (ga - 1) * mc * rapt.c, (gb - 1) * mc * rapt.c
        K0 = (g[owner] - 1) * mc * rapt.c
        red = lambda x: np.maximum.reduceat(x, first)
        errs[name] = {"particles": int(len(idx)),
                      "gc_position_rel_err_fp32_vs_fp64": _pct(red(d)),
                      "gc_position_rel_err_beta0_vs_default_fp64": _pct(red(dc)),
                      "gc_energy_drift_fp64": _pct(red(np.abs(Ka - K0) / K0)),
                      "gc_energy_drift_fp32": _pct(red(np.abs(Kb - K0) / K0)),
                      "gc_energy_rel_err_fp32_vs_fp64": _pct(red(np.abs(Kb - Ka) / Ka))}
    # Particle-mode sub-batch: first invariant per row as Particle.mu() defines it
    # (utils.guidingcenter then utils.magnetic_moment, on the device)
    for sp, name, delta in ((0, "proton", 30.0), (1, "electron", 0.02)):
        sel = np.nonzero(wl["species"] == sp)[0][:256]
        k = sel.size
        last7 = np.zeros((k, 7))
        last7[:, 1:4] = wl["pos"][sel]
        last7[:, 4:7] = (m[sel] * g[sel])[:, None] * wl["vel"][sel]
        mus = {}
        rows = {}
        for prec in ("fp64", "fp32_field"):
            r = BT.advance_particles(f, last7, m[sel], q[sel], delta, params=prm,
                                     precision=prec, beta=0.0)
            rows[prec] = r
            allrows = [np.vstack([last7[i:i + 1], r.rows[i]]) for i in range(k)]
            cnt = np.array([len(x) for x in allrows])
            t0d = time.perf_counter()
            gco, gst = ctx.particle_gc(f.descriptor(), np.vstack(allrows),
                                       np.repeat(m[sel], cnt), np.repeat(q[sel], cnt))
            diag_s = time.perf_counter() - t0d
            mus[prec] = np.split(gco[:, 5], np.cumsum(cnt)[:-1])
            log("%s %s: mu of %d rows on the device in %.3f s (%d not converged)"
                % (name, prec, cnt.sum(), diag_s, int(np.sum(gst < 0))))
        drift64 = [np.max(np.abs(x - x[0])) / x[0] for x in mus["fp64"]]
        drift32 = [np.max(np.abs(x - x[0])) / x[0] for x in mus["fp32_field"]]
        mu_err = [np.max(np.abs(x[:min(len(x), len(y))] - y[:min(len(x), len(y))]) /
                         y[:min(len(x), len(y))])
                  for x, y in zip(mus["fp32_field"], mus["fp64"])]
        pe = []
        for i in range(k):
            x64, x32 = rows["fp64"].rows[i], rows["fp32_field"].rows[i]
            kk = min(len(x64), len(x32))
            pe.append(np.max(np.linalg.norm(x64[:kk, 1:4] - x32[:kk, 1:4], axis=1) /
                             np.linalg.norm(x64[:kk, 1:4], axis=1)))
        errs[name].update({"particle_subbatch": {"particles": k, "delta_s": delta,
                                                 "mu_drift_fp64": _pct(drift64),
                                                 "mu_drift_fp32": _pct(drift32),
                                                 "mu_rel_err_fp32_vs_fp64": _pct(mu_err),
                                                 "position_rel_err_fp32_vs_fp64": _pct(pe)}})
    value = res_gc["fp32_field"]["particle_steps_per_s"]
    res = base_line("FP32-field guiding-centre particle-steps/s (1 GPU)", value,
                    "particle-steps/s", args, args.steps * res_gc["fp32_field"]["kernel_ms_avg"]
                    / 1e3, "f32 field / f64 state",
                    {"workload": "EarthDipole + (0,0,-10 nT) GuidingCenter batch: %d particles "
                                 "(half protons, half electrons), log10 K[eV] U[5,log10 5e6], "
                                 "L U[3,7], pitch U[10,90] deg; DOPRI5 default tolerances, "
                                 "advance(30), output every 1 s" % n, "particles": n})
    res["gc_runs"] = res_gc
    res["errors_per_species"] = errs
    if not args.no_cpu:
        res.update(mixed_oracle_leg(f, wl, last5, m, q, mu, v, prm, outs["fp64"]))
    return res


def mixed_oracle_leg(f, wl, last5, m, q, mu, v, prm, out64, k=512):
    """CPU baseline and full-length parity of the FP64 guiding-centre run of
    the mixed batch: the oracle's advance(30) of the first k protons and the
    first k electrons (C oracle, OpenMP on the host cores) against the rows
    of the same particles from the timed FP64 launch, per species."""
    from oracle import oracle as O
    fo = O.field_from_desc(f.descriptor())
    rtol, atol = prm["solvertolerances"]
    po = O.gc_params(rtol=rtol, atol=atol)
    threads = min(16, os.cpu_count() or 1)
    out = {"parity_fp64": {}}
    steps, t_cpu = 0, 0.0
    for sp, name in ((0, "proton"), (1, "electron")):
        sel = np.nonzero(wl["species"] == sp)[0][:k]
        t0 = time.perf_counter()
        rows, off, nrows, st, cnt = O.batch_gc(fo, po, 0, last5[sel], m[sel], q[sel], mu[sel],
                                               v[sel], 1.0, 30.0, np.full(sel.size, 33),
                                               nthreads=threads)
        t_cpu += time.perf_counter() - t0
        steps += cnt.naccpt
        ref = [rows[off[j]:off[j] + nrows[j]] for j in range(sel.size)]
        gpu = [out64["rows"][out64["row_off"][i]:out64["row_off"][i] + out64["nrows"][i]]
               for i in sel]

        def truth(j, sel=sel):
            i = sel[j]
            r, o, kk, _, _ = O.batch_gc(fo, O.gc_params(rtol=1e-12, atol=1e-30, nsteps=10 ** 7),
                                        0, last5[i], m[i], q[i], mu[i], v[i], 1.0, 30.0,
                                        np.full(len(i), 33), nthreads=threads)
            return [r[o[a]:o[a] + kk[a]] for a in range(len(i))]
        out["parity_fp64"][name] = parity_summary(gpu, ref, truth)
    out["parity_fp64"]["what"] = ("first %d protons and %d electrons of the timed FP64 launch, "
                                  "full advance(30), against the oracle" % (k, k))
    out["cpu_baseline"] = {"value": steps / t_cpu, "unit": "particle-steps/s", "cores": threads,
                           "kind": "port",
                           "sample": "first %d protons and %d electrons, full advance(30), C "
                                     "oracle with OpenMP (%.1f s)" % (k, k, t_cpu)}
    return out



def run_surrogate(args):
    """Learned bounce-averaged drift surrogate in EarthDipole + (0,0,-10 nT):
    protons; queries (L U[2,8], alpha_eq U[10,89] deg, log10 K[eV] U[4,7]).
    Training set 2^20 (seed 4), held-out 2^16 (seed 104), inference 2^24
    (seed 204), and a second 2^16 check set (seed 105) that no setting was
    chosen on.  Labels and fallback: traced bounce quantities on the GPU.
    value = inference queries/s including the traced fallback of flagged
    queries; compared with the traced rate."""
    import rapt_amd as rapt
    from rapt_amd import surrogate as SG
    f = rapt.fields.DipolePlusUniform(Bu=(0.0, 0.0, -10e-9))

    def qs(n, seed):
        rng = np.random.default_rng(seed)
        return np.c_[rng.uniform(2, 8, n), rng.uniform(10, 89, n), rng.uniform(4, 7, n)]

    if args.infer_only:
        # the inference kernels alone (their work does not depend on the
        # weights): an untrained ensemble, 2^24 queries, for profiling
        qin = qs(1 << 24, 204)
        s = SG.Surrogate(f, members=args.members, seed=0)
        s.infer(qin[:1 << 16])
        kms = []
        for i in range(max(args.steps, 1)):
            kms.append(s.infer(qin)[3])
            log("inference %d: %.2f ms" % (i, kms[-1]))
        s.close()
        ms = float(np.mean(kms))
        res = base_line("surrogate inference kernel time, untrained ensemble (profiling)",
                        len(qin) / (ms / 1e3), "queries/s", args, ms / 1e3 * len(kms),
                        "bf16 MFMA / fp32 accumulate",
                        {"workload": "MLP ensemble %d x 3-256-256-256-3 SiLU, 2^24 queries, "
                                     "untrained weights" % args.members, "queries": len(qin)})
        res["ms_per_step"] = ms
        res["inference_kernel_ms"] = kms
        return res
    ntr = args.n or (1 << 20)
    qtr, qte, qin = qs(ntr, 4), qs(1 << 16, 104), qs(1 << 24, 204)
    t0 = time.perf_counter()
    lab, st = SG.traced_labels(f, qtr)
    t_lab = time.perf_counter() - t0
    traced_rate = ntr / t_lab
    log("traced labels: %d in %.2f s (%.3g queries/s), ok %.4f" % (ntr, t_lab, traced_rate,
                                                                    np.mean(st == 1)))
    truth, st_te = SG.traced_labels(f, qte)
    s = SG.Surrogate(f, members=args.members, seed=0)
    t0 = time.perf_counter()
    hist = s.fit(qtr[st == 1], lab[st == 1], epochs=args.epochs, batch=4096, log=log)
    t_fit = time.perf_counter() - t0
    # held-out: calibrate the fallback threshold on the members' deviation so
    # that at least 99.9% of the final answers are within 1e-3 of the traced
    # labels (flagged queries are traced, i.e. exact)
    v_te, _, raw, _ = s.infer(qte, 1e30, raw=True)
    good = st_te == 1
    e_s = np.abs(v_te - truth) / np.abs(truth)
    sc = s.norm[3:]
    dev = np.expm1(sc * np.max(np.abs(raw - raw.mean(0)), axis=0)).max(axis=1)
    # the ensemble-disagreement threshold of the configuration (1e-3); the
    # table shows what lower thresholds would flag and gain
    thresh, calib = 1e-3, []
    for t in (1e-3, 5e-4, 2e-4, 1e-4, 5e-5, 2e-5):
        ok = np.all(np.where((dev > t)[:, None], 0.0, e_s) <= 1e-3, axis=1)[good].mean()
        calib.append({"thresh": t, "frac_le_1e-3": float(ok),
                      "flagged": float(np.mean(dev > t))})
    log("threshold table: %s" % calib)
    if args.dump:
        np.savez(args.dump, query=qte, truth=truth, status=st_te, surrogate=v_te, dev=dev)
    p_te, fl_te, _ = s.predict(qte, thresh)
    e_f = np.abs(p_te - truth) / np.abs(truth)
    # inference throughput at 2^24 queries (warmup on a slice first)
    s.infer(qin[:1 << 16], thresh)
    t0 = time.perf_counter()
    v_in, fl_in, _, kms = s.infer(qin, thresh)
    t_inf = time.perf_counter() - t0
    idx = np.nonzero(fl_in)[0]
    t0 = time.perf_counter()
    if idx.size:
        SG.traced_labels(f, qin[idx])
    t_fb = time.perf_counter() - t0
    # inputs resident in HBM: the inference kernel time, plus the traced
    # fallback of the flagged queries at the traced rate
    value = len(qin) / (kms / 1e3 + idx.size / traced_rate)
    flop = args.members * 265216.0 * len(qin)
    per_out = lambda e: {"median": [float(x) for x in np.median(e, axis=0)],
                         "p99": [float(x) for x in np.percentile(e, 99, axis=0)],
                         "max": [float(x) for x in np.max(e, axis=0)],
                         "frac_all_le_1e-3": float(np.mean(np.all(e <= 1e-3, axis=1)))}
    res = base_line("surrogate bounce-drift evals/s at <=1e-3 err (1 GPU)", value, "queries/s",
                    args, t_inf + t_fb, "bf16 MFMA / fp32 accumulate",
                    {"workload": "EarthDipole + (0,0,-10 nT), protons; MLP ensemble %d x "
                                 "3-256-256-256-3 SiLU; train 2^20 traced samples, held-out "
                                 "2^16, inference 2^24 queries, fallback when a member deviates "
                                 "> 1e-3 from the ensemble mean" % args.members,
                     "queries": len(qin)})
    res["steps"], res["warmup"] = 1, 1
    res["ms_per_step"] = (kms / 1e3 + idx.size / traced_rate) * 1e3
    res["traced_queries_per_s"] = traced_rate
    res["speedup_vs_traced"] = value / traced_rate
    res["inference_kernel_ms"] = kms
    res["inference_wall_ms"] = t_inf * 1e3
    res["fallback_queries"] = int(idx.size)
    res["fallback_ms"] = t_fb * 1e3
    res["fallback_threshold"] = thresh
    res["threshold_calibration"] = calib
    res["value_wall_incl_copies"] = len(qin) / (t_inf + t_fb)
    res["train_s"] = t_fit
    res["train_epochs"] = args.epochs
    res["final_loss"] = [float(x) for x in hist[-1]]
    res["heldout_surrogate_rel_err"] = per_out(e_s[good])
    res["heldout_with_fallback_rel_err"] = per_out(e_f[good])
    res["heldout_flagged_frac"] = float(np.mean(fl_te))
    res["roofline"] = {"bound": "mfma", "achieved": flop / (kms / 1e3) / 1e12,
                       "peak": 2500.0, "unit": "TFLOP/s",
                       "frac": flop / (kms / 1e3) / 2.5e15, "traffic": None,
                       "flop_per_query": args.members * 265216}
    # an independent check set (seed 105) that chose nothing: neither the
    # threshold nor the member count and epochs
    qva = qs(1 << 16, 105)
    truth_va, st_va = SG.traced_labels(f, qva)
    p_va, fl_va, _ = s.predict(qva, thresh)
    e_va = np.abs(p_va - truth_va) / np.abs(truth_va)
    res["check_set_with_fallback_rel_err"] = per_out(e_va[st_va == 1])
    res["check_set_flagged_frac"] = float(np.mean(fl_va))
    s.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="gc_doubledipole",
                    choices=("gc_doubledipole", "particle_dipole", "adaptive_sheet",
                             "mixed_precision", "surrogate", "grid", "user_field"))
    ap.add_argument("--epochs", type=int, default=200, help="surrogate training epochs")
    ap.add_argument("--members", type=int, default=2, choices=(2, 3, 4),
                    help="surrogate: ensemble members")
    ap.add_argument("--dump", default="", help="surrogate: write held-out results to this .npz")
    ap.add_argument("--infer-only", action="store_true",
                    help="surrogate: time only the inference kernels (untrained ensemble)")
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=0)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--mixed-quads", type=int, default=0,
                    help="with --lpp 3: particles traced on lane quads (0: n/8)")
    ap.add_argument("--mixed-octets", type=int, default=0,
                    help="with --lpp 5: particles traced on lane octets (0: n/64)")
    ap.add_argument("--lpp", type=int, default=4, choices=(1, 2, 3, 4, 5, 8),
                    help="lanes per guiding centre (4: a lane quad, 2: a lane pair per "
                         "particle); with --no-tune")
    ap.add_argument("--block", type=int, default=0,
                    help="threads per block (0: library default); with --no-tune")
    ap.add_argument("--top", type=int, default=0,
                    help="trace only this many of the most expensive particles (critical "
                         "path of the launch); implies --no-cpu")
    ap.add_argument("--skip-top", type=int, default=0,
                    help="trace all but this many of the most expensive particles (the "
                         "batch without its critical path); implies --no-cpu")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-shape autotuner (use --lpp/--lanes as given)")
    ap.add_argument("--cpu-sample", type=int, default=8192)
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the other configurations' summaries in the default line")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--order", choices=("cost", "dt", "input"), default="cost",
                    help="library work order: descending cost measured by a trial launch "
                         "over half the advance (default), shortest output step first, "
                         "or input order")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="only run a device copy of this many bytes (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the tracer shards particles with no exchange, "
              "and this build runs and reports a single MI355X only (--gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        a = torch.empty(args.copy_probe // 4, dtype=torch.float32, device="cuda")
        a.fill_(1.0)
        bb = torch.empty_like(a)
        torch.cuda.synchronize()
        bb.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": args.copy_probe}))
        return

    runners = {"particle_dipole": run_particle_dipole, "adaptive_sheet": run_adaptive_sheet,
               "grid": run_grid, "mixed_precision": run_mixed_precision,
               "surrogate": run_surrogate, "gc_doubledipole": run_gc_doubledipole,
               "user_field": run_user_field}
    res = runners[args.workload](args)
    if args.workload == "gc_doubledipole" and not args.headline_only and not args.top \
            and not args.skip_top:
        res["configs"] = other_configs(args)
    print(json.dumps(res), flush=True)


def other_configs(args):
    """Short runs of the other configurations for the same JSON line (each
    its own seeded workload at full size; a failure is recorded, not raised)."""
    import copy
    out = {}
    for name, fn, keys in (
            ("particle_dipole", run_particle_dipole,
             ("value", "ms_per_step", "kernel_ms_avg", "accepted_steps_per_step", "roofline",
              "parity", "mu_drift", "cpu_baseline", "gpu_vs_cpu", "particles_done", "config")),
            ("adaptive_sheet", run_adaptive_sheet,
             ("value", "ms_per_step", "kernel_ms_avg", "accepted_steps_per_step",
              "switches_per_step", "status_counts", "parity", "cpu_baseline", "config")),
            ("mixed_precision", run_mixed_precision,
             ("value", "gc_runs", "errors_per_species", "parity_fp64", "cpu_baseline",
              "config")),
            ("user_field", run_user_field,
             ("value", "kernel_ms_avg", "particle_steps_per_s", "gc_doubledipole",
              "first_compiled_run_s",
              "compiled_equals_interpreted", "final_position_rel_diff_vs_builtin", "roofline",
              "config")),
            ("surrogate", run_surrogate,
             ("value", "unit", "speedup_vs_traced", "traced_queries_per_s", "inference_kernel_ms",
              "fallback_queries", "fallback_ms", "heldout_surrogate_rel_err",
              "heldout_with_fallback_rel_err", "heldout_flagged_frac", "check_set_with_fallback_rel_err",
              "check_set_flagged_frac", "roofline", "train_s",
              "config"))):
        a = copy.copy(args)
        a.n, a.lanes, a.block = 0, 0, 0
        a.steps, a.warmup = min(args.steps, 3), 1
        t0 = time.perf_counter()
        try:
            r = fn(a)
            out[name] = {k: r[k] for k in keys if k in r}
            out[name]["metric"] = r.get("metric")
        except Exception as e:  # keep the headline line whatever an extra does
            out[name] = {"error": "%s: %s" % (type(e).__name__, e)}
        out[name]["wall_s"] = time.perf_counter() - t0
        log("%s: %.1f s" % (name, out[name]["wall_s"]))
    return out


def run_gc_doubledipole(args):
    """The headline: the DoubleDipole guiding-centre batch (module docstring)."""
    args.n = args.n or 65536
    from rapt_amd import _native as N
    w = build_workload(args.n)
    n = args.n
    ctx = N.context(0)
    ctx.set_field(w["field"].descriptor())
    from rapt_amd import batch as BT
    S = BT.gc_state(w["last5"], w["m"], w["q"], w["mu"], w["v"], w["dt"])
    flags = np.zeros(n, np.int32)
    cap = BT._rows_needed(w["delta"], w["dt"])
    desc = N.make_desc(N.MODE_GC, w["params"], w["delta"], lanes=args.lanes, block=args.block,
                       lanes_per_particle=args.lpp, work_order=1 if args.order == "input" else 0,
                       mixed_quads=args.mixed_quads, mixed_octets=args.mixed_octets)
    if args.top or args.skip_top:
        # only the `top` most expensive particles (accepted steps of a full-advance
        # trial): the launch's critical path without the rest of the batch; or
        # all but the `skip_top` most expensive: the rest without its critical path
        from rapt_amd import autotune as AT
        full = AT.trial_cost(w["field"], S, flags, cap, desc, trial_fraction=1.0)
        rank = np.argsort(-full, kind="stable")
        idx = rank[:args.top] if args.top else rank[args.skip_top:]
        S, flags, cap = S[:, idx], flags[idx], cap[idx]
        n = len(idx)
        args.no_cpu = True
        log("%d particles: %.0f-%.0f accepted steps each" % (n, full[idx].min(),
                                                             full[idx].max()))
    cost = None
    cost_s = None
    if args.order == "cost":
        from rapt_amd import autotune as AT
        t0 = time.perf_counter()
        cost = AT.trial_cost(w["field"], S, flags, cap, desc, trial_fraction=0.5)
        cost_s = time.perf_counter() - t0
        log("work order: descending accepted steps of a trial launch over half the "
            "advance (mean %.0f per particle, %.2f s)" % (cost.mean(), cost_s))
    tuned = None
    if not args.no_tune:
        from rapt_amd import autotune as AT
        # trials over the whole advance: the mixed layouts differ in how the
        # longest traces finish, which a shorter trial does not reach
        shape, table = AT.tune(w["field"], S, flags, cap, desc, trial_fraction=1.0, log=log,
                               cost=cost)
        AT.apply(desc, shape)
        tuned = {"chosen": shape, "trials": [{"shape": t, "steps_per_s": r} for t, r in table],
                 "trial_fraction_of_delta": 1.0, "trial_work_order": args.order,
                 "cost_trial_fraction_of_delta": 0.5 if args.order == "cost" else None,
                 "trial_repeats": 2}
        log("autotune chose %s" % shape)
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")

    def timed(b, steps, warmup, label):
        for i in range(warmup):
            b.reset()
            c = b.run(desc)
            log("%s warmup %d: %.1f ms, %d accepted steps" % (label, i, c.kernel_ms, c.naccpt))
        hip.hipDeviceSynchronize()
        tot = dict(naccpt=0, nrejct=0, nstep=0, nfcn=0, kernel_ms=0.0)
        t0 = time.perf_counter()
        for i in range(steps):
            b.reset()
            c = b.run(desc)
            for k in ("naccpt", "nrejct", "nstep", "nfcn"):
                tot[k] += getattr(c, k)
            tot["kernel_ms"] += c.kernel_ms
            log("%s step %d: %.1f ms, %d accepted steps" % (label, i, c.kernel_ms, c.naccpt))
        hip.hipDeviceSynchronize()
        return tot, time.perf_counter() - t0

    b = N.Batch(ctx, n)
    b.upload(S, flags)
    if cost is not None:
        b.set_cost(cost)
    b.set_rows(5, cap)
    log("rows allocated: %d (%.2f GB)" % (cap.sum(), cap.sum() * 40 / 1e9))
    tot, el = timed(b, args.steps, args.warmup, "timed")
    # rows of the last timed advance delivered to the host (the reference's
    # advance hands back host trajectories): timed apart from the device rate
    t0 = time.perf_counter()
    out = b.download(rows=not args.no_cpu)
    dl_s = time.perf_counter() - t0
    out["row_off"] = np.asarray(b.row_off, np.int64).copy()
    ok = int(np.sum(out["status"] == 1))
    q = np.percentile(out["naccpt"], [50, 90, 99, 100])
    log("accepted steps per particle: mean %.0f, p50 %.0f, p90 %.0f, p99 %.0f, max %.0f"
        % (out["naccpt"].mean(), *q))
    # the library's default work order (shortest output step first), same shape
    dflt = None
    if args.order == "cost" and not args.top and not args.skip_top:
        b2 = N.Batch(ctx, n)
        try:
            b2.upload(S, flags)
            b2.set_rows(5, cap)
            t2, e2 = timed(b2, min(args.steps, 5), 1, "output-step order")
            dflt = {"value": t2["naccpt"] / e2, "ms_per_step": e2 / min(args.steps, 5) * 1e3,
                    "kernel_ms_avg": t2["kernel_ms"] / min(args.steps, 5)}
        finally:
            b2.close()
    value = tot["naccpt"] / el
    launches = args.steps
    flop_launch = (tot["nfcn"] * FLOP_RHS_TCB_DOUBLEDIPOLE +
                   tot["nstep"] * FLOP_DOPRI5_STEP_N4) / launches
    avg_s = tot["kernel_ms"] / launches / 1e3
    achieved = flop_launch / avg_s
    rows_launch = float(out["nrows"].sum())
    res = {
        "metric": "FP64 guiding-centre particle-steps/s (1 GPU)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded, seed 1)",
        "config": {
            "workload": "DoubleDipole GuidingCenter batch: %d electrons, log10 K[eV] U[4,7], "
                        "L U[3,8], eq. pitch U[15,85] deg; TCB EOM, DOPRI5 rtol=atol=1e-8, "
                        "advance(60), dt=bounceperiod/10" % n,
            "particles": n, "delta_s": 60.0, "rtol": 1e-8, "atol": 1e-8,
            "lanes": desc.lanes or n * max(desc.lanes_per_particle, 1),
            "lanes_per_particle": desc.lanes_per_particle or 4, "block": desc.block or 256,
            "order": args.order,
            "top": args.top or None,
        },
        "particles_done": ok,
        "autotune": tuned,
        "accepted_steps_per_step": tot["naccpt"] / launches,
        "rhs_evals_per_step": tot["nfcn"] / launches,
        "rejected_steps_per_step": tot["nrejct"] / launches,
        "kernel_ms_avg": tot["kernel_ms"] / launches,
        "cost_trial_s": cost_s,
        "value_output_step_order": dflt,
        "row_delivery": {"rows": rows_launch, "bytes": rows_launch * ROW_BYTES_GC,
                         "download_s": dl_s if not args.no_cpu else None,
                         "value_incl_download": (tot["naccpt"] / launches) /
                         (el / launches + dl_s) if not args.no_cpu else None},
        "roofline": {
            "bound": "valu_fp64", "compute": "FP64 vector ALU (78.6 TFLOP/s)",
            "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64, "traffic": PROFILED_TRAFFIC.get(n),
            "flop_per_launch": flop_launch,
            "flop_per_particle_step": flop_launch / (tot["naccpt"] / launches),
            "reference_formulation_flop_per_launch":
                (tot["nfcn"] * FLOP_RHS_TCB_REFERENCE + tot["nstep"] * FLOP_DOPRI5_STEP_N4)
                / launches,
            "trajectory_write_GBps": rows_launch * ROW_BYTES_GC / avg_s / 1e9,
        },
    }
    b.close()
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        sample = min(args.cpu_sample, n)
        steps_cpu, el_cpu, (rows_o, off_o, nrows_o, st_o) = cpu_baseline(w, sample, threads)
        res["cpu_baseline"] = {
            "value": steps_cpu / el_cpu, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": "first %d particles of the same workload, full advance(60), C oracle "
                      "with OpenMP (%.1f s)" % (sample, el_cpu),
        }
        res["gpu_vs_cpu"] = value / (steps_cpu / el_cpu)
        from oracle import oracle as O
        fo = O.field_from_desc(w["field"].descriptor())
        ref = [rows_o[off_o[i]:off_o[i] + nrows_o[i]] for i in range(sample)]
        gpu = [out["rows"][out["row_off"][i]:out["row_off"][i] + out["nrows"][i]]
               for i in range(sample)]

        def truth(sel):
            c = (np.ceil(w["delta"] / w["dt"][sel]) + 3).astype(np.int64)
            r, o, k, st, _ = O.batch_gc(fo, O.gc_params(rtol=1e-12, atol=1e-30, nsteps=10 ** 7),
                                        0, w["last5"][sel], w["m"], w["q"], w["mu"][sel],
                                        w["v"][sel], w["dt"][sel], w["delta"], c,
                                        nthreads=threads)
            return [r[o[j]:o[j] + k[j]] for j in range(len(sel))]
        t0 = time.perf_counter()
        res["parity"] = parity_summary(gpu, ref, truth)
        res["parity"]["what"] = ("rows of the last timed advance against the oracle's, the "
                                 "CPU baseline's %d particles, full advance(60)" % sample)
        log("parity: %.1f s" % (time.perf_counter() - t0))
    return res


if __name__ == "__main__":
    main()
