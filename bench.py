"""Throughput benchmark: FP64 guiding-centre particle-steps/s on one MI355X.

Workload (the batched GuidingCenter configuration): 65,536 electrons in the
DoubleDipole field, log10 K[eV] ~ U[4,7], L ~ U[3,8], azimuth ~ U[0,2pi),
equatorial pitch angle ~ U[15,85] deg (seed 1); GuidingCenter.advance(60)
with the Tao-Chan-Brizard equations, DOPRI5 at rtol = atol = 1e-8, output
step bounceperiod()/bounceresolution per particle.  One bench step is one
full advance(60) of the whole batch; the state is restored on the device
between steps, so inputs stay resident in HBM.  A particle-step is one
accepted integrator step of one particle.

Prints one JSON line.  With --gpus > 1 it exits: this build measures one GPU.
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
PEAK_FP64 = 78.6e12                  # MI355X FP64 vector (and matrix) peak, FLOP/s
PEAK_HBM = 8.0e12                    # bytes/s
ROW_BYTES_GC = 40                    # one (t, x, y, z, ppar) output row
# HBM bytes per k_trace launch at the default batch, from rocprofv3 FETCH_SIZE
# and WRITE_SIZE (KiB) scaled by a 256 MiB copy probe (read x2.00, write x1.00);
# profiles/k_trace_gc_doubledipole_65536.md has the commands and raw values
PROFILED_TRAFFIC = {65536: (5172.6875 * 2.0 + 3057523.5 * 1.0) * 1024}


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
    the first `sample` particles of the same workload."""
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
    return cnt.naccpt, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=65536)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--lpp", type=int, default=4, choices=(1, 4),
                    help="lanes per guiding centre (4: one lane quad per particle)")
    ap.add_argument("--cpu-sample", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--order", choices=("input", "dt"), default="dt",
                    help="library work order: shortest output step (bounce period) first "
                         "(default), or input order")
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

    from rapt_amd import _native as N
    w = build_workload(args.n)
    n = args.n
    ctx = N.context(0)
    ctx.set_field(w["field"].descriptor())
    from rapt_amd import batch as BT
    S = BT.gc_state(w["last5"], w["m"], w["q"], w["mu"], w["v"], w["dt"])
    flags = np.zeros(n, np.int32)
    cap = BT._rows_needed(w["delta"], w["dt"])
    desc = N.make_desc(N.MODE_GC, w["params"], w["delta"], lanes=args.lanes,
                       lanes_per_particle=args.lpp, work_order=0 if args.order == "dt" else 1)
    b = N.Batch(ctx, n)
    b.upload(S, flags)
    b.set_rows(5, cap)
    log("rows allocated: %d (%.2f GB)" % (cap.sum(), cap.sum() * 40 / 1e9))

    for i in range(args.warmup):
        b.reset()
        c = b.run(desc)
        log("warmup %d: %.1f ms, %d accepted steps" % (i, c.kernel_ms, c.naccpt))
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipDeviceSynchronize()
    tot = dict(naccpt=0, nrejct=0, nstep=0, nfcn=0, kernel_ms=0.0)
    t0 = time.perf_counter()
    for i in range(args.steps):
        b.reset()
        c = b.run(desc)
        for k in ("naccpt", "nrejct", "nstep", "nfcn"):
            tot[k] += getattr(c, k)
        tot["kernel_ms"] += c.kernel_ms
        log("step %d: %.1f ms, %d accepted steps" % (i, c.kernel_ms, c.naccpt))
    hip.hipDeviceSynchronize()
    el = time.perf_counter() - t0
    out = b.download(rows=False)
    ok = int(np.sum(out["status"] == 1))
    q = np.percentile(out["naccpt"], [50, 90, 99, 100])
    log("accepted steps per particle: mean %.0f, p50 %.0f, p90 %.0f, p99 %.0f, max %.0f"
        % (out["naccpt"].mean(), *q))
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
            "lanes": args.lanes or n * args.lpp, "lanes_per_particle": args.lpp,
            "order": args.order,
        },
        "particles_done": ok,
        "accepted_steps_per_step": tot["naccpt"] / launches,
        "rhs_evals_per_step": tot["nfcn"] / launches,
        "rejected_steps_per_step": tot["nrejct"] / launches,
        "kernel_ms_avg": tot["kernel_ms"] / launches,
        "roofline": {
            "bound": "mfma", "compute": "FP64 vector ALU (78.6 TFLOP/s, same as FP64 matrix)",
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
        steps_cpu, el_cpu = cpu_baseline(w, min(args.cpu_sample, n), threads)
        res["cpu_baseline"] = {
            "value": steps_cpu / el_cpu, "unit": "particle-steps/s", "cores": threads,
            "kind": "port",
            "sample": "first %d particles of the same workload, full advance(60), C oracle "
                      "with OpenMP (%.1f s)" % (min(args.cpu_sample, n), el_cpu),
        }
        res["gpu_vs_cpu"] = value / (steps_cpu / el_cpu)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
