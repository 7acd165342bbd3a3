"""ctypes binding of librapt_hip.so (include/rapt_hip.h).

The library is required: importing the tracers works without it, but the
first call that needs the GPU raises RuntimeError if the library is missing or
no HIP device is visible.  There is no CPU integration path.
"""
import ctypes as C
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# RAPT_HIP_LIB names another build of the same library (e.g. one compiled
# with different launch bounds)
LIB_PATH = os.environ.get("RAPT_HIP_LIB") or os.path.join(_HERE, "lib", "librapt_hip.so")

MODE_PARTICLE, MODE_GC, MODE_ADAPTIVE = 0, 1, 2
EOM = {"TaoChanBrizardEOM": 0, "BrizardChanEOM": 1, "NorthropTellerEOM": 2}
STEP_DOP853, STEP_DOPRI5, STEP_RKF45 = 0, 1, 2
PREC_FP64, PREC_FP32_FIELD = 0, 1
EXPR_INTERPRETED, EXPR_COMPILED = 0, 1
NSLOT = 16
S_T, S_Y0, S_MASS, S_CHARGE, S_MU, S_V, S_DT, S_H, S_TEND, S_GM, S_TCUR = \
    0, 1, 7, 8, 9, 10, 11, 12, 13, 14, 15
FLAG_STARTED, FLAG_GCMODE, FLAG_AINIT, FLAG_FINAL = 1, 2, 4, 8
ST_PENDING, ST_DONE, ST_SWITCH = 0, 1, 2
ST_NSTEPS, ST_HSMALL, ST_NAN, ST_GCITER, ST_CAP, ST_HMIN = -2, -3, -4, -5, -6, -7
GC_NOCONV = -30  # rapt_particle_gc: fixed-point iteration did not reach tol
STATUS_TEXT = {1: "ok", 2: "adiabaticity switch", -2: "larger nsteps is needed",
               -3: "step size becomes too small", -4: "non-finite state",
               -5: "guiding-centre iteration did not converge", -6: "row capacity reached",
               -7: "step below hmin", -10: "bad configuration"}
EVAL = dict(B=0, E=1, unitb=2, magB=3, gradB=4, jacobianB=5, curvature=6, curlb=7,
            dBdt=8, dbdt=9, lengthscale=10, timescale=11)
EVAL_WIDTH = dict(B=3, E=3, unitb=3, magB=1, gradB=3, jacobianB=9, curvature=1, curlb=3,
                  dBdt=1, dbdt=3, lengthscale=1, timescale=1)
ROW_WIDTH = {MODE_PARTICLE: 7, MODE_GC: 5, MODE_ADAPTIVE: 8}


class FieldDesc(C.Structure):
    _fields_ = [("model", C.c_int32), ("is_static", C.c_int32), ("gstep", C.c_double),
                ("tstep", C.c_double), ("p", C.c_double * 16)]


class TraceDesc(C.Structure):
    _fields_ = [("mode", C.c_int32), ("eom", C.c_int32), ("stepper", C.c_int32),
                ("precision", C.c_int32), ("nsteps", C.c_int32),
                ("enforce_equatorial", C.c_int32), ("check_adiabaticity", C.c_int32),
                ("block", C.c_int32), ("rtol", C.c_double), ("atol", C.c_double),
                ("safety", C.c_double), ("dfactor", C.c_double), ("ifactor", C.c_double),
                ("beta", C.c_double), ("max_step", C.c_double), ("first_step", C.c_double),
                ("gc_rtol", C.c_double), ("gc_atol", C.c_double), ("rkf_tol", C.c_double),
                ("rkf_hmax", C.c_double), ("rkf_hmin", C.c_double),
                ("cyclotronresolution", C.c_double), ("gctimestep", C.c_double),
                ("epss", C.c_double), ("epst", C.c_double), ("delta", C.c_double),
                ("lanes", C.c_int64), ("lanes_per_particle", C.c_int32),
                ("work_order", C.c_int32), ("mixed_quads", C.c_int64),
                ("mixed_octets", C.c_int64)]


class Counters(C.Structure):
    _fields_ = [("naccpt", C.c_int64), ("nrejct", C.c_int64), ("nstep", C.c_int64),
                ("nfcn", C.c_int64), ("nswitch", C.c_int64), ("nrows", C.c_int64),
                ("kernel_ms", C.c_double)]


_lib = None
_lock = threading.Lock()
_ctx = {}


def lib():
    """Load librapt_hip.so (raises RuntimeError when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("librapt_hip.so is not built (%s); run __graft_entry__.build() or "
                           "make -C rapt_amd/csrc" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    vp, dp, i32p, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int32), \
        C.POINTER(C.c_int64)
    L.rapt_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    L.rapt_ctx_destroy.argtypes = [vp]
    L.rapt_last_error.argtypes = [vp]
    L.rapt_last_error.restype = C.c_char_p
    L.rapt_field_set.argtypes = [vp, C.POINTER(FieldDesc)]
    L.rapt_field_eval.argtypes = [vp, C.c_int, C.c_int64, dp, dp]
    L.rapt_halfbounce.argtypes = [vp, C.c_int64, dp, dp, C.c_double, dp, i32p]
    L.rapt_bounce_labels.argtypes = [vp, C.c_int64, dp, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, dp, i32p]
    L.rapt_batch_epochs_begin.argtypes = [vp, C.POINTER(TraceDesc), C.c_int32]
    L.rapt_batch_epoch_next.argtypes = [vp, C.POINTER(dp), C.POINTER(i64p), C.POINTER(i32p),
                                        C.POINTER(C.c_int32), C.POINTER(Counters)]
    L.rapt_batch_epochs_end.argtypes = [vp]
    L.rapt_bounce_center_advance.argtypes = [vp, C.c_int64, dp, dp, dp, dp, dp, i64p, dp, i32p,
                                             i64p, C.c_double, C.c_double, C.c_double,
                                             C.c_double, C.c_int32]
    L.rapt_bounce_center_equatorial.argtypes = [vp, C.c_int64, dp, dp, dp, dp, i64p, dp, i32p,
                                                i64p, C.c_double, C.c_double, C.c_int32]
    fp = C.POINTER(C.c_float)
    L.rapt_mlp_create.argtypes = [vp, C.c_int32, fp, C.c_double, C.c_double, C.POINTER(vp)]
    L.rapt_mlp_destroy.argtypes = [vp]
    L.rapt_mlp_set_norm.argtypes = [vp, fp]
    L.rapt_mlp_set_split.argtypes = [vp, C.c_double, C.c_double]
    L.rapt_mlp_targets.argtypes = [vp, C.c_int64, fp, dp, C.c_int, fp, fp]
    L.rapt_mlp_train.argtypes = [vp, C.c_int64, fp, fp, C.c_int32, C.c_int32, i64p,
                                 C.c_float, C.c_float, dp, fp]
    L.rapt_mlp_get_params.argtypes = [vp, fp]
    L.rapt_mlp_forward.argtypes = [vp, C.c_int64, fp, fp]
    L.rapt_mlp_infer.argtypes = [vp, C.c_int64, fp, C.c_float, fp, i32p, fp, dp]
    L.rapt_batch_create.argtypes = [vp, C.c_int64, C.POINTER(vp)]
    L.rapt_batch_destroy.argtypes = [vp]
    L.rapt_batch_upload.argtypes = [vp, dp, i32p]
    L.rapt_batch_set_rows.argtypes = [vp, C.c_int32, i64p]
    L.rapt_batch_reset.argtypes = [vp]
    L.rapt_batch_set_cost.argtypes = [vp, dp]
    L.rapt_grid_set.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, dp, dp, dp, dp,
                                dp, dp]
    L.rapt_field_expr_set.argtypes = [vp, C.c_int32, C.c_int32, i32p, C.c_int32, dp]
    L.rapt_field_expr_mode.argtypes = [vp, C.c_int32]
    L.rapt_rhs_eval.argtypes = [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int64, dp, dp, dp]
    L.rapt_field_expr_compile.argtypes = [C.c_int32, C.c_int32, i32p, C.c_char_p, C.c_int32,
                                          C.c_char_p, C.c_int64]
    L.rapt_eye.argtypes = [vp, C.c_int64, dp, dp, C.c_double, dp, i32p]
    L.rapt_particle_gc.argtypes = [vp, C.c_int64, dp, dp, dp, C.c_int32, C.c_double, C.c_int32,
                                   dp, i32p]
    L.rapt_batch_run.argtypes = [vp, C.POINTER(TraceDesc), C.POINTER(Counters)]
    L.rapt_batch_download.argtypes = [vp, dp, i32p, i32p, i64p, dp, i64p]
    L.rapt_trace.argtypes = [vp, C.POINTER(TraceDesc), C.c_int64, dp, i32p, i64p, dp, i64p,
                             i32p, C.POINTER(Counters)]
    # device pointers (as integers) for the allocation-free entry point
    L.rapt_trace_device_noalloc.argtypes = [vp, C.POINTER(TraceDesc), C.c_int64, vp, vp, vp,
                                            vp, vp, vp, vp, vp, vp, C.POINTER(Counters)]
    for nm in ("rapt_ctx_create", "rapt_ctx_destroy", "rapt_field_set", "rapt_field_eval",
               "rapt_halfbounce", "rapt_bounce_labels", "rapt_bounce_center_advance",
               "rapt_bounce_center_equatorial",
               "rapt_batch_epochs_begin", "rapt_batch_epoch_next", "rapt_batch_epochs_end",
               "rapt_mlp_create", "rapt_mlp_destroy",
               "rapt_mlp_set_norm", "rapt_mlp_set_split", "rapt_mlp_targets", "rapt_mlp_train", "rapt_mlp_get_params",
               "rapt_mlp_forward", "rapt_mlp_infer",
               "rapt_batch_create", "rapt_batch_destroy", "rapt_batch_upload",
               "rapt_batch_set_rows", "rapt_batch_run", "rapt_batch_download", "rapt_trace",
               "rapt_batch_reset", "rapt_batch_set_cost", "rapt_eye", "rapt_particle_gc",
               "rapt_grid_set", "rapt_field_expr_set", "rapt_trace_device_noalloc",
               "rapt_field_expr_mode", "rapt_field_expr_compile", "rapt_rhs_eval"):
        getattr(L, nm).restype = C.c_int
    _lib = L
    return L


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _ip32(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _ip64(a):
    return a.ctypes.data_as(C.POINTER(C.c_int64))


class Context:
    """One library context (HIP stream + events) on one device."""

    def __init__(self, device=0):
        L = lib()
        h = C.c_void_p()
        rc = L.rapt_ctx_create(device, C.byref(h))
        if rc != 0:
            raise RuntimeError("rapt_ctx_create(device=%d) failed with %d: no usable HIP "
                               "device" % (device, rc))
        self.h = h
        self.device = device
        self._field_key = None
        self._grid_id = 0
        self._expr_key = None

    def check(self, rc, what):
        if rc != 0:
            msg = lib().rapt_last_error(self.h)
            raise RuntimeError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))

    def set_field(self, desc):
        grid = desc.get("grid")
        prog = desc.get("expr")
        key = (desc["model"], desc["static"], desc["gstep"], desc["tstep"], desc["params"],
               grid._grid_id if grid is not None else 0,
               prog.key() if prog is not None else None)
        if key == self._field_key:
            return
        if prog is not None and prog.key() != self._expr_key:
            code = np.ascontiguousarray(prog.code, dtype=np.int32)
            consts = np.ascontiguousarray(prog.consts, dtype=np.float64)
            self.check(lib().rapt_field_expr_set(
                self.h, prog.nb, prog.ne, code.ctypes.data_as(C.POINTER(C.c_int32)),
                len(consts), _dp(consts) if len(consts) else None), "rapt_field_expr_set")
            self._expr_key = prog.key()
        if grid is not None and self._grid_id != grid._grid_id:
            x, y, z, t, B, E = (np.ascontiguousarray(a, dtype=np.float64)
                                for a in grid.grid_arrays())
            nt = B.shape[0]
            # an all-zero E is not uploaded: the device then skips E (E = 0)
            self.check(lib().rapt_grid_set(self.h, len(x), len(y), len(z), nt, _dp(x), _dp(y),
                                           _dp(z), _dp(t), _dp(B),
                                           _dp(E) if np.any(E) else None), "rapt_grid_set")
            self._grid_id = grid._grid_id
        f = FieldDesc()
        f.model = desc["model"]
        f.is_static = 1 if desc["static"] else 0
        f.gstep = desc["gstep"]
        f.tstep = desc["tstep"]
        for i, v in enumerate(desc["params"]):
            f.p[i] = v
        self.check(lib().rapt_field_set(self.h, C.byref(f)), "rapt_field_set")
        self._field_key = key

    def set_expr_mode(self, compiled):
        """How tracer runs evaluate a user-written field: compiled at run time
        (True, the default) or interpreted (False); the trajectories are the
        same bit for bit."""
        self.check(lib().rapt_field_expr_mode(self.h, EXPR_COMPILED if compiled else
                                              EXPR_INTERPRETED), "rapt_field_expr_mode")

    def rhs_eval(self, desc, mode, ty, k, eom=0, enforce_equatorial=False):
        """The tracers' right-hand sides at the rows ty ([n][7] Particle,
        [n][5] GuidingCenter) with per-row constants k ([n][3] (mass, charge,
        gamma m) or [n][4] (mass, charge, mu, v))."""
        self.set_field(desc)
        wi, wk, wo = (7, 3, 6) if mode == MODE_PARTICLE else (5, 4, 4)
        ty = np.ascontiguousarray(np.atleast_2d(ty), dtype=np.float64)
        n = ty.shape[0]
        k = np.ascontiguousarray(np.broadcast_to(np.atleast_2d(k), (n, wk)), dtype=np.float64)
        assert ty.shape == (n, wi)
        out = np.zeros((n, wo))
        e = EOM[eom] if isinstance(eom, str) else int(eom)
        self.check(lib().rapt_rhs_eval(self.h, mode, e, 1 if enforce_equatorial else 0, n,
                                       _dp(ty), _dp(k), _dp(out)), "rapt_rhs_eval")
        return out

    def field_eval(self, desc, what, tpos):
        self.set_field(desc)
        tpos = np.ascontiguousarray(np.atleast_2d(tpos), dtype=np.float64)
        n = tpos.shape[0]
        w = EVAL_WIDTH[what]
        out = np.zeros((n, w))
        self.check(lib().rapt_field_eval(self.h, EVAL[what], n, _dp(tpos), _dp(out)),
                   "rapt_field_eval")
        return out


    def halfbounce(self, desc, tpos, Bm, fieldlineresolution):
        """S_b per particle on the device; returns (sb, status)."""
        self.set_field(desc)
        tpos = np.ascontiguousarray(np.atleast_2d(tpos), dtype=np.float64)
        n = tpos.shape[0]
        Bm = np.ascontiguousarray(np.broadcast_to(Bm, (n,)), dtype=np.float64)
        sb = np.zeros(n)
        st = np.zeros(n, np.int32)
        self.check(lib().rapt_halfbounce(self.h, n, _dp(tpos), _dp(Bm),
                                         float(fieldlineresolution), _dp(sb), _ip32(st)),
                   "rapt_halfbounce")
        return sb, st


    def eye(self, desc, tpos, Bm, fieldlineresolution):
        """Second invariant I per guiding centre on the device; returns (I, status)."""
        self.set_field(desc)
        tpos = np.ascontiguousarray(np.atleast_2d(tpos), dtype=np.float64)
        n = tpos.shape[0]
        Bm = np.ascontiguousarray(np.broadcast_to(Bm, (n,)), dtype=np.float64)
        out = np.zeros(n)
        st = np.zeros(n, np.int32)
        self.check(lib().rapt_eye(self.h, n, _dp(tpos), _dp(Bm), float(fieldlineresolution),
                                  _dp(out), _ip32(st)), "rapt_eye")
        return out, st

    def particle_gc(self, desc, rows, mass, charge, tol=1e-3, maxiter=20):
        """Guiding centre and mu of every Particle row (t, x, y, z, px, py, pz) on the
        device; mass and charge scalars or per row.  Returns (out (n, 6) = X, Y, Z,
        v_par, speed, mu; status = iterations used, or GC_NOCONV)."""
        self.set_field(desc)
        rows = np.ascontiguousarray(np.atleast_2d(rows), dtype=np.float64)
        n = rows.shape[0]
        assert rows.shape[1] == 7
        per_row = np.ndim(mass) > 0 or np.ndim(charge) > 0
        m = np.ascontiguousarray(np.broadcast_to(mass, (n,)) if per_row else [mass],
                                 dtype=np.float64)
        q = np.ascontiguousarray(np.broadcast_to(charge, (n,)) if per_row else [charge],
                                 dtype=np.float64)
        out = np.zeros((n, 6))
        st = np.zeros(n, np.int32)
        self.check(lib().rapt_particle_gc(self.h, n, _dp(rows), _dp(m), _dp(q), int(per_row),
                                          float(tol), int(maxiter), _dp(out), _ip32(st)),
                   "rapt_particle_gc")
        return out, st

    def bounce_labels(self, desc, query, mass, charge, eyegradientstep, fieldlineresolution):
        """(T_b, omega_d, mirror latitude) per (L, alpha_eq [deg], log10 K [eV]) query;
        returns (out (n, 3), status)."""
        from .constants import Re, e
        self.set_field(desc)
        query = np.ascontiguousarray(np.atleast_2d(query), dtype=np.float64)
        n = query.shape[0]
        out = np.zeros((n, 3))
        st = np.zeros(n, np.int32)
        self.check(lib().rapt_bounce_labels(self.h, n, _dp(query), float(mass), float(charge),
                                            float(Re), float(e), float(eyegradientstep),
                                            float(fieldlineresolution), _dp(out), _ip32(st)),
                   "rapt_bounce_labels")
        return out, st


    def bounce_center_advance(self, desc, tpos, Bm, coef, dt, tcur, nrow, rtol, atol,
                              eyegradientstep, fieldlineresolution, nsteps=500):
        """Advance n bounce centres; returns (rows list, final tpos, status, naccpt)."""
        self.set_field(desc)
        tp = np.ascontiguousarray(np.atleast_2d(tpos), dtype=np.float64).copy()
        n = tp.shape[0]
        f = lambda a: np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=np.float64)
        Bm, coef, dt, tcur = f(Bm), f(coef), f(dt), f(tcur)
        nrow = np.ascontiguousarray(np.broadcast_to(nrow, (n,)), dtype=np.int64)
        rows = np.zeros((max(int(nrow.sum()), 1), 4))
        st = np.zeros(n, np.int32)
        acc = np.zeros(n, np.int64)
        self.check(lib().rapt_bounce_center_advance(
            self.h, n, _dp(tp), _dp(Bm), _dp(coef), _dp(dt), _dp(tcur), _ip64(nrow), _dp(rows),
            _ip32(st), _ip64(acc), float(rtol), float(atol), float(eyegradientstep),
            float(fieldlineresolution), int(nsteps)), "rapt_bounce_center_advance")
        off = np.r_[0, np.cumsum(nrow)]
        return [rows[off[i]:off[i + 1]] for i in range(n)], tp, st, acc


    def bounce_center_equatorial(self, desc, tpos, k, dt, tcur, nrow, rtol, atol, nsteps=500):
        """Equatorial bounce centres (gradient + inertial drift); k: (n, 4) rows of
        (mu, mass, charge, gamma).  Returns (rows list, final tpos, status, naccpt)."""
        self.set_field(desc)
        tp = np.ascontiguousarray(np.atleast_2d(tpos), dtype=np.float64).copy()
        n = tp.shape[0]
        k = np.ascontiguousarray(np.broadcast_to(k, (n, 4)), dtype=np.float64)
        f = lambda a: np.ascontiguousarray(np.broadcast_to(a, (n,)), dtype=np.float64)
        dt, tcur = f(dt), f(tcur)
        nrow = np.ascontiguousarray(np.broadcast_to(nrow, (n,)), dtype=np.int64)
        rows = np.zeros((max(int(nrow.sum()), 1), 4))
        st = np.zeros(n, np.int32)
        acc = np.zeros(n, np.int64)
        self.check(lib().rapt_bounce_center_equatorial(
            self.h, n, _dp(tp), _dp(k), _dp(dt), _dp(tcur), _ip64(nrow), _dp(rows), _ip32(st),
            _ip64(acc), float(rtol), float(atol), int(nsteps)), "rapt_bounce_center_equatorial")
        off = np.r_[0, np.cumsum(nrow)]
        return [rows[off[i]:off[i + 1]] for i in range(n)], tp, st, acc


def expr_compile(prog, arch="gfx950", kernels=0):
    """Compiles kernels of a traced user-field program for arch (no device
    needed; the code is cached for the process).  kernels: the library's bit
    mask (0: all nineteen; 1: the one-lane Particle kernel; ...).  Returns the
    library's message; raises RuntimeError with the compiler's log on
    failure."""
    code = np.ascontiguousarray(prog.code, dtype=np.int32)
    buf = C.create_string_buffer(8192)
    rc = lib().rapt_field_expr_compile(prog.nb, prog.ne, _ip32(code), arch.encode(),
                                       int(kernels), buf, len(buf))
    if rc != 0:
        raise RuntimeError("rapt_field_expr_compile failed (%d): %s" % (rc, buf.value.decode()))
    return buf.value.decode()


def context(device=0):
    with _lock:
        if device not in _ctx:
            _ctx[device] = Context(device)
        return _ctx[device]


class Batch:
    """Device-resident particle batch (rapt_batch_*)."""

    def __init__(self, ctx, n):
        self.ctx = ctx
        self.n = int(n)
        h = C.c_void_p()
        ctx.check(lib().rapt_batch_create(ctx.h, self.n, C.byref(h)), "rapt_batch_create")
        self.h = h
        self.width = None

    def close(self):
        if self.h:
            lib().rapt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, state, flags):
        state = np.ascontiguousarray(state, dtype=np.float64)
        flags = np.ascontiguousarray(flags, dtype=np.int32)
        assert state.shape == (NSLOT, self.n) and flags.shape == (self.n,)
        self.ctx.check(lib().rapt_batch_upload(self.h, _dp(state), _ip32(flags)),
                       "rapt_batch_upload")

    def epochs(self, desc, rows_per_epoch, max_epochs=1 << 30):
        """Run the batch in output epochs (rapt_batch_epochs_begin/epoch_next):
        yields (rows [n][rows_per_epoch][width] copy, nrows [n], status [n],
        running, Counters) per epoch until no particle is left running."""
        width = ROW_WIDTH[desc.mode]
        self.ctx.check(lib().rapt_batch_epochs_begin(self.h, C.byref(desc), int(rows_per_epoch)),
                       "rapt_batch_epochs_begin")
        try:
            for _ in range(max_epochs):
                r = C.POINTER(C.c_double)()
                nr, st = C.POINTER(C.c_int64)(), C.POINTER(C.c_int32)()
                left = C.c_int32(0)
                cnt = Counters()
                self.ctx.check(lib().rapt_batch_epoch_next(self.h, C.byref(r), C.byref(nr),
                                                           C.byref(st), C.byref(left),
                                                           C.byref(cnt)), "rapt_batch_epoch_next")
                n, k = self.n, int(rows_per_epoch)
                rows = np.ctypeslib.as_array(r, shape=(n * k * width,)).reshape(n, k, width).copy()
                nrows = np.ctypeslib.as_array(nr, shape=(n,)).copy()
                status = np.ctypeslib.as_array(st, shape=(n,)).copy()
                yield rows, nrows, status, left.value, cnt
                if left.value == 0:
                    break
        finally:
            lib().rapt_batch_epochs_end(self.h)

    def set_cost(self, cost):
        """Take particles in descending cost (e.g. accepted steps of a trial
        launch) instead of by output step; the next upload restores the
        output-step order."""
        cost = np.ascontiguousarray(cost, dtype=np.float64)
        assert cost.shape == (self.n,)
        self.ctx.check(lib().rapt_batch_set_cost(self.h, _dp(cost)), "rapt_batch_set_cost")

    def reset(self):
        """Restore the uploaded state on the device (no host transfer)."""
        self.ctx.check(lib().rapt_batch_reset(self.h), "rapt_batch_reset")

    def set_rows(self, width, row_cap):
        row_cap = np.ascontiguousarray(row_cap, dtype=np.int64)
        assert row_cap.shape == (self.n,)
        self.ctx.check(lib().rapt_batch_set_rows(self.h, width, _ip64(row_cap)),
                       "rapt_batch_set_rows")
        self.width = width
        self.row_cap = row_cap
        self.row_off = np.concatenate([[0], np.cumsum(row_cap)[:-1]]).astype(np.int64)

    def run(self, desc):
        cnt = Counters()
        self.ctx.check(lib().rapt_batch_run(self.h, C.byref(desc), C.byref(cnt)),
                       "rapt_batch_run")
        return cnt

    def download(self, rows=True):
        n = self.n
        state = np.zeros((NSLOT, n))
        flags = np.zeros(n, np.int32)
        status = np.zeros(n, np.int32)
        nrows = np.zeros(n, np.int64)
        naccpt = np.zeros(n, np.int64)
        rbuf = np.zeros((int(self.row_cap.sum()), self.width)) if rows else None
        self.ctx.check(lib().rapt_batch_download(
            self.h, _dp(state), _ip32(flags), _ip32(status), _ip64(nrows),
            _dp(rbuf) if rows and rbuf.size else None, _ip64(naccpt)), "rapt_batch_download")
        return dict(state=state, flags=flags, status=status, nrows=nrows, naccpt=naccpt,
                    rows=rbuf)


def make_desc(mode, params, delta, eom="TaoChanBrizardEOM", stepper=None, check=False,
              lanes=0, block=0, **kw):
    """TraceDesc from a rapt params dict (rapt/__init__.py:21-34) and the scipy
    integrator settings the tracers use (rapt/Particle.py:300,
    rapt/GuidingCenter.py:450)."""
    d = TraceDesc()
    d.mode = mode
    d.eom = EOM[eom] if isinstance(eom, str) else int(eom)
    if stepper is None:
        stepper = STEP_DOPRI5 if mode == MODE_GC else STEP_DOP853
    d.stepper = stepper
    d.precision = {"fp64": PREC_FP64, "fp32_field": PREC_FP32_FIELD}[kw.get("precision", "fp64")]
    d.nsteps = int(kw.get("nsteps", 500))
    d.enforce_equatorial = 1 if params.get("enforce equatorial", False) else 0
    d.check_adiabaticity = 1 if check else 0
    d.block = block
    rtol, atol = params["solvertolerances"]
    d.rtol, d.atol = float(rtol), float(atol)
    d.gc_rtol, d.gc_atol = float(rtol), float(atol)
    if mode == MODE_GC:
        d.safety, d.dfactor, d.ifactor, d.beta = 0.9, 0.2, 10.0, 0.0
    else:
        d.safety, d.dfactor, d.ifactor, d.beta = 0.9, 0.3, 6.0, 0.1
    for k in ("safety", "dfactor", "ifactor", "beta"):
        if k in kw:
            setattr(d, k, float(kw[k]))
    d.max_step = float(kw.get("max_step", 0.0))
    d.first_step = float(kw.get("first_step", 0.0))
    d.rkf_tol = float(kw.get("rkf_tol", 1e-3))
    d.rkf_hmax = float(kw.get("rkf_hmax", 0.0))
    d.rkf_hmin = float(kw.get("rkf_hmin", 1e-6))
    d.cyclotronresolution = float(params["cyclotronresolution"])
    d.gctimestep = float(params.get("GCtimestep", 0.0))
    d.epss = float(params["epss"])
    d.epst = float(params["epst"])
    d.delta = float(delta)
    d.lanes = int(lanes)
    d.lanes_per_particle = int(kw.get("lanes_per_particle", 0))
    d.work_order = int(kw.get("work_order", 0))
    d.mixed_quads = int(kw.get("mixed_quads", 0))
    d.mixed_octets = int(kw.get("mixed_octets", 0))
    return d
