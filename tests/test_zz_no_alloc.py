"""The allocation-free entry point rapt_trace_device_noalloc (caller-owned
device buffers, counters in the library's static scratch) against
rapt_batch_run, through the C-ABI: two calls in a row, the second with a
smaller batch of another mode, each bit-identical to the batch path (rows,
row counts, statuses, final state, accepted-step counts and totals)."""
import ctypes as C

import numpy as np
import pytest

import rapt_amd as rapt
import rapt_amd.fields as F
from rapt_amd import _native as N
from rapt_amd import batch
from rapt_amd import constants as K
from rapt_amd import workloads as W

pytestmark = pytest.mark.gpu


class _Dev:
    """Device buffers through the HIP runtime (freed at close)."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def put(self, arr):
        arr = np.ascontiguousarray(arr)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(arr.nbytes, 8)) == 0
        self.bufs.append(p)
        if arr.nbytes:
            assert self.hip.hipMemcpy(p, arr.ctypes.data, arr.nbytes, 1) == 0
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.hip.hipFree(p)
        self.bufs = []


def _noalloc(field, mode, S, flags, row_cap, desc):
    ctx = N.context()
    ctx.set_field(field.descriptor())
    n = S.shape[1]
    width = N.ROW_WIDTH[mode]
    row_off = np.r_[0, np.cumsum(row_cap)[:-1]].astype(np.int64)
    dev = _Dev()
    try:
        d_state, d_flags = dev.put(S), dev.put(flags.astype(np.int32))
        d_off, d_cap = dev.put(row_off), dev.put(row_cap.astype(np.int64))
        rows0 = np.zeros((int(row_cap.sum()), width))
        d_rows = dev.put(rows0)
        d_nrows, d_status, d_acc = (dev.put(np.zeros(n, np.int64)), dev.put(np.zeros(n, np.int32)),
                                    dev.put(np.zeros(n, np.int64)))
        cnt = N.Counters()
        ctx.check(N.lib().rapt_trace_device_noalloc(ctx.h, C.byref(desc), n, d_state, d_flags,
                                                    d_off, d_cap, d_rows, d_nrows, d_status,
                                                    d_acc, None, C.byref(cnt)),
                  "rapt_trace_device_noalloc")
        rows = dev.get(d_rows, rows0)
        nrows = dev.get(d_nrows, np.zeros(n, np.int64))
        out = dict(state=dev.get(d_state, S), flags=dev.get(d_flags, flags.astype(np.int32)),
                   status=dev.get(d_status, np.zeros(n, np.int32)), nrows=nrows,
                   naccpt=dev.get(d_acc, np.zeros(n, np.int64)), cnt=cnt,
                   rows=[rows[row_off[i]:row_off[i] + nrows[i]] for i in range(n)])
    finally:
        dev.close()
    return out


def _same(got, ref):
    np.testing.assert_array_equal(got["status"], ref.status)
    np.testing.assert_array_equal(got["state"], ref.state)
    np.testing.assert_array_equal(got["naccpt"], ref.naccpt)
    for x, y in zip(got["rows"], ref.rows):
        np.testing.assert_array_equal(x, y)
    assert got["cnt"].naccpt == ref.counters["naccpt"]
    assert got["cnt"].nfcn == ref.counters["nfcn"]
    assert got["cnt"].nrejct == ref.counters["nrejct"]


def test_noalloc_matches_batch_run_twice():
    # call 1: 64 DoubleDipole electrons as guiding centres (lane-quad kernel)
    wl = W.doubledipole_electrons_gc(n=64)
    f = F.DoubleDipole()
    v = wl["v"]
    g = 1 / np.sqrt(1 - (v / K.c) ** 2)
    ppar = g * K.m_el * v * np.cos(np.radians(wl["pa"]))
    last5 = np.c_[np.zeros(64), wl["pos"], ppar]
    B = N.context().field_eval(f.descriptor(), "magB", last5[:, :4])[:, 0]
    vpar = ppar / (K.m_el * g)
    mu = g ** 2 * K.m_el * (v - vpar) * (v + vpar) / (2 * B)
    prm = dict(rapt.params, solvertolerances=(1e-8, 1e-8))
    dt = np.full(64, 0.05)
    ref = batch.advance_gcs(f, last5, K.m_el, -K.e, mu, v, dt, 0.5, params=prm)
    S = batch.gc_state(last5, K.m_el, -K.e, mu, v, dt)
    cap = batch._rows_needed(0.5, dt)
    got = _noalloc(f, N.MODE_GC, S, np.zeros(64, np.int32), cap,
                   N.make_desc(N.MODE_GC, prm, 0.5))
    assert np.all(got["status"] == N.ST_DONE)
    _same(got, ref)
    # call 2: a smaller batch, 16 dipole protons in Particle mode
    wl = W.dipole_protons_1mev(n=16)
    g = 1 / np.sqrt(1 - np.sum(wl["vel"] ** 2, axis=1) / K.c ** 2)
    last7 = np.c_[np.zeros(16), wl["pos"], (K.m_pr * g)[:, None] * wl["vel"]]
    fd = F.EarthDipole()
    prm = dict(rapt.params, cyclotronresolution=20)
    ref = batch.advance_particles(fd, last7, K.m_pr, K.e, 0.5, params=prm)
    S = batch.particle_state(last7, K.m_pr, K.e)
    cap = batch._rows_needed(0.5, batch.particle_dt(fd, last7, K.m_pr, K.e, 20))
    got = _noalloc(fd, N.MODE_PARTICLE, S, np.zeros(16, np.int32), cap,
                   N.make_desc(N.MODE_PARTICLE, prm, 0.5))
    assert np.all(got["status"] == N.ST_DONE)
    _same(got, ref)


def test_noalloc_rejects_missing_buffers():
    ctx = N.context()
    ctx.set_field(F.EarthDipole().descriptor())
    d = N.make_desc(N.MODE_PARTICLE, dict(rapt.params), 0.5)
    rc = N.lib().rapt_trace_device_noalloc(ctx.h, C.byref(d), 4, None, None, None, None, None,
                                           None, None, None, None, None)
    assert rc == -1
    assert b"device buffer" in N.lib().rapt_last_error(ctx.h)
