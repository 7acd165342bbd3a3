// rapt_bounce.hip -- launches and entry points of the field-line kernels
// (rapt_bounce_kernel.hpp): rapt_halfbounce, rapt_eye, rapt_bounce_labels,
// rapt_bounce_center_advance, rapt_bounce_center_equatorial.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <cmath>

#include <vector>

#include "../../include/rapt_hip.h"
#include "rapt_bounce_kernel.hpp"
#include "rapt_internal.hpp"

using namespace rapt;

// 64-lane blocks: a run holds at most 16,384 lanes (256 waves), which 64-lane
// blocks spread over every CU instead of a quarter of them.
template <int M>
static void launch_hb(rapt_ctx* c, const BounceArgs& A) {
  const int B = 64;
  hipLaunchKernelGGL(k_halfbounce<M>, dim3((unsigned)((A.lanes + B - 1) / B)), dim3(B), 0,
                     c->stream, A);
}

template <int M>
static void launch_lab(rapt_ctx* c, const BounceArgs& A, const LabelArgs& L) {
  const int B = 64;
  hipLaunchKernelGGL(k_bounce_labels<M>, dim3((unsigned)((A.lanes + B - 1) / B)), dim3(B), 0,
                     c->stream, A, L);
}

// True when every one of the n values is finite (a NaN or infinite start point
// or mirror field has no field line to trace).
static bool all_finite(const double* v, int64_t n) {
  for (int64_t i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

// Shared host side: device copies of the inputs, the per-lane trace scratch,
// one launch, results back.  in_w doubles in per query, out_w doubles out.
static int run_bounce(rapt_ctx* c, int64_t n, const double* in, int in_w, const double* Bm,
                      double res, bool positions, const LabelArgs* lab, double* out,
                      int out_w, int32_t* status, int32_t what = 0) {
  if (!all_finite(in, (int64_t)in_w * n) || (Bm && !all_finite(Bm, n)))
    return set_err(c, RAPT_E_ARG, "non-finite start point or mirror field");
  HIP_OK(c, hipSetDevice(c->device));
  BounceArgs A;
  memset(&A, 0, sizeof A);
  A.F = c->field;
  A.n = n;
  A.res = res;
  A.cap = 2048;
  A.what = what;
  A.lanes = n < 16384 ? n : 16384;
  const size_t npts = (size_t)2 * A.cap * A.lanes;
  const int narr = positions ? 6 : 4;
  double *d_in = nullptr, *d_bm = nullptr, *d_out = nullptr, *scratch = nullptr;
  int32_t* d_st = nullptr;
  int rc = RAPT_OK;
  if (hipMalloc(&d_in, sizeof(double) * in_w * n) != hipSuccess ||
      (Bm && hipMalloc(&d_bm, sizeof(double) * n) != hipSuccess) ||
      hipMalloc(&d_out, sizeof(double) * out_w * n) != hipSuccess ||
      hipMalloc(&d_st, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc(&scratch, sizeof(double) * narr * npts) != hipSuccess) {
    rc = set_err(c, RAPT_E_OOM, "bounce scratch allocation failed");
  } else {
    A.tpos = d_in;
    A.Bm = d_bm;
    A.out = d_out;
    A.status = d_st;
    A.S = scratch;
    A.Bv = scratch + npts;
    A.Cc = scratch + 2 * npts;
    A.Wk = scratch + 3 * npts;
    A.Zs = positions ? scratch + 4 * npts : nullptr;
    A.Rs = positions ? scratch + 5 * npts : nullptr;
    if (hipMemcpyAsync(d_in, in, sizeof(double) * in_w * n, hipMemcpyHostToDevice,
                       c->stream) != hipSuccess ||
        (Bm && hipMemcpyAsync(d_bm, Bm, sizeof(double) * n, hipMemcpyHostToDevice,
                              c->stream) != hipSuccess)) {
      rc = set_err(c, RAPT_E_HIP, "bounce input copy failed");
    } else {
      switch (c->field.model) {
        case 1: lab ? launch_lab<1>(c, A, *lab) : launch_hb<1>(c, A); break;
        case 2: lab ? launch_lab<2>(c, A, *lab) : launch_hb<2>(c, A); break;
        case 5: lab ? launch_lab<5>(c, A, *lab) : launch_hb<5>(c, A); break;
        case 8: lab ? launch_lab<8>(c, A, *lab) : launch_hb<8>(c, A); break;
        case 9: lab ? launch_lab<9>(c, A, *lab) : launch_hb<9>(c, A); break;
        case 10:
          if (c->expr_mode == RAPT_EXPR_COMPILED) {  // compiled at run time
            BounceArgs a = A;
            LabelArgs lb = lab ? *lab : LabelArgs{};
            void* args[] = {&a, &lb};
            rc = rapt_jit_run_slot(c, lab ? RAPT_JIT_LABELS : RAPT_JIT_HALFBOUNCE,
                                   (unsigned)((A.lanes + 63) / 64), 64, args);
          } else {
            lab ? launch_lab<10>(c, A, *lab) : launch_hb<10>(c, A);
          }
          break;
        default:
          rc = set_err(c, RAPT_E_ARG, "field-line integrals need a dipole-type, gridded or "
                       "user-written field (model %d)", c->field.model);
      }
      if (rc == RAPT_OK) {
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(out, d_out, sizeof(double) * out_w * n, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipMemcpyAsync(status, d_st, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
          rc = set_err(c, RAPT_E_HIP, "bounce kernel failed");
      }
    }
  }
  (void)hipFree(d_in);
  (void)hipFree(d_bm);
  (void)hipFree(d_out);
  (void)hipFree(d_st);
  (void)hipFree(scratch);
  return rc;
}

extern "C" int rapt_halfbounce(rapt_ctx* c, int64_t n, const double* tpos, const double* Bm,
                               double fieldlineresolution, double* sb, int32_t* status) {
  if (!c || n < 0 || (n && (!tpos || !Bm || !sb || !status))) return RAPT_E_ARG;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  if (!(fieldlineresolution > 0.0))
    return set_err(c, RAPT_E_ARG, "fieldlineresolution must be > 0");
  if (n == 0) return RAPT_OK;
  return run_bounce(c, n, tpos, 4, Bm, fieldlineresolution, false, nullptr, sb, 1, status);
}

extern "C" int rapt_eye(rapt_ctx* c, int64_t n, const double* tpos, const double* Bm,
                        double fieldlineresolution, double* eye, int32_t* status) {
  if (!c || n < 0 || (n && (!tpos || !Bm || !eye || !status))) return RAPT_E_ARG;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  if (!(fieldlineresolution > 0.0))
    return set_err(c, RAPT_E_ARG, "fieldlineresolution must be > 0");
  if (n == 0) return RAPT_OK;
  return run_bounce(c, n, tpos, 4, Bm, fieldlineresolution, false, nullptr, eye, 1, status, 1);
}

extern "C" int rapt_bounce_labels(rapt_ctx* c, int64_t n, const double* query, double mass,
                                  double charge, double Re, double eV, double eyegradientstep,
                                  double fieldlineresolution, double* out, int32_t* status) {
  if (!c || n < 0 || (n && (!query || !out || !status))) return RAPT_E_ARG;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  if (!(fieldlineresolution > 0.0) || !(eyegradientstep > 0.0) || !(mass > 0.0) ||
      charge == 0.0 || !(Re > 0.0) || !(eV > 0.0))
    return set_err(c, RAPT_E_ARG, "bad label parameters");
  if (!c->field.is_static)
    return set_err(c, RAPT_E_ARG, "bounce-averaged drift needs a static field");
  if (n == 0) return RAPT_OK;
  const LabelArgs L{mass, charge, eyegradientstep, Re, eV};
  return run_bounce(c, n, query, 3, nullptr, fieldlineresolution, true, &L, out, 3, status);
}

template <int M>
static void launch_bc(rapt_ctx* c, const BounceArgs& A, const BCArgs& C) {
  const int B = 64;
  hipLaunchKernelGGL(k_bounce_center<M>, dim3((unsigned)((A.lanes + B - 1) / B)), dim3(B), 0,
                     c->stream, A, C);
}

extern "C" int rapt_bounce_center_advance(rapt_ctx* c, int64_t n, double* tpos, const double* Bm,
                                          const double* coef, const double* dt,
                                          const double* tcur, const int64_t* nrow, double* rows,
                                          int32_t* status, int64_t* naccpt, double rtol,
                                          double atol, double eyegradientstep,
                                          double fieldlineresolution, int32_t nsteps) {
  if (!c || n < 0 || (n && (!tpos || !Bm || !coef || !dt || !tcur || !nrow || !rows ||
                            !status || !naccpt)))
    return RAPT_E_ARG;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  if (!c->field.is_static)
    return set_err(c, RAPT_E_ARG, "BounceCenter does not work with nonstatic fields");
  if (!(fieldlineresolution > 0.0) || !(eyegradientstep > 0.0) || !(rtol > 0.0) ||
      !(atol > 0.0))
    return set_err(c, RAPT_E_ARG, "bad BounceCenter parameters");
  if (n == 0) return RAPT_OK;
  if (!all_finite(tpos, 4 * n) || !all_finite(Bm, n) || !all_finite(coef, n) ||
      !all_finite(tcur, n))
    return set_err(c, RAPT_E_ARG, "non-finite BounceCenter state, mirror field or coefficient");
  int64_t total = 0;
  std::vector<int64_t> off((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    if (nrow[i] < 0 || !(dt[i] > 0.0)) return set_err(c, RAPT_E_ARG, "bad row count or step");
    off[(size_t)i] = total;
    total += nrow[i];
  }
  HIP_OK(c, hipSetDevice(c->device));
  BounceArgs A;
  memset(&A, 0, sizeof A);
  A.F = c->field;
  A.n = n;
  A.res = fieldlineresolution;
  A.cap = 2048;
  A.lanes = n < 16384 ? n : 16384;
  const size_t npts = (size_t)2 * A.cap * A.lanes;
  double *d_tpos = nullptr, *d_in = nullptr, *d_rows = nullptr, *scratch = nullptr;
  int64_t *d_nrow = nullptr, *d_off = nullptr, *d_acc = nullptr;
  int32_t* d_st = nullptr;
  int rc = RAPT_OK;
  if (hipMalloc(&d_tpos, sizeof(double) * 4 * n) != hipSuccess ||
      hipMalloc(&d_in, sizeof(double) * 4 * n) != hipSuccess ||
      hipMalloc(&d_rows, sizeof(double) * 4 * (total > 0 ? total : 1)) != hipSuccess ||
      hipMalloc(&d_nrow, sizeof(int64_t) * n) != hipSuccess ||
      hipMalloc(&d_off, sizeof(int64_t) * n) != hipSuccess ||
      hipMalloc(&d_acc, sizeof(int64_t) * n) != hipSuccess ||
      hipMalloc(&d_st, sizeof(int32_t) * n) != hipSuccess ||
      hipMalloc(&scratch, sizeof(double) * 4 * npts) != hipSuccess) {
    rc = set_err(c, RAPT_E_OOM, "BounceCenter allocation failed");
  } else {
    A.S = scratch;
    A.Bv = scratch + npts;
    A.Cc = scratch + 2 * npts;
    A.Wk = scratch + 3 * npts;
    BCArgs C;
    memset(&C, 0, sizeof C);
    C.n = n;
    C.tpos = d_tpos;
    C.Bm = d_in;
    C.coef = d_in + n;
    C.dt = d_in + 2 * n;
    C.tcur = d_in + 3 * n;
    C.nrow = d_nrow;
    C.row_off = d_off;
    C.rows = d_rows;
    C.status = d_st;
    C.naccpt = d_acc;
    C.rtol = rtol;
    C.atol = atol;
    C.dstep = eyegradientstep;
    C.nmax = nsteps > 0 ? nsteps : 500;
    // scipy's dopri5 defaults: safety 0.9, fac1 0.2, fac2 10, beta 0 -> 0.04
    C.c5.safe = 0.9;
    C.c5.facc1 = 1.0 / 0.2;
    C.c5.facc2 = 1.0 / 10.0;
    C.c5.beta = 0.04;
    C.c5.expo1 = 0.2 - 0.04 * 0.75;
    bool ok = hipMemcpyAsync(d_tpos, tpos, sizeof(double) * 4 * n, hipMemcpyHostToDevice,
                             c->stream) == hipSuccess &&
              hipMemcpyAsync(d_in, Bm, sizeof(double) * n, hipMemcpyHostToDevice,
                             c->stream) == hipSuccess &&
              hipMemcpyAsync(d_in + n, coef, sizeof(double) * n, hipMemcpyHostToDevice,
                             c->stream) == hipSuccess &&
              hipMemcpyAsync(d_in + 2 * n, dt, sizeof(double) * n, hipMemcpyHostToDevice,
                             c->stream) == hipSuccess &&
              hipMemcpyAsync(d_in + 3 * n, tcur, sizeof(double) * n, hipMemcpyHostToDevice,
                             c->stream) == hipSuccess &&
              hipMemcpyAsync(d_nrow, nrow, sizeof(int64_t) * n, hipMemcpyHostToDevice,
                             c->stream) == hipSuccess &&
              hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
                             c->stream) == hipSuccess;
    if (!ok) {
      rc = set_err(c, RAPT_E_HIP, "BounceCenter input copy failed");
    } else {
      switch (c->field.model) {
        case 1: launch_bc<1>(c, A, C); break;
        case 2: launch_bc<2>(c, A, C); break;
        case 8: launch_bc<8>(c, A, C); break;
        case 10:
          if (c->expr_mode == RAPT_EXPR_COMPILED) {
            BounceArgs a = A;
            BCArgs cc = C;
            void* args[] = {&a, &cc};
            rc = rapt_jit_run_slot(c, RAPT_JIT_BC, (unsigned)((A.lanes + 63) / 64), 64, args);
          } else {
            launch_bc<10>(c, A, C);
          }
          break;
        default:
          rc = set_err(c, RAPT_E_ARG, "BounceCenter needs a static dipole-type or user-written "
                       "field (model %d)", c->field.model);
      }
      if (rc == RAPT_OK) {
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(tpos, d_tpos, sizeof(double) * 4 * n, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            (total > 0 && hipMemcpyAsync(rows, d_rows, sizeof(double) * 4 * total,
                                         hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
            hipMemcpyAsync(status, d_st, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipMemcpyAsync(naccpt, d_acc, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
          rc = set_err(c, RAPT_E_HIP, "BounceCenter kernel failed");
      }
    }
  }
  (void)hipFree(d_tpos);
  (void)hipFree(d_in);
  (void)hipFree(d_rows);
  (void)hipFree(d_nrow);
  (void)hipFree(d_off);
  (void)hipFree(d_acc);
  (void)hipFree(d_st);
  (void)hipFree(scratch);
  return rc;
}

template <int M>
static void launch_bc_eq(rapt_ctx* c, const BCArgs& C, const EqArgs& E) {
  const int B = 64;
  hipLaunchKernelGGL(k_bounce_center_eq<M>, dim3((unsigned)((C.n + B - 1) / B)), dim3(B), 0,
                     c->stream, c->field, C, E);
}

extern "C" int rapt_bounce_center_equatorial(rapt_ctx* c, int64_t n, double* tpos,
                                             const double* k, const double* dt,
                                             const double* tcur, const int64_t* nrow,
                                             double* rows, int32_t* status, int64_t* naccpt,
                                             double rtol, double atol, int32_t nsteps) {
  if (!c || n < 0 || (n && (!tpos || !k || !dt || !tcur || !nrow || !rows || !status ||
                            !naccpt)))
    return RAPT_E_ARG;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  if (!c->field.is_static)
    return set_err(c, RAPT_E_ARG, "BounceCenter does not work with nonstatic fields");
  if (!(rtol > 0.0) || !(atol > 0.0)) return set_err(c, RAPT_E_ARG, "bad BounceCenter tolerances");
  if (n == 0) return RAPT_OK;
  if (!all_finite(tpos, 4 * n) || !all_finite(k, 4 * n) || !all_finite(tcur, n))
    return set_err(c, RAPT_E_ARG, "non-finite BounceCenter state or constants");
  int64_t total = 0;
  std::vector<int64_t> off((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    if (nrow[i] < 0 || !(dt[i] > 0.0)) return set_err(c, RAPT_E_ARG, "bad row count or step");
    off[(size_t)i] = total;
    total += nrow[i];
  }
  HIP_OK(c, hipSetDevice(c->device));
  double *d_tpos = nullptr, *d_in = nullptr, *d_rows = nullptr;
  int64_t *d_nrow = nullptr, *d_off = nullptr, *d_acc = nullptr;
  int32_t* d_st = nullptr;
  int rc = RAPT_OK;
  if (hipMalloc(&d_tpos, sizeof(double) * 4 * n) != hipSuccess ||
      hipMalloc(&d_in, sizeof(double) * 6 * n) != hipSuccess ||
      hipMalloc(&d_rows, sizeof(double) * 4 * (total > 0 ? total : 1)) != hipSuccess ||
      hipMalloc(&d_nrow, sizeof(int64_t) * n) != hipSuccess ||
      hipMalloc(&d_off, sizeof(int64_t) * n) != hipSuccess ||
      hipMalloc(&d_acc, sizeof(int64_t) * n) != hipSuccess ||
      hipMalloc(&d_st, sizeof(int32_t) * n) != hipSuccess) {
    rc = set_err(c, RAPT_E_OOM, "BounceCenter allocation failed");
  } else {
    BCArgs C;
    memset(&C, 0, sizeof C);
    C.n = n;
    C.tpos = d_tpos;
    C.dt = d_in + 4 * n;
    C.tcur = d_in + 5 * n;
    C.nrow = d_nrow;
    C.row_off = d_off;
    C.rows = d_rows;
    C.status = d_st;
    C.naccpt = d_acc;
    C.rtol = rtol;
    C.atol = atol;
    C.nmax = nsteps > 0 ? nsteps : 500;
    // scipy's dopri5 defaults: safety 0.9, fac1 0.2, fac2 10, beta 0 -> 0.04
    C.c5.safe = 0.9;
    C.c5.facc1 = 1.0 / 0.2;
    C.c5.facc2 = 1.0 / 10.0;
    C.c5.beta = 0.04;
    C.c5.expo1 = 0.2 - 0.04 * 0.75;
    const EqArgs E{d_in};
    const bool ok =
        hipMemcpyAsync(d_tpos, tpos, sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->stream) ==
            hipSuccess &&
        hipMemcpyAsync(d_in, k, sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->stream) ==
            hipSuccess &&
        hipMemcpyAsync(d_in + 4 * n, dt, sizeof(double) * n, hipMemcpyHostToDevice, c->stream) ==
            hipSuccess &&
        hipMemcpyAsync(d_in + 5 * n, tcur, sizeof(double) * n, hipMemcpyHostToDevice,
                       c->stream) == hipSuccess &&
        hipMemcpyAsync(d_nrow, nrow, sizeof(int64_t) * n, hipMemcpyHostToDevice, c->stream) ==
            hipSuccess &&
        hipMemcpyAsync(d_off, off.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice,
                       c->stream) == hipSuccess;
    if (!ok) {
      rc = set_err(c, RAPT_E_HIP, "BounceCenter input copy failed");
    } else {
      switch (c->field.model) {
        case 0: launch_bc_eq<0>(c, C, E); break;
        case 1: launch_bc_eq<1>(c, C, E); break;
        case 2: launch_bc_eq<2>(c, C, E); break;
        case 3: launch_bc_eq<3>(c, C, E); break;
        case 4: launch_bc_eq<4>(c, C, E); break;
        case 6: launch_bc_eq<6>(c, C, E); break;
        case 7: launch_bc_eq<7>(c, C, E); break;
        case 8: launch_bc_eq<8>(c, C, E); break;
        case 9: launch_bc_eq<9>(c, C, E); break;
        case 10:
          if (c->expr_mode == RAPT_EXPR_COMPILED) {
            FieldC f = c->field;
            BCArgs cc = C;
            EqArgs e = E;
            void* args[] = {&f, &cc, &e};
            rc = rapt_jit_run_slot(c, RAPT_JIT_BC_EQ, (unsigned)((C.n + 63) / 64), 64, args);
          } else {
            launch_bc_eq<10>(c, C, E);
          }
          break;
        default:
          rc = set_err(c, RAPT_E_ARG, "BounceCenter needs a static field (model %d)",
                       c->field.model);
      }
      if (rc == RAPT_OK) {
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(tpos, d_tpos, sizeof(double) * 4 * n, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            (total > 0 && hipMemcpyAsync(rows, d_rows, sizeof(double) * 4 * total,
                                         hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
            hipMemcpyAsync(status, d_st, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipMemcpyAsync(naccpt, d_acc, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                           c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
          rc = set_err(c, RAPT_E_HIP, "BounceCenter kernel failed");
      }
    }
  }
  (void)hipFree(d_tpos);
  (void)hipFree(d_in);
  (void)hipFree(d_rows);
  (void)hipFree(d_nrow);
  (void)hipFree(d_off);
  (void)hipFree(d_acc);
  (void)hipFree(d_st);
  return rc;
}
