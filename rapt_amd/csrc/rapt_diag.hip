// rapt_diag.hip -- rapt_particle_gc: launches of the per-row diagnostics
// kernel (rapt_diag_kernel.hpp).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rapt_hip.h"
#include "rapt_diag_kernel.hpp"
#include "rapt_internal.hpp"

using namespace rapt;

template <int M>
static void launch_pgc(hipStream_t st, const GcDiagArgs& A) {
  const int B = 256;
  hipLaunchKernelGGL(k_particle_gc<M>, dim3((unsigned)((A.n + B - 1) / B)), dim3(B), 0, st, A);
}

extern "C" int rapt_particle_gc(rapt_ctx* c, int64_t n, const double* rows, const double* mass,
                                const double* charge, int32_t per_row, double tol,
                                int32_t maxiter, double* out, int32_t* status) {
  if (!c || n < 0 || (n && (!rows || !mass || !charge || !out || !status)))
    return RAPT_E_ARG;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  if (!(tol > 0.0) || maxiter < 1) return set_err(c, RAPT_E_ARG, "need tol > 0, maxiter >= 1");
  if (n == 0) return RAPT_OK;
  HIP_OK(c, hipSetDevice(c->device));
  const int64_t nmq = per_row ? n : 1;
  double *d_rows = nullptr, *d_m = nullptr, *d_q = nullptr, *d_out = nullptr;
  int32_t* d_st = nullptr;
  int rc = RAPT_OK;
  if (hipMalloc(&d_rows, sizeof(double) * 7 * n) != hipSuccess ||
      hipMalloc(&d_m, sizeof(double) * nmq) != hipSuccess ||
      hipMalloc(&d_q, sizeof(double) * nmq) != hipSuccess ||
      hipMalloc(&d_out, sizeof(double) * 6 * n) != hipSuccess ||
      hipMalloc(&d_st, sizeof(int32_t) * n) != hipSuccess) {
    rc = set_err(c, RAPT_E_OOM, "diagnostics allocation failed");
  } else if (hipMemcpyAsync(d_rows, rows, sizeof(double) * 7 * n, hipMemcpyHostToDevice,
                            c->stream) != hipSuccess ||
             hipMemcpyAsync(d_m, mass, sizeof(double) * nmq, hipMemcpyHostToDevice,
                            c->stream) != hipSuccess ||
             hipMemcpyAsync(d_q, charge, sizeof(double) * nmq, hipMemcpyHostToDevice,
                            c->stream) != hipSuccess) {
    rc = set_err(c, RAPT_E_HIP, "diagnostics input copy failed");
  } else {
    GcDiagArgs A;
    A.F = c->field;
    A.n = n;
    A.rows = d_rows;
    A.mass = d_m;
    A.charge = d_q;
    A.mq_stride = per_row ? 1 : 0;
    A.tol = tol;
    A.maxiter = maxiter;
    A.pad = 0;
    A.out = d_out;
    A.status = d_st;
    switch (c->field.model) {
      case 0: launch_pgc<0>(c->stream, A); break;
      case 1: launch_pgc<1>(c->stream, A); break;
      case 2: launch_pgc<2>(c->stream, A); break;
      case 3: launch_pgc<3>(c->stream, A); break;
      case 4: launch_pgc<4>(c->stream, A); break;
      case 5: launch_pgc<5>(c->stream, A); break;
      case 6: launch_pgc<6>(c->stream, A); break;
      case 7: launch_pgc<7>(c->stream, A); break;
      case 8: launch_pgc<8>(c->stream, A); break;
      case 9: launch_pgc<9>(c->stream, A); break;
      case 10:
        if (c->expr_mode == RAPT_EXPR_COMPILED) {  // compiled at run time
          GcDiagArgs a = A;
          void* args[] = {&a};
          rc = rapt_jit_run_slot(c, RAPT_JIT_PARTICLE_GC, (unsigned)((A.n + 255) / 256), 256,
                                 args);
        } else {
          launch_pgc<10>(c->stream, A);
        }
        break;
    }
    if (rc == RAPT_OK &&
        (hipGetLastError() != hipSuccess ||
         hipMemcpyAsync(out, d_out, sizeof(double) * 6 * n, hipMemcpyDeviceToHost, c->stream) !=
             hipSuccess ||
         hipMemcpyAsync(status, d_st, sizeof(int32_t) * n, hipMemcpyDeviceToHost, c->stream) !=
             hipSuccess ||
         hipStreamSynchronize(c->stream) != hipSuccess))
      rc = set_err(c, RAPT_E_HIP, "diagnostics kernel failed");
  }
  (void)hipFree(d_rows);
  (void)hipFree(d_m);
  (void)hipFree(d_q);
  (void)hipFree(d_out);
  (void)hipFree(d_st);
  return rc;
}
