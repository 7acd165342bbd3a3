// rapt_tracer.hip -- the C-ABI of librapt_hip.so (include/rapt_hip.h): contexts,
// field evaluation, batches, and the launch of the FP64 tracer kernels.  The
// kernel itself is k_trace in rapt_trace_kernel.hpp; the FP32-field
// instantiations live in rapt_tracer_f32.hip and the grouped guiding-centre
// kernels (two, four or eight lanes per particle) in rapt_gcn.hip.
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <math.h>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/rapt_hip.h"
#include "rapt_fields.hpp"
#include "rapt_internal.hpp"
#include "rapt_physics.hpp"
#include "rapt_steppers.hpp"
#include "rapt_trace_kernel.hpp"

namespace rapt {
// epoch graph nodes: counter reset (a fill kernel, not a memset node) and the
// count of particles that stopped on a full epoch row buffer
__global__ void k_epoch_prep(unsigned long long* work, int32_t* left) {
  if (threadIdx.x < 8) work[threadIdx.x] = 0ULL;
  if (threadIdx.x == 0) *left = 0;
}

// Gridded fields: expand the point layout [nt][nx][ny][nz][4] into the cell
// layout [nt][nx-1][ny-1][nz-1][32] (the 8 corners of each cell, corner
// (dx, dy, dz) at 3 (4 dx + 2 dy + dz), the last 8 doubles zero).  One thread
// per (cell, corner).
__global__ void k_grid_cells(const double* P, double* C, int32_t nx, int32_t ny, int32_t nz,
                             int64_t ncell) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= ncell * 8) return;
  const int corner = (int)(idx & 7);
  int64_t cell = idx >> 3;
  const int dz = corner & 1, dy = (corner >> 1) & 1, dx = corner >> 2;
  const int64_t k = cell % (nz - 1);
  int64_t r = cell / (nz - 1);
  const int64_t j = r % (ny - 1);
  r /= (ny - 1);
  const int64_t i = r % (nx - 1);
  const int64_t it = r / (nx - 1);
  const double* p = P + (((it * nx + i + dx) * ny + j + dy) * nz + k + dz) * 4;
  double* c = C + cell * 32 + 3 * corner;
  c[0] = p[0];
  c[1] = p[1];
  c[2] = p[2];
  if (corner == 7)
    for (int m = 0; m < 8; ++m) C[cell * 32 + 24 + m] = 0.0;
}
__global__ void k_epoch_left(int64_t n, const int32_t* status, int32_t* left) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (status[i] == RAPT_ST_CAP) atomicAdd(left, 1);
}
}  // namespace rapt

// ============================================================================
// host side
// ============================================================================
using namespace rapt;

struct rapt_batch {
  rapt_ctx* ctx = nullptr;
  int64_t n = 0;
  double* d_state = nullptr;
  double* d_state0 = nullptr;  // state as uploaded (rapt_batch_reset)
  int32_t* d_flags0 = nullptr;
  int32_t* d_flags = nullptr;
  int32_t* d_status = nullptr;
  int64_t* d_nrows = nullptr;
  int64_t* d_naccpt = nullptr;
  int64_t* d_rowoff = nullptr;
  int64_t* d_rowcap = nullptr;
  int64_t* d_order = nullptr;  // work order: particle indices by output step
  double* d_rows = nullptr;
  int32_t width = 0;
  int64_t total_rows = 0;
  unsigned long long* d_work = nullptr;
  // output epochs: a captured graph that runs one epoch of at most ep_rows
  // rows per particle and copies its rows to pinned host memory
  int32_t ep_rows = 0;
  double* d_erows = nullptr;  // [n * ep_rows][width]
  int64_t* d_eoff = nullptr;  // [n]: i * ep_rows
  int64_t* d_ecap = nullptr;  // [n]: ep_rows
  int32_t* d_left = nullptr;  // particles still running after an epoch
  double* h_rows = nullptr;   // pinned copies of the epoch's outputs
  int64_t* h_nrows = nullptr;
  int32_t* h_status = nullptr;
  int32_t* h_left = nullptr;
  unsigned long long* h_work = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};


static int field_width(int what) {
  switch (what) {
    case 0: case 1: case 2: case 4: case 7: case 9: return 3;
    case 5: return 9;
    case 3: case 6: case 8: case 10: case 11: return 1;
    default: return -1;
  }
}

template <int M>
static void launch_field_eval(rapt_ctx* c, int what, int64_t n, const double* dt, double* dout,
                              int w) {
  const int B = 256;
  const int64_t g = (n + B - 1) / B;
  hipLaunchKernelGGL(k_field_eval<M>, dim3((unsigned)g), dim3(B), 0, c->stream, c->field, what,
                     n, dt, dout, w);
}

template <int M>
static void launch_rhs_eval(rapt_ctx* c, int mode, int eom, int eq, int64_t n, const double* ty,
                            const double* k, double* out) {
  const int B = 256;
  hipLaunchKernelGGL(k_rhs_eval<M>, dim3((unsigned)((n + B - 1) / B)), dim3(B), 0, c->stream,
                     c->field, mode, eom, eq, n, ty, k, out);
}

static CtlParams ctl(double safety, double dfactor, double ifactor, double beta, bool dop853) {
  CtlParams p;
  p.safe = safety == 0.0 ? 0.9 : safety;
  const double fac1 = dfactor == 0.0 ? (dop853 ? 0.333 : 0.2) : dfactor;
  const double fac2 = ifactor == 0.0 ? (dop853 ? 6.0 : 10.0) : ifactor;
  p.facc1 = 1.0 / fac1;
  p.facc2 = 1.0 / fac2;
  if (dop853) {
    p.beta = beta <= 0.0 ? 0.0 : beta;
    p.expo1 = 1.0 / 8.0 - p.beta * 0.2;
  } else {
    p.beta = beta == 0.0 ? 0.04 : (beta < 0.0 ? 0.0 : beta);
    p.expo1 = 0.2 - p.beta * 0.75;
  }
  return p;
}

void rapt_launch_gcn(hipStream_t st, int model, int lanes_per_particle, const KArgs& A,
                     int block, int64_t groups, int64_t quads, int64_t pairs);

void rapt_launch_trace_f32(hipStream_t st, int model, int kind, const KArgs& A, int block,
                           int64_t lanes);

// ---- scratch of the allocation-free entry point -----------------------------
// The work-queue counter(s) and step counters that rapt_batch_run keeps in the
// batch's d_work: 8 words (queue 0, accepted, rejected, attempts, RHS
// evaluations, switches, the mixed layouts' queues 1 and 2).  One copy per
// device, held until the process ends; host code uses the address that
// hipGetSymbolAddress gives for the current device, never this name.
#define RAPT_NOALLOC_WORDS 8
static __device__ __align__(256) unsigned long long scratch[RAPT_NOALLOC_WORDS];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned long long* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on `device` (current), looked up once per
// device.  Returns nullptr when the lookup fails.
unsigned long long* scratch_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
    g_scratch[device] = static_cast<unsigned long long*>(p);
  }
  return g_scratch[device];
}
}  // namespace

// clears n words (a fill kernel rather than a memset node, so a caller may
// capture the entry point in a graph)
__global__ void k_fill_u64(unsigned long long* p, int n) {
  const int i = (int)threadIdx.x;
  if (i < n) p[i] = 0ULL;
}

template <int M, int K>
static void launch_trace(rapt_ctx* c, const KArgs& A, int block, int64_t lanes) {
  const int64_t g = (lanes + block - 1) / block;
  if constexpr (K == KIND_GC && M != F_VAR_EARTH_DIPOLE) {
    if (trace_use_sp<M>(A)) {
      hipLaunchKernelGGL((k_trace<M, K, RAPT_PREC_FP64, true>), dim3((unsigned)g), dim3(block),
                         0, c->stream, A);
      return;
    }
  }
  hipLaunchKernelGGL((k_trace<M, K, RAPT_PREC_FP64>), dim3((unsigned)g), dim3(block), 0,
                     c->stream, A);
}

template <int K>
static void launch_model(rapt_ctx* c, const KArgs& A, int block, int64_t lanes) {
  // a compiled user-written field, once its kernel is loaded (otherwise the
  // built-in kernel interprets the program: the same results)
  if (c->field.model == RAPT_FIELD_EXPR && c->expr_mode == RAPT_EXPR_COMPILED &&
      rapt_jit_launch(c, K, K == KIND_GC && trace_use_sp<F_EXPR>(A), A, block, lanes))
    return;
  switch (c->field.model) {
    case 0: launch_trace<0, K>(c, A, block, lanes); break;
    case 1: launch_trace<1, K>(c, A, block, lanes); break;
    case 2: launch_trace<2, K>(c, A, block, lanes); break;
    case 3: launch_trace<3, K>(c, A, block, lanes); break;
    case 4: launch_trace<4, K>(c, A, block, lanes); break;
    case 5: launch_trace<5, K>(c, A, block, lanes); break;
    case 6: launch_trace<6, K>(c, A, block, lanes); break;
    case 7: launch_trace<7, K>(c, A, block, lanes); break;
    case 8: launch_trace<8, K>(c, A, block, lanes); break;
    case 9: launch_trace<9, K>(c, A, block, lanes); break;
    case 10: launch_trace<10, K>(c, A, block, lanes); break;
  }
}

extern "C" {

int rapt_ctx_create(int device, rapt_ctx** out) {
  if (!out) return RAPT_E_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return RAPT_E_HIP;
  if (device < 0 || device >= ndev) return RAPT_E_ARG;
  rapt_ctx* c = new (std::nothrow) rapt_ctx();
  if (!c) return RAPT_E_OOM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return RAPT_E_HIP;
  }
  // this device's copy of the allocation-free scratch (device current)
  c->noalloc_scratch = scratch_address(device);
  if (!c->noalloc_scratch) {
    (void)hipEventDestroy(c->ev0);
    (void)hipEventDestroy(c->ev1);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return RAPT_E_HIP;
  }
  *out = c;
  return RAPT_OK;
}

int rapt_ctx_destroy(rapt_ctx* c) {
  if (!c) return RAPT_E_ARG;
  (void)hipSetDevice(c->device);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  (void)hipFree(c->grid_mem);
  (void)hipFree(c->grid_cells);
  (void)hipFree(c->expr_mem);
  delete c;
  return RAPT_OK;
}

const char* rapt_last_error(rapt_ctx* c) { return c ? c->err.c_str() : "null context"; }

int rapt_field_set(rapt_ctx* c, const rapt_field_desc* f) {
  if (!c || !f) return RAPT_E_ARG;
  if (f->model < 0 || f->model >= RAPT_FIELD_NMODELS)
    return set_err(c, RAPT_E_ARG, "unknown field model %d", f->model);
  if (!(f->gstep > 0.0)) return set_err(c, RAPT_E_ARG, "gradientstepsize must be > 0");
  if (!f->is_static && !(f->tstep > 0.0))
    return set_err(c, RAPT_E_ARG, "timederivstepsize must be > 0");
  if (f->model == RAPT_FIELD_GRID && !c->grid_set)
    return set_err(c, RAPT_E_STATE, "gridded field: no grid uploaded (rapt_grid_set)");
  if (f->model == RAPT_FIELD_EXPR && !c->expr_set)
    return set_err(c, RAPT_E_STATE, "user-written field: no program uploaded "
                                    "(rapt_field_expr_set)");
  static_assert(offsetof(FieldC, i2d) == sizeof(rapt_field_desc), "descriptor layout");
  memcpy(&c->field, f, sizeof(rapt_field_desc));
  c->field.i2d = 1.0 / (2.0 * f->gstep);
  c->field.ihts = f->is_static ? 0.0 : 0.5 / f->tstep;
  if (f->model == RAPT_FIELD_EXPR)
    c->field.expr = c->expr;
  else
    c->field.grid = f->model == RAPT_FIELD_GRID ? c->grid : GridC{};
  c->field_set = true;
  return RAPT_OK;
}

int rapt_field_expr_set(rapt_ctx* c, int32_t nb, int32_t ne, const int32_t* code, int32_t nconst,
                        const double* consts) {
  if (!c || !code || (nconst > 0 && !consts)) return RAPT_E_ARG;
  if (nb < 1 || ne < 0 || nb + ne > RAPT_EXPR_MAXCODE || nconst < 0 ||
      nconst > RAPT_EXPR_MAXCONST)
    return set_err(c, RAPT_E_ARG, "field program: %d + %d instructions, %d constants (at most "
                   "%d instructions, %d constants)", nb, ne, nconst, RAPT_EXPR_MAXCODE,
                   RAPT_EXPR_MAXCONST);
  // every operand register is an input (0-3) or written earlier in its own
  // program, and each program writes its three results (registers 4-6), so
  // the interpreter never reads a register it did not set
  for (int prog = 0; prog < 2; ++prog) {
    const int32_t lo = prog == 0 ? 0 : nb, hi = prog == 0 ? nb : nb + ne;
    if (lo == hi) continue;
    uint32_t set = 0xFu;
    for (int32_t i = lo; i < hi; ++i) {
      const int32_t op = code[4 * i], d = code[4 * i + 1], a = code[4 * i + 2],
                    b = code[4 * i + 3];
      const bool binary = op >= 1 && op <= 4 ? true : (op == 6 || op == 20 || op == 23 ||
                                                      op == 24 || op == 25);
      const bool regs = d >= 4 && d < EXPR_NREG && a >= 0 && b >= 0 &&
                        (op == 0 ? a < nconst : a < EXPR_NREG) && b < EXPR_NREG;
      const bool defined = op == 0 || (((set >> a) & 1u) && (!binary || ((set >> b) & 1u)));
      if (op < 0 || op > 26 || !regs || !defined)
        return set_err(c, RAPT_E_ARG, "field program: bad instruction %d (%d %d %d %d)", i, op,
                       d, a, b);
      set |= 1u << d;
    }
    if ((set & 0x70u) != 0x70u)
      return set_err(c, RAPT_E_ARG, "field program %d leaves a result register unset", prog);
  }
  HIP_OK(c, hipSetDevice(c->device));
  const size_t kb = sizeof(double) * (size_t)(nconst > 0 ? nconst : 1);
  const size_t cb = sizeof(int32_t) * 4 * (size_t)(nb + ne);
  void* mem = nullptr;
  if (hipMalloc(&mem, kb + cb) != hipSuccess)
    return set_err(c, RAPT_E_OOM, "field program allocation failed");
  if ((nconst > 0 && hipMemcpy(mem, consts, sizeof(double) * nconst, hipMemcpyHostToDevice) !=
                         hipSuccess) ||
      hipMemcpy((char*)mem + kb, code, cb, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(mem);
    return set_err(c, RAPT_E_HIP, "field program upload failed");
  }
  // work still queued on the stream may read the previous program
  HIP_OK(c, hipStreamSynchronize(c->stream));
  (void)hipFree(c->expr_mem);
  c->expr_mem = mem;
  c->expr = ExprC{(const int32_t*)((char*)mem + kb), (const double*)mem, nb, ne};
  c->expr_set = true;
  c->expr_code.assign(code, code + 4 * (size_t)(nb + ne));
  c->expr_nb = nb;
  c->expr_ne = ne;
  for (hipFunction_t& f : c->jit_fn) f = nullptr;  // the compiled kernels of another program
  if (c->field_set && c->field.model == RAPT_FIELD_EXPR) c->field.expr = c->expr;
  return RAPT_OK;
}

int rapt_grid_set(rapt_ctx* c, int32_t nx, int32_t ny, int32_t nz, int32_t nt, const double* x,
                  const double* y, const double* z, const double* t, const double* B,
                  const double* E) {
  if (!c || !x || !y || !z || !B || (nt > 1 && !t)) return RAPT_E_ARG;
  if (nx < 2 || ny < 2 || nz < 2 || nt < 1)
    return set_err(c, RAPT_E_ARG, "grid needs at least 2 points per axis and 1 time");
  const double* ax[4] = {x, y, z, t};
  const int32_t n[4] = {nx, ny, nz, nt};
  for (int k = 0; k < (nt > 1 ? 4 : 3); ++k)
    for (int32_t i = 0; i + 1 < n[k]; ++i)
      if (!(ax[k][i] < ax[k][i + 1]))
        return set_err(c, RAPT_E_ARG, "grid axis %d is not strictly increasing", k);
  HIP_OK(c, hipSetDevice(c->device));
  const size_t npt = (size_t)nx * ny * nz * nt;
  const size_t naxis = (size_t)nx + ny + nz + nt;
  const size_t total = naxis + (E ? 8 : 4) * npt;
  std::vector<double> h(total, 0.0);
  size_t o = 0;
  size_t off[4];
  for (int k = 0; k < 4; ++k) {
    off[k] = o;
    if (k < 3 || nt > 1) memcpy(&h[o], ax[k], sizeof(double) * n[k]);
    o += n[k];
  }
  const size_t offB = o, offE = o + 4 * npt;
  for (size_t i = 0; i < npt; ++i) {
    for (int m = 0; m < 3; ++m) {
      h[offB + 4 * i + m] = B[3 * i + m];
      if (E) h[offE + 4 * i + m] = E[3 * i + m];
    }
  }
  double* mem = nullptr;
  if (hipMalloc(&mem, sizeof(double) * total) != hipSuccess)
    return set_err(c, RAPT_E_OOM, "grid allocation of %zu bytes failed", sizeof(double) * total);
  if (hipMemcpy(mem, h.data(), sizeof(double) * total, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(mem);
    return set_err(c, RAPT_E_HIP, "grid upload failed");
  }
  HIP_OK(c, hipStreamSynchronize(c->stream));
  // Cell layout when it fits (8x the point data; at most 64 GiB and half the
  // free device memory): one aligned 192-byte run per evaluation instead of
  // four 64-byte runs.  Otherwise the point layout stays.
  const size_t ncell = (size_t)(nx - 1) * (ny - 1) * (nz - 1) * nt;
  const size_t cell_bytes = sizeof(double) * 32 * ncell * (E ? 2 : 1);
  size_t free_b = 0, total_b = 0;
  (void)hipMemGetInfo(&free_b, &total_b);
  double* cells = nullptr;
  const char* pts = getenv("RAPT_GRID_POINTS");  // "1": keep the point layout
  const bool want_cells = !(pts && pts[0] == '1');
  if (want_cells && cell_bytes <= ((size_t)64 << 30) && cell_bytes <= free_b / 2 &&
      hipMalloc(&cells, cell_bytes) == hipSuccess) {
    const int64_t nthr = (int64_t)ncell * 8;
    const unsigned nblk = (unsigned)((nthr + 255) / 256);
    hipLaunchKernelGGL(k_grid_cells, dim3(nblk), dim3(256), 0, c->stream, mem + offB, cells,
                       nx, ny, nz, (int64_t)ncell);
    if (E)
      hipLaunchKernelGGL(k_grid_cells, dim3(nblk), dim3(256), 0, c->stream, mem + offE,
                         cells + 32 * ncell, nx, ny, nz, (int64_t)ncell);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
      (void)hipFree(cells);
      (void)hipFree(mem);
      return set_err(c, RAPT_E_HIP, "grid cell layout failed");
    }
  } else {
    (void)hipGetLastError();  // a failed allocation only selects the point layout
    cells = nullptr;
  }
  (void)hipFree(c->grid_mem);
  (void)hipFree(c->grid_cells);
  c->grid_mem = mem;
  c->grid_cells = cells;
  GridC& G = c->grid;
  G = GridC{};
  for (int k = 0; k < 4; ++k) {
    G.ax[k] = mem + off[k];
    G.n[k] = n[k];
    if (k == 3 && nt == 1) continue;
    G.lo[k] = ax[k][0];
    G.hi[k] = ax[k][n[k] - 1];
    // uniform spacing: every coordinate within 1e-12 of the spacing of a0 + i h
    const double h0 = (ax[k][n[k] - 1] - ax[k][0]) / (n[k] - 1);
    bool uni = true;
    for (int32_t i = 0; i < n[k] && uni; ++i)
      uni = fabs(ax[k][i] - (ax[k][0] + i * h0)) <= 1e-12 * fabs(h0) * n[k];
    if (uni) {
      G.uniform |= 1 << k;
      G.a0[k] = ax[k][0];
      G.ida[k] = 1.0 / h0;
    }
  }
  if (cells) {
    G.brick = 1;
    G.B = cells;
    G.E = E ? cells + 32 * ncell : nullptr;
  } else {
    G.B = mem + offB;
    G.E = E ? mem + offE : nullptr;
  }
  c->grid_set = true;
  if (c->field_set && c->field.model == RAPT_FIELD_GRID) c->field.grid = G;
  return RAPT_OK;
}

int rapt_rhs_eval(rapt_ctx* c, int32_t mode, int32_t eom, int32_t enforce_equatorial, int64_t n,
                  const double* ty, const double* k, double* out) {
  if (!c || n < 0 || (n && (!ty || !k || !out))) return RAPT_E_ARG;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  if (mode != RAPT_MODE_PARTICLE && mode != RAPT_MODE_GC)
    return set_err(c, RAPT_E_ARG, "rhs: mode must be PARTICLE or GC");
  if (eom < 0 || eom > 2) return set_err(c, RAPT_E_ARG, "bad eom %d", eom);
  if (n == 0) return RAPT_OK;
  const int wi = mode == RAPT_MODE_PARTICLE ? 7 : 5, wk = mode == RAPT_MODE_PARTICLE ? 3 : 4,
            wo = mode == RAPT_MODE_PARTICLE ? 6 : 4;
  HIP_OK(c, hipSetDevice(c->device));
  double* mem = nullptr;
  if (hipMalloc(&mem, sizeof(double) * (size_t)(wi + wk + wo) * n) != hipSuccess)
    return set_err(c, RAPT_E_OOM, "rhs: allocation failed");
  double *dty = mem, *dk = mem + (size_t)wi * n, *dout = mem + (size_t)(wi + wk) * n;
  int rc = RAPT_OK;
  if (hipMemcpyAsync(dty, ty, sizeof(double) * wi * n, hipMemcpyHostToDevice, c->stream) !=
          hipSuccess ||
      hipMemcpyAsync(dk, k, sizeof(double) * wk * n, hipMemcpyHostToDevice, c->stream) !=
          hipSuccess) {
    rc = set_err(c, RAPT_E_HIP, "rhs: H2D copy failed");
  } else {
    const int kind = mode == RAPT_MODE_PARTICLE ? KIND_PARTICLE : KIND_GC;
    const int eq = enforce_equatorial ? 1 : 0;
    switch (c->field.model) {
      case 0: launch_rhs_eval<0>(c, kind, eom, eq, n, dty, dk, dout); break;
      case 1: launch_rhs_eval<1>(c, kind, eom, eq, n, dty, dk, dout); break;
      case 2: launch_rhs_eval<2>(c, kind, eom, eq, n, dty, dk, dout); break;
      case 3: launch_rhs_eval<3>(c, kind, eom, eq, n, dty, dk, dout); break;
      case 4: launch_rhs_eval<4>(c, kind, eom, eq, n, dty, dk, dout); break;
      case 5: launch_rhs_eval<5>(c, kind, eom, eq, n, dty, dk, dout); break;
      case 6: launch_rhs_eval<6>(c, kind, eom, eq, n, dty, dk, dout); break;
      case 7: launch_rhs_eval<7>(c, kind, eom, eq, n, dty, dk, dout); break;
      case 8: launch_rhs_eval<8>(c, kind, eom, eq, n, dty, dk, dout); break;
      case 9: launch_rhs_eval<9>(c, kind, eom, eq, n, dty, dk, dout); break;
      case 10:
        if (c->expr_mode == RAPT_EXPR_COMPILED) {  // compiled at run time
          FieldC f = c->field;
          int32_t km = kind, e = eom, q = eq;
          int64_t nn = n;
          const double *a0 = dty, *a1 = dk;
          double* a2 = dout;
          void* args[] = {&f, &km, &e, &q, &nn, &a0, &a1, &a2};
          rc = rapt_jit_run_slot(c, RAPT_JIT_RHS_EVAL, (unsigned)((n + 255) / 256), 256, args);
        } else {
          launch_rhs_eval<10>(c, kind, eom, eq, n, dty, dk, dout);
        }
        break;
    }
    if (rc == RAPT_OK && (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, dout, sizeof(double) * wo * n, hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess))
      rc = set_err(c, RAPT_E_HIP, "rhs evaluation failed");
  }
  (void)hipFree(mem);
  return rc;
}

int rapt_field_eval(rapt_ctx* c, int what, int64_t n, const double* tpos, double* out) {
  if (!c || (!tpos && n) || (!out && n) || n < 0) return RAPT_E_ARG;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  const int w = field_width(what);
  if (w < 0) return set_err(c, RAPT_E_ARG, "bad field quantity %d", what);
  if (n == 0) return RAPT_OK;
  HIP_OK(c, hipSetDevice(c->device));
  double *dt = nullptr, *dout = nullptr;
  HIP_OK(c, hipMalloc(&dt, sizeof(double) * 4 * n));
  if (hipMalloc(&dout, sizeof(double) * w * n) != hipSuccess) {
    (void)hipFree(dt);
    return set_err(c, RAPT_E_OOM, "hipMalloc failed");
  }
  int rc = RAPT_OK;
  if (hipMemcpyAsync(dt, tpos, sizeof(double) * 4 * n, hipMemcpyHostToDevice, c->stream) !=
      hipSuccess) {
    rc = set_err(c, RAPT_E_HIP, "H2D copy failed");
  } else {
    switch (c->field.model) {
      case 0: launch_field_eval<0>(c, what, n, dt, dout, w); break;
      case 1: launch_field_eval<1>(c, what, n, dt, dout, w); break;
      case 2: launch_field_eval<2>(c, what, n, dt, dout, w); break;
      case 3: launch_field_eval<3>(c, what, n, dt, dout, w); break;
      case 4: launch_field_eval<4>(c, what, n, dt, dout, w); break;
      case 5: launch_field_eval<5>(c, what, n, dt, dout, w); break;
      case 6: launch_field_eval<6>(c, what, n, dt, dout, w); break;
      case 7: launch_field_eval<7>(c, what, n, dt, dout, w); break;
      case 8: launch_field_eval<8>(c, what, n, dt, dout, w); break;
      case 9: launch_field_eval<9>(c, what, n, dt, dout, w); break;
      case 10:
        if (c->expr_mode == RAPT_EXPR_COMPILED) {  // compiled at run time
          FieldC f = c->field;
          int wh = what, ww = w;
          int64_t nn = n;
          const double* a0 = dt;
          double* a1 = dout;
          void* args[] = {&f, &wh, &nn, &a0, &a1, &ww};
          rc = rapt_jit_run_slot(c, RAPT_JIT_FIELD_EVAL, (unsigned)((n + 255) / 256), 256, args);
        } else {
          launch_field_eval<10>(c, what, n, dt, dout, w);
        }
        break;
    }
    // (rc set: compiling or loading the user field's kernel failed)
    if (rc == RAPT_OK && hipGetLastError() != hipSuccess) {
      rc = set_err(c, RAPT_E_HIP, "field kernel launch failed");
    } else if (rc == RAPT_OK &&
               (hipMemcpyAsync(out, dout, sizeof(double) * w * n, hipMemcpyDeviceToHost,
                               c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess)) {
      rc = set_err(c, RAPT_E_HIP, "field eval failed");
    }
  }
  (void)hipFree(dt);
  (void)hipFree(dout);
  return rc;
}

int rapt_batch_create(rapt_ctx* c, int64_t n, rapt_batch** out) {
  if (!c || !out || n <= 0) return c ? set_err(c, RAPT_E_ARG, "bad batch size") : RAPT_E_ARG;
  HIP_OK(c, hipSetDevice(c->device));
  rapt_batch* b = new (std::nothrow) rapt_batch();
  if (!b) return RAPT_E_OOM;
  b->ctx = c;
  b->n = n;
  bool ok = hipMalloc(&b->d_state, sizeof(double) * RAPT_NSLOT * n) == hipSuccess &&
            hipMalloc(&b->d_state0, sizeof(double) * RAPT_NSLOT * n) == hipSuccess &&
            hipMalloc(&b->d_flags0, sizeof(int32_t) * n) == hipSuccess &&
            hipMalloc(&b->d_flags, sizeof(int32_t) * n) == hipSuccess &&
            hipMalloc(&b->d_status, sizeof(int32_t) * n) == hipSuccess &&
            hipMalloc(&b->d_nrows, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_naccpt, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_rowoff, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_rowcap, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_order, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_work, sizeof(unsigned long long) * 8) == hipSuccess;
  if (!ok) {
    rapt_batch_destroy(b);
    return set_err(c, RAPT_E_OOM, "device allocation of a %lld-particle batch failed",
                   (long long)n);
  }
  *out = b;
  return RAPT_OK;
}

int rapt_batch_destroy(rapt_batch* b) {
  if (!b) return RAPT_E_ARG;
  if (b->ctx) (void)hipSetDevice(b->ctx->device);
  (void)hipFree(b->d_state);
  (void)hipFree(b->d_state0);
  (void)hipFree(b->d_flags0);
  (void)hipFree(b->d_flags);
  (void)hipFree(b->d_status);
  (void)hipFree(b->d_nrows);
  (void)hipFree(b->d_naccpt);
  (void)hipFree(b->d_rowoff);
  (void)hipFree(b->d_rowcap);
  (void)hipFree(b->d_order);
  (void)hipFree(b->d_rows);
  rapt_batch_epochs_end(b);
  (void)hipFree(b->d_work);
  delete b;
  return RAPT_OK;
}

int rapt_batch_upload(rapt_batch* b, const double* state, const int32_t* flags) {
  if (!b || !state || !flags) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  HIP_OK(c, hipMemcpyAsync(b->d_state, state, sizeof(double) * RAPT_NSLOT * b->n,
                           hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, hipMemcpyAsync(b->d_flags, flags, sizeof(int32_t) * b->n, hipMemcpyHostToDevice,
                           c->stream));
  HIP_OK(c, hipMemcpyAsync(b->d_state0, b->d_state, sizeof(double) * RAPT_NSLOT * b->n,
                           hipMemcpyDeviceToDevice, c->stream));
  HIP_OK(c, hipMemcpyAsync(b->d_flags0, b->d_flags, sizeof(int32_t) * b->n,
                           hipMemcpyDeviceToDevice, c->stream));
  // Work order: shortest output step first.  For guiding centres the output
  // step is a fixed fraction of the bounce period and the step count scales
  // with delta / bounce period, so this starts the longest traces first and
  // puts traces of similar length into the same wavefront.  Entries without
  // an output step (0: computed in the kernel) keep their input order after
  // the others.
  std::vector<int64_t> order((size_t)b->n);
  const double* dts = state + (size_t)RAPT_S_DT * b->n;
  for (int64_t i = 0; i < b->n; ++i) order[(size_t)i] = i;
  std::stable_sort(order.begin(), order.end(), [dts](int64_t a, int64_t q) {
    const double ka = dts[a] > 0.0 ? dts[a] : HUGE_VAL, kq = dts[q] > 0.0 ? dts[q] : HUGE_VAL;
    return ka < kq;
  });
  HIP_OK(c, hipMemcpyAsync(b->d_order, order.data(), sizeof(int64_t) * b->n,
                           hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return RAPT_OK;
}

int rapt_batch_set_cost(rapt_batch* b, const double* cost) {
  if (!b || !cost) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  // Longest-processing-time-first: the most expensive traces start first,
  // cheap ones fill the queue's end, and traces of similar cost share a
  // wavefront.  Ties keep the input order.
  std::vector<int64_t> order((size_t)b->n);
  for (int64_t i = 0; i < b->n; ++i) order[(size_t)i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [cost](int64_t a, int64_t q) { return cost[a] > cost[q]; });
  HIP_OK(c, hipMemcpyAsync(b->d_order, order.data(), sizeof(int64_t) * b->n,
                           hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return RAPT_OK;
}

int rapt_batch_reset(rapt_batch* b) {
  if (!b) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  HIP_OK(c, hipMemcpyAsync(b->d_state, b->d_state0, sizeof(double) * RAPT_NSLOT * b->n,
                           hipMemcpyDeviceToDevice, c->stream));
  HIP_OK(c, hipMemcpyAsync(b->d_flags, b->d_flags0, sizeof(int32_t) * b->n,
                           hipMemcpyDeviceToDevice, c->stream));
  return RAPT_OK;
}

int rapt_batch_set_rows(rapt_batch* b, int32_t width, const int64_t* row_cap) {
  if (!b || !row_cap || width <= 0 || width > 16) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * b->n);
  if (!off) return set_err(c, RAPT_E_OOM, "host allocation failed");
  int64_t tot = 0;
  for (int64_t i = 0; i < b->n; ++i) {
    if (row_cap[i] < 0) {
      free(off);
      return set_err(c, RAPT_E_ARG, "negative row capacity");
    }
    off[i] = tot;
    tot += row_cap[i];
  }
  if (tot * width != b->total_rows * b->width || !b->d_rows) {
    (void)hipFree(b->d_rows);
    b->d_rows = nullptr;
    if (tot > 0 && hipMalloc(&b->d_rows, sizeof(double) * width * tot) != hipSuccess) {
      free(off);
      b->total_rows = 0;
      return set_err(c, RAPT_E_OOM, "trajectory buffer of %lld rows failed", (long long)tot);
    }
  }
  b->total_rows = tot;
  b->width = width;
  hipError_t e1 = hipMemcpyAsync(b->d_rowoff, off, sizeof(int64_t) * b->n,
                                 hipMemcpyHostToDevice, c->stream);
  hipError_t e2 = hipMemcpyAsync(b->d_rowcap, row_cap, sizeof(int64_t) * b->n,
                                 hipMemcpyHostToDevice, c->stream);
  hipError_t e3 = hipStreamSynchronize(c->stream);
  free(off);
  if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess)
    return set_err(c, RAPT_E_HIP, "row table upload failed");
  return RAPT_OK;
}

}  // extern "C"

// Validates a descriptor and builds the kernel arguments of a batch run.
static int prepare_run(rapt_batch* b, const rapt_trace_desc* d, KArgs& A, int& block_out,
                       int64_t& lanes_out, bool& grouped_out) {
  if (!b || !d) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  if (d->mode < 0 || d->mode > 2) return set_err(c, RAPT_E_ARG, "bad mode %d", d->mode);
  const int width = d->mode == RAPT_MODE_PARTICLE ? 7 : (d->mode == RAPT_MODE_GC ? 5 : 8);
  if (b->width != width)
    return set_err(c, RAPT_E_STATE, "row width %d does not match mode %d", b->width, d->mode);
  if (d->eom < 0 || d->eom > 2) return set_err(c, RAPT_E_ARG, "bad eom %d", d->eom);
  if (d->stepper < 0 || d->stepper > 2) return set_err(c, RAPT_E_ARG, "bad stepper");
  if (d->mode == RAPT_MODE_PARTICLE && d->stepper == RAPT_STEP_DOPRI5)
    return set_err(c, RAPT_E_ARG, "PARTICLE mode steps with DOP853 or RKF45");
  if (d->mode == RAPT_MODE_GC && d->stepper == RAPT_STEP_DOP853)
    return set_err(c, RAPT_E_ARG, "GC mode steps with DOPRI5 or RKF45");
  if (d->lanes_per_particle < 0 || (d->lanes_per_particle > 5 && d->lanes_per_particle != 8))
    return set_err(c, RAPT_E_ARG, "lanes_per_particle must be 0-5 or 8");
  if (d->mixed_quads < 0 || d->mixed_octets < 0)
    return set_err(c, RAPT_E_ARG, "mixed_quads and mixed_octets must be >= 0");
  if (d->work_order != 0 && d->work_order != 1)
    return set_err(c, RAPT_E_ARG, "work_order must be 0 or 1");
  if (d->precision != RAPT_PREC_FP64 && d->precision != RAPT_PREC_FP32_FIELD)
    return set_err(c, RAPT_E_ARG, "precision must be RAPT_PREC_FP64 or RAPT_PREC_FP32_FIELD");
  if (d->precision == RAPT_PREC_FP32_FIELD && c->field.model == RAPT_FIELD_GRID)
    return set_err(c, RAPT_E_ARG, "gridded fields are traced in FP64 only");
  if (d->precision == RAPT_PREC_FP32_FIELD && c->field.model == RAPT_FIELD_EXPR)
    return set_err(c, RAPT_E_ARG, "user-written fields are traced in FP64 only");
  if (d->mode == RAPT_MODE_ADAPTIVE && !(d->gctimestep > 0.0))
    return set_err(c, RAPT_E_ARG, "ADAPTIVE needs params['GCtimestep'] > 0");
  if (d->mode == RAPT_MODE_PARTICLE && !(d->cyclotronresolution > 0.0))
    return set_err(c, RAPT_E_ARG, "cyclotronresolution must be > 0");
  // rkf (rapt/rkf.py:13-143) takes any hmin; 0 never stops it, which the
  // fixed-step mode (tol = inf) needs: the step that ends an output interval
  // can be a rounding remainder far below any positive hmin
  if (d->stepper == RAPT_STEP_RKF45 && !(d->rkf_hmax > 0.0 && d->rkf_hmin >= 0.0))
    return set_err(c, RAPT_E_ARG, "RKF45 needs rkf_hmax > 0 and rkf_hmin >= 0");
  if (d->block > 0 && (d->block % 64 != 0 || d->block > 256))
    return set_err(c, RAPT_E_ARG, "block must be 64..256, x64");
  HIP_OK(c, hipSetDevice(c->device));
  memset(&A, 0, sizeof A);
  A.F = c->field;
  A.kind = d->mode;
  A.eom = d->eom;
  A.stepper = d->stepper;
  A.nmax = d->nsteps > 0 ? d->nsteps : 500;
  A.eq = d->enforce_equatorial;
  A.check = d->check_adiabaticity;
  A.width = width;
  A.rtol = d->rtol;
  A.atol = d->atol;
  A.gc_rtol = d->gc_rtol > 0 ? d->gc_rtol : d->rtol;
  A.gc_atol = d->gc_atol > 0 ? d->gc_atol : d->atol;
  if (d->mode == RAPT_MODE_ADAPTIVE) {
    // Particle segments: scipy dop853 as set by Particle.advance; GC segments: dopri5 defaults
    A.c853 = ctl(d->safety, d->dfactor, d->ifactor, d->beta, true);
    A.c5 = ctl(0.9, 0.2, 10.0, 0.0, false);
  } else {
    A.c853 = ctl(d->safety, d->dfactor, d->ifactor, d->beta, true);
    A.c5 = ctl(d->safety, d->dfactor, d->ifactor, d->beta, false);
  }
  if (d->precision == RAPT_PREC_FP32_FIELD) {
    // Float forces give an error estimate with a rounding floor that grows
    // only like h; the PI term (beta > 0) then settles at err ~ 1e-4 and
    // shrinks h without bound.  FP32-field runs use the I-controller.
    A.c853.beta = 0.0;
    A.c853.expo1 = 1.0 / 8.0;
    A.c5.beta = 0.0;
    A.c5.expo1 = 0.2;
  }
  A.max_step = d->max_step;
  A.first_step = d->first_step;
  A.rkf_tol = d->rkf_tol;
  A.rkf_hmax = d->rkf_hmax;
  A.rkf_hmin = d->rkf_hmin;
  A.cycres = d->cyclotronresolution;
  A.gcdt = d->gctimestep;
  A.epss = d->epss;
  A.epst = d->epst;
  A.delta = d->delta;
  A.n = b->n;
  A.state = b->d_state;
  A.flags = b->d_flags;
  A.status = b->d_status;
  A.nrows = b->d_nrows;
  A.naccpt = b->d_naccpt;
  A.row_off = b->d_rowoff;
  A.row_cap = b->d_rowcap;
  A.rows = b->d_rows;
  A.work = b->d_work;
  A.order = d->work_order == 0 ? b->d_order : nullptr;
  // work queues of the grouped kernels: one queue [0, n) for a single layout;
  // for the mixed layouts [0, split) octets, [split, split2) quads and
  // [split2, n) pairs (lanes_per_particle 3: no octets, at least one quad)
  A.split = b->n;
  A.split2 = b->n;
  if (d->lanes_per_particle == 3 || d->lanes_per_particle == 5) {
    const int64_t n = b->n;
    int64_t ko = 0;
    if (d->lanes_per_particle == 5) {
      ko = d->mixed_octets > 0 ? d->mixed_octets : n / 64;
      ko = ko < 1 ? 1 : (ko > n ? n : ko);
    }
    int64_t kq = d->mixed_quads > 0 ? d->mixed_quads : n / 8;
    const int64_t kq_min = d->lanes_per_particle == 3 ? 1 : 0;
    kq = kq < kq_min ? kq_min : (kq > n - ko ? n - ko : kq);
    A.split = ko;
    A.split2 = ko + kq;
  }
  int64_t lanes = d->lanes > 0 ? d->lanes : b->n;
  if (lanes > b->n) lanes = b->n;
  if (lanes < 1) lanes = 1;
  // the grouped kernels (two, four or eight lanes, mixed) are FP64 only; an
  // interpreted user-written field runs one lane per particle (the grouped
  // kernels keep too many registers live across the interpreter's calls:
  // 1-1.2 KB of spills per lane), a compiled one takes every layout
  const bool grouped = d->mode == RAPT_MODE_GC && d->stepper == RAPT_STEP_DOPRI5 &&
                       d->lanes_per_particle != 1 && d->precision == RAPT_PREC_FP64 &&
                       (c->field.model != RAPT_FIELD_EXPR ||
                        c->expr_mode == RAPT_EXPR_COMPILED);
  // Default block: 256 lanes, except for launches of fewer than four waves
  // per CU, which take 64 so that the waves spread over every CU (a CU's
  // waves share its scalar unit and instruction fetch): 1,024 Particle-mode
  // protons advance in 39-43 ms instead of 52, 16,384 in 48 instead of 63;
  // at 65,536 (four waves per CU) the two are equal for the analytic dipole
  // and 64 was slower on a gridded one.  The grouped kernels count the lanes
  // of all their groups.
  int block = d->block > 0 ? d->block : 0;
  if (block == 0) {
    int64_t launched = lanes;
    if (grouped) {
      const int lpp = d->lanes_per_particle;
      if (lpp == 3 || lpp == 5) {
        launched = 8 * A.split + 4 * (A.split2 - A.split) + 2 * (b->n - A.split2);
        if (d->lanes > 0 && d->lanes < launched) launched = d->lanes;
      } else {
        const int L = lpp == 2 ? 2 : (lpp == 8 ? 8 : 4);
        int64_t groups = d->lanes > 0 ? (d->lanes + L - 1) / L : b->n;
        launched = (groups < b->n ? groups : b->n) * L;
      }
    }
    int cus = 0;
    HIP_OK(c, hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    block = launched < (int64_t)4 * 64 * cus ? 64 : 256;
  }
  block_out = block;
  lanes_out = lanes;
  grouped_out = grouped;
  return RAPT_OK;
}

// Compiles (once per program) and loads the run-time tracer kernels of a
// compiled user-written field before a launch that uses them; not called by
// the allocation-free entry point (loading a module allocates).
static int prepare_jit(rapt_ctx* c, const rapt_trace_desc* d, const KArgs& A, bool grouped) {
  if (c->field.model != RAPT_FIELD_EXPR || c->expr_mode != RAPT_EXPR_COMPILED ||
      d->precision != RAPT_PREC_FP64)
    return RAPT_OK;
  if (grouped) return rapt_jit_prepare_gcn(c, d->lanes_per_particle, trace_use_sp<F_EXPR>(A));
  const int kind = d->mode == RAPT_MODE_PARTICLE ? KIND_PARTICLE
                                                 : (d->mode == RAPT_MODE_GC ? KIND_GC
                                                                            : KIND_ADAPTIVE);
  return rapt_jit_prepare(c, kind, kind == KIND_GC && trace_use_sp<F_EXPR>(A));
}

// The tracer kernel of a prepared run on the context's stream.
static void launch_tracer(rapt_batch* b, const rapt_trace_desc* d, const KArgs& A, int block,
                          int64_t lanes, bool grouped) {
  rapt_ctx* c = b->ctx;
  if (grouped) {
    if (d->lanes_per_particle == 3 || d->lanes_per_particle == 5) {
      // octets and quads for their queues; pairs for the rest, all resident
      // unless d->lanes caps the resident lanes (all groups' together): then
      // fewer pair groups refill from the queues
      const int64_t octets = A.split, quads = A.split2 - A.split;
      int64_t pairs = b->n - A.split2;
      if (d->lanes > 0) {
        const int64_t p = (d->lanes - 8 * octets - 4 * quads) / 2;
        pairs = p < 1 ? 1 : (p < pairs ? p : pairs);
      }
      if (c->field.model == RAPT_FIELD_EXPR)  // compiled (prepare_jit loaded it)
        rapt_jit_launch_gcn(c, d->lanes_per_particle, trace_use_sp<F_EXPR>(A), A, block, octets,
                            quads, pairs);
      else
        rapt_launch_gcn(c->stream, c->field.model, d->lanes_per_particle, A, block, octets,
                        quads, pairs);
    } else {
      const int L = d->lanes_per_particle == 2 ? 2 : (d->lanes_per_particle == 8 ? 8 : 4);
      int64_t groups = d->lanes > 0 ? (d->lanes + L - 1) / L : b->n;
      if (groups > b->n) groups = b->n;
      if (c->field.model == RAPT_FIELD_EXPR)
        rapt_jit_launch_gcn(c, L, trace_use_sp<F_EXPR>(A), A, block, groups, 0, 0);
      else
        rapt_launch_gcn(c->stream, c->field.model, L, A, block, groups, 0, 0);
    }
  } else if (d->precision == RAPT_PREC_FP32_FIELD) {
    const int kind = d->mode == RAPT_MODE_PARTICLE ? KIND_PARTICLE
                                                   : (d->mode == RAPT_MODE_GC ? KIND_GC
                                                                              : KIND_ADAPTIVE);
    rapt_launch_trace_f32(c->stream, c->field.model, kind, A, block, lanes);
  } else {
    switch (d->mode) {
      case RAPT_MODE_PARTICLE: launch_model<KIND_PARTICLE>(c, A, block, lanes); break;
      case RAPT_MODE_GC: launch_model<KIND_GC>(c, A, block, lanes); break;
      default: launch_model<KIND_ADAPTIVE>(c, A, block, lanes); break;
    }
  }
}

extern "C" {

int rapt_batch_run(rapt_batch* b, const rapt_trace_desc* d, rapt_counters* cnt) {
  if (!b || !d) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  KArgs A;
  int block = 256;
  int64_t lanes = 0;
  bool grouped = false;
  int rc = prepare_run(b, d, A, block, lanes, grouped);
  if (!rc) rc = prepare_jit(c, d, A, grouped);
  if (rc) return rc;
  HIP_OK(c, hipMemsetAsync(b->d_work, 0, sizeof(unsigned long long) * 8, c->stream));
  HIP_OK(c, hipEventRecord(c->ev0, c->stream));
  launch_tracer(b, d, A, block, lanes, grouped);
  HIP_OK(c, hipGetLastError());
  HIP_OK(c, hipEventRecord(c->ev1, c->stream));
  unsigned long long w[8];
  HIP_OK(c, hipMemcpyAsync(w, b->d_work, sizeof w, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  if (cnt) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    cnt->naccpt = (int64_t)w[1];
    cnt->nrejct = (int64_t)w[2];
    cnt->nstep = (int64_t)w[3];
    cnt->nfcn = (int64_t)w[4];
    cnt->nswitch = (int64_t)w[5];
    cnt->nrows = 0;
    cnt->kernel_ms = ms;
  }
  return RAPT_OK;
}

int rapt_batch_download(rapt_batch* b, double* state, int32_t* flags, int32_t* status,
                        int64_t* nrows, double* rows, int64_t* naccpt) {
  if (!b) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  const int64_t n = b->n;
  if (state)
    HIP_OK(c, hipMemcpyAsync(state, b->d_state, sizeof(double) * RAPT_NSLOT * n,
                             hipMemcpyDeviceToHost, c->stream));
  if (flags)
    HIP_OK(c, hipMemcpyAsync(flags, b->d_flags, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                             c->stream));
  if (status)
    HIP_OK(c, hipMemcpyAsync(status, b->d_status, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                             c->stream));
  if (nrows)
    HIP_OK(c, hipMemcpyAsync(nrows, b->d_nrows, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                             c->stream));
  if (naccpt)
    HIP_OK(c, hipMemcpyAsync(naccpt, b->d_naccpt, sizeof(int64_t) * n, hipMemcpyDeviceToHost,
                             c->stream));
  if (rows && b->total_rows > 0)
    HIP_OK(c, hipMemcpyAsync(rows, b->d_rows, sizeof(double) * b->width * b->total_rows,
                             hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return RAPT_OK;
}

int rapt_trace_device_noalloc(rapt_ctx* c, const rapt_trace_desc* d, int64_t n,
                              double* state, int32_t* flags, const int64_t* row_off,
                              const int64_t* row_cap, double* rows, int64_t* nrows,
                              int32_t* status, int64_t* naccpt, const int64_t* order,
                              rapt_counters* cnt) {
  if (!c || !d) return RAPT_E_ARG;
  if (n <= 0 || !state || !flags || !row_off || !row_cap || !rows || !nrows || !status ||
      !naccpt)
    return set_err(c, RAPT_E_ARG, "noalloc: n > 0 and every device buffer are required");
  if (!c->noalloc_scratch) return set_err(c, RAPT_E_STATE, "noalloc: no scratch on this device");
  // the kernels use counters 0-7 (queue 0, five step counters, the mixed
  // layouts' queues 1 and 2): the scratch must hold them all
  static_assert(RAPT_NOALLOC_WORDS >= 8, "scratch too small for the work counters");
  // a view of the caller's buffers with the batch layout the launch code reads
  // (nothing here is owned or freed)
  rapt_batch view;
  view.ctx = c;
  view.n = n;
  view.d_state = state;
  view.d_flags = flags;
  view.d_status = status;
  view.d_nrows = nrows;
  view.d_naccpt = naccpt;
  view.d_rowoff = const_cast<int64_t*>(row_off);
  view.d_rowcap = const_cast<int64_t*>(row_cap);
  view.d_order = const_cast<int64_t*>(order);
  view.d_rows = rows;
  view.width = d->mode == RAPT_MODE_PARTICLE ? 7 : (d->mode == RAPT_MODE_GC ? 5 : 8);
  view.d_work = c->noalloc_scratch;
  rapt_trace_desc dd = *d;
  if (!order) dd.work_order = 1;  // no order given: input order
  // a user-written field: one lane per particle (compiled if an earlier run
  // loaded that kernel, else interpreted; loading a module allocates)
  if (c->field.model == RAPT_FIELD_EXPR) dd.lanes_per_particle = 1;
  KArgs A;
  int block = 256;
  int64_t lanes = 0;
  bool grouped = false;
  const int rc = prepare_run(&view, &dd, A, block, lanes, grouped);
  if (rc) return rc;
  // the scratch keeps the last call's counters: clear them
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, c->stream, c->noalloc_scratch,
                     (int)RAPT_NOALLOC_WORDS);
  HIP_OK(c, hipGetLastError());
  HIP_OK(c, hipEventRecord(c->ev0, c->stream));
  launch_tracer(&view, &dd, A, block, lanes, grouped);
  HIP_OK(c, hipGetLastError());
  HIP_OK(c, hipEventRecord(c->ev1, c->stream));
  unsigned long long w[RAPT_NOALLOC_WORDS];
  HIP_OK(c, hipMemcpyAsync(w, c->noalloc_scratch, sizeof w, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  if (cnt) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    cnt->naccpt = (int64_t)w[1];
    cnt->nrejct = (int64_t)w[2];
    cnt->nstep = (int64_t)w[3];
    cnt->nfcn = (int64_t)w[4];
    cnt->nswitch = (int64_t)w[5];
    cnt->nrows = 0;
    cnt->kernel_ms = ms;
  }
  return RAPT_OK;
}

int rapt_trace(rapt_ctx* c, const rapt_trace_desc* d, int64_t n, double* state, int32_t* flags,
               const int64_t* row_cap, double* rows, int64_t* nrows, int32_t* status,
               rapt_counters* cnt) {
  if (!c || !d || !state || !flags || !row_cap || !nrows || !status || n <= 0)
    return RAPT_E_ARG;
  rapt_batch* b = nullptr;
  int rc = rapt_batch_create(c, n, &b);
  if (rc) return rc;
  const int width = d->mode == RAPT_MODE_PARTICLE ? 7 : (d->mode == RAPT_MODE_GC ? 5 : 8);
  if (!(rc = rapt_batch_upload(b, state, flags)) &&
      !(rc = rapt_batch_set_rows(b, width, row_cap)) && !(rc = rapt_batch_run(b, d, cnt)))
    rc = rapt_batch_download(b, state, flags, status, nrows, rows, nullptr);
  rapt_batch_destroy(b);
  return rc;
}


int rapt_batch_epochs_end(rapt_batch* b) {
  if (!b) return RAPT_E_ARG;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  (void)hipFree(b->d_erows);
  (void)hipFree(b->d_eoff);
  (void)hipFree(b->d_ecap);
  (void)hipFree(b->d_left);
  (void)hipHostFree(b->h_rows);
  (void)hipHostFree(b->h_nrows);
  (void)hipHostFree(b->h_status);
  (void)hipHostFree(b->h_left);
  (void)hipHostFree(b->h_work);
  b->d_erows = nullptr;
  b->d_eoff = b->d_ecap = nullptr;
  b->d_left = nullptr;
  b->h_rows = nullptr;
  b->h_nrows = nullptr;
  b->h_status = nullptr;
  b->h_left = nullptr;
  b->h_work = nullptr;
  b->ep_rows = 0;
  return RAPT_OK;
}

int rapt_batch_epochs_begin(rapt_batch* b, const rapt_trace_desc* d, int32_t rows_per_epoch) {
  if (!b || !d) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  if (rows_per_epoch < 1) return set_err(c, RAPT_E_ARG, "rows_per_epoch must be >= 1");
  const int width = d->mode == RAPT_MODE_PARTICLE ? 7 : (d->mode == RAPT_MODE_GC ? 5 : 8);
  rapt_batch_epochs_end(b);
  HIP_OK(c, hipSetDevice(c->device));
  const int64_t n = b->n;
  const size_t nr = (size_t)n * rows_per_epoch;
  bool ok = hipMalloc(&b->d_erows, sizeof(double) * nr * width) == hipSuccess &&
            hipMalloc(&b->d_eoff, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_ecap, sizeof(int64_t) * n) == hipSuccess &&
            hipMalloc(&b->d_left, sizeof(int32_t)) == hipSuccess &&
            hipHostMalloc(&b->h_rows, sizeof(double) * nr * width, 0) == hipSuccess &&
            hipHostMalloc(&b->h_nrows, sizeof(int64_t) * n, 0) == hipSuccess &&
            hipHostMalloc(&b->h_status, sizeof(int32_t) * n, 0) == hipSuccess &&
            hipHostMalloc(&b->h_left, sizeof(int32_t), 0) == hipSuccess &&
            hipHostMalloc(&b->h_work, sizeof(unsigned long long) * 8, 0) == hipSuccess;
  if (!ok) {
    rapt_batch_epochs_end(b);
    return set_err(c, RAPT_E_OOM, "epoch buffers (%lld rows) do not fit", (long long)nr);
  }
  b->ep_rows = rows_per_epoch;
  {
    std::vector<int64_t> off((size_t)n), cap((size_t)n, rows_per_epoch);
    for (int64_t i = 0; i < n; ++i) off[(size_t)i] = i * rows_per_epoch;
    HIP_OK(c, hipMemcpy(b->d_eoff, off.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
    HIP_OK(c, hipMemcpy(b->d_ecap, cap.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
  }
  // the run's arguments, pointed at the epoch row buffer
  const int32_t w0 = b->width;
  b->width = width;
  KArgs A;
  int block = 256;
  int64_t lanes = 0;
  bool grouped = false;
  int rc = prepare_run(b, d, A, block, lanes, grouped);
  b->width = w0;
  if (!rc) rc = prepare_jit(c, d, A, grouped);  // before the capture: loading allocates
  if (rc) {
    rapt_batch_epochs_end(b);
    return rc;
  }
  A.rows = b->d_erows;
  A.row_off = b->d_eoff;
  A.row_cap = b->d_ecap;
  hipStream_t st = c->stream;
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rapt_batch_epochs_end(b);
    return set_err(c, RAPT_E_HIP, "graph capture failed to start");
  }
  hipLaunchKernelGGL(k_epoch_prep, dim3(1), dim3(64), 0, st, b->d_work, b->d_left);
  launch_tracer(b, d, A, block, lanes, grouped);
  hipLaunchKernelGGL(k_epoch_left, dim3(64), dim3(256), 0, st, n, b->d_status, b->d_left);
  (void)hipMemcpyAsync(b->h_rows, b->d_erows, sizeof(double) * nr * width, hipMemcpyDeviceToHost,
                       st);
  (void)hipMemcpyAsync(b->h_nrows, b->d_nrows, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st);
  (void)hipMemcpyAsync(b->h_status, b->d_status, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st);
  (void)hipMemcpyAsync(b->h_left, b->d_left, sizeof(int32_t), hipMemcpyDeviceToHost, st);
  (void)hipMemcpyAsync(b->h_work, b->d_work, sizeof(unsigned long long) * 8,
                       hipMemcpyDeviceToHost, st);
  if (hipStreamEndCapture(st, &b->graph) != hipSuccess ||
      hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
    rapt_batch_epochs_end(b);
    return set_err(c, RAPT_E_HIP, "graph capture failed");
  }
  return RAPT_OK;
}

int rapt_batch_epoch_next(rapt_batch* b, const double** rows, const int64_t** nrows,
                          const int32_t** status, int32_t* running, rapt_counters* cnt) {
  if (!b) return RAPT_E_ARG;
  rapt_ctx* c = b->ctx;
  if (!b->exec) return set_err(c, RAPT_E_STATE, "no epochs started (rapt_batch_epochs_begin)");
  HIP_OK(c, hipSetDevice(c->device));
  HIP_OK(c, hipEventRecord(c->ev0, c->stream));
  HIP_OK(c, hipGraphLaunch(b->exec, c->stream));
  HIP_OK(c, hipEventRecord(c->ev1, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  if (rows) *rows = b->h_rows;
  if (nrows) *nrows = b->h_nrows;
  if (status) *status = b->h_status;
  if (running) *running = *b->h_left;
  if (cnt) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    cnt->naccpt = (int64_t)b->h_work[1];
    cnt->nrejct = (int64_t)b->h_work[2];
    cnt->nstep = (int64_t)b->h_work[3];
    cnt->nfcn = (int64_t)b->h_work[4];
    cnt->nswitch = (int64_t)b->h_work[5];
    cnt->nrows = 0;
    cnt->kernel_ms = ms;
  }
  return RAPT_OK;
}

}  // extern "C"
