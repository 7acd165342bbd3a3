// rapt_jit.hip -- user-written fields compiled at run time.
//
// A _Field subclass that overrides B and/or E in Python (rapt/fields.py:5-74)
// reaches the library as a traced register program (rapt_field_expr_set).
// The built-in kernels interpret that program (expr_eval: one scalar
// dispatch per instruction, the registers in scratch).  Here the program
// becomes C++, one statement per instruction -- the same operation on the
// same values as the interpreter's case for its opcode, the constants read
// from the program's table in device memory -- and the one-lane tracer
// kernels k_trace<F_EXPR, kind, FP64> and the lane-group GuidingCenter
// kernels k_trace_gcn / k_trace_gcmix, the field-line kernels (half-bounce
// path, eye, surrogate labels, BounceCenter), field evaluation, right-hand
// sides and the row diagnostics are compiled for it with hipRTC from
// the library's own headers (embedded when the library is built), so the
// field is inline code like an analytic model's and the kernels run at two
// waves per SIMD.  The trajectories are the interpreter's one-lane kernel's,
// bit for bit, in every layout.
//
// Each kernel is compiled when a run first needs it (a few seconds).  Code
// objects are cached for the process by program text and kernel (the
// constants are not part of the text, so a changed constant table reuses
// the code), loaded modules by device as well; both live until the process
// ends.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <stdint.h>
#include <string.h>

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/rapt_hip.h"
#include "rapt_internal.hpp"
#include "rapt_jit_headers.inc"

using namespace rapt;

namespace {

// the instantiations the launch code uses for F_EXPR (FP64): one lane per
// particle for Particle, GuidingCenter (generic, and the static
// Tao-Chan-Brizard one) and Adaptive (slots 0-3, rapt_trace_kernel.hpp);
// lane pairs, quads and octets (4-9) and the mixed layout (10-11) for
// GuidingCenter (rapt_gcn_kernel.hpp); the field-line kernels (12-15,
// rapt_bounce_kernel.hpp); field evaluation and right-hand sides (16-17,
// rapt_trace_kernel.hpp) and the row diagnostics (18, rapt_diag_kernel.hpp)
static_assert(F_EXPR == 10 && KIND_PARTICLE == 0 && KIND_GC == 1 && KIND_ADAPTIVE == 2 &&
                  RAPT_PREC_FP64 == 0,
              "kernel names below spell these values");
constexpr int kSlots = RAPT_JIT_SLOTS;
constexpr int kFirstGroupSlot = 4;
constexpr int kFirstLineSlot = RAPT_JIT_HALFBOUNCE;
constexpr int kFirstEvalSlot = RAPT_JIT_FIELD_EVAL;
static_assert(RAPT_JIT_HALFBOUNCE == 12 && RAPT_JIT_LABELS == 13 && RAPT_JIT_BC == 14 &&
                  RAPT_JIT_BC_EQ == 15 && RAPT_JIT_FIELD_EVAL == 16 && RAPT_JIT_RHS_EVAL == 17 &&
                  RAPT_JIT_PARTICLE_GC == 18 && kSlots == 19,
              "slot numbers below");
const char* const kKernelNames[kSlots] = {
    "rapt::k_trace<10, 0, 0, false>", "rapt::k_trace<10, 1, 0, false>",
    "rapt::k_trace<10, 1, 0, true>",  "rapt::k_trace<10, 2, 0, false>",
    "rapt::k_trace_gcn<10, 2, false>", "rapt::k_trace_gcn<10, 2, true>",
    "rapt::k_trace_gcn<10, 4, false>", "rapt::k_trace_gcn<10, 4, true>",
    "rapt::k_trace_gcn<10, 8, false>", "rapt::k_trace_gcn<10, 8, true>",
    "rapt::k_trace_gcmix<10, false>",  "rapt::k_trace_gcmix<10, true>",
    "rapt::k_halfbounce<10>",          "rapt::k_bounce_labels<10>",
    "rapt::k_bounce_center<10>",       "rapt::k_bounce_center_eq<10>",
    "rapt::k_field_eval<10>",          "rapt::k_rhs_eval<10>",
    "rapt::k_particle_gc<10>"};

int jit_slot(int kind, bool sp) {
  return kind == KIND_PARTICLE ? 0 : (kind == KIND_GC ? (sp ? 2 : 1) : 3);
}
// lanes_per_particle as rapt_trace_desc has it (2, 8; 3 and 5 mixed; 0 and
// 4 quads)
int jit_group_slot(int lanes_per_particle, bool sp) {
  const int s = sp ? 1 : 0;
  switch (lanes_per_particle) {
    case 2: return 4 + s;
    case 8: return 8 + s;
    case 3:
    case 5: return 10 + s;
    default: return 6 + s;
  }
}

std::string reg(int i) { return "r" + std::to_string(i); }

// One program (instructions lo .. lo + n - 1) as the body of a function of
// t, x, y, z returning registers 4-6.  Opcodes as ExprC documents them; the
// program has passed rapt_field_expr_set's checks or the same ones here.
bool emit_body(const int32_t* code, int lo, int n, std::string& s) {
  s += "  double r0 = t, r1 = x, r2 = y, r3 = z";
  for (int i = 4; i < EXPR_NREG; ++i) s += ", " + reg(i) + " = 0.0";
  s += ";\n";
  for (int i = lo; i < lo + n; ++i) {
    const int op = code[4 * i], d = code[4 * i + 1], ia = code[4 * i + 2], ib = code[4 * i + 3];
    const std::string a = reg(ia), b = reg(ib);
    const char* f1 = nullptr;
    std::string e;
    switch (op) {
      case 0: e = "k[" + std::to_string(ia) + "]"; break;
      case 1: e = a + " + " + b; break;
      case 2: e = a + " - " + b; break;
      case 3: e = a + " * " + b; break;
      case 4: e = a + " / " + b; break;
      case 5: e = "-" + a; break;
      case 6: e = "expr_pow(" + a + ", " + b + ")"; break;
      case 7: f1 = "sqrt"; break;
      case 8: f1 = "exp"; break;
      case 9: f1 = "log"; break;
      case 10: f1 = "sin"; break;
      case 11: f1 = "cos"; break;
      case 12: f1 = "tan"; break;
      case 13: f1 = "tanh"; break;
      case 14: f1 = "sinh"; break;
      case 15: f1 = "cosh"; break;
      case 16: f1 = "atan"; break;
      case 17: f1 = "asin"; break;
      case 18: f1 = "acos"; break;
      case 19: f1 = "fabs"; break;
      case 20: e = "atan2(" + a + ", " + b + ")"; break;
      case 21: e = a; break;
      case 22: f1 = "log10"; break;
      case 23: e = "hypot(" + a + ", " + b + ")"; break;
      // numpy's maximum / minimum (a NaN propagates, ties give b) and sign
      case 24: e = "(" + a + " > " + b + " || " + a + " != " + a + ") ? " + a + " : " + b; break;
      case 25: e = "(" + a + " < " + b + " || " + a + " != " + a + ") ? " + a + " : " + b; break;
      case 26:
        e = a + " > 0.0 ? 1.0 : (" + a + " < 0.0 ? -1.0 : (" + a + " == 0.0 ? 0.0 : " + a + "))";
        break;
      default: return false;
    }
    if (f1) e = std::string(f1) + "(" + a + ")";
    s += "  " + reg(d) + " = " + e + ";\n";
  }
  s += "  return V3{r4, r5, r6};\n";
  return true;
}

// The whole translation unit of a kernel slot: the field functions between
// the field header and the kernel header that calls them.
bool make_source(int32_t nb, int32_t ne, const int32_t* code, int slot, std::string& src) {
  src = "#define RAPT_JIT_FIELD 1\n#include \"rapt_fields.hpp\"\nnamespace rapt {\n";
  const char* fn[2] = {"jit_B", "jit_E"};
  for (int p = 0; p < 2; ++p) {
    src += "__device__ __forceinline__ V3 " + std::string(fn[p]) +
           "(const double* kg, double t, double x, double y, double z) {\n"
           "  using CD = const __attribute__((address_space(4))) double;\n"
           "  CD* k = (CD*)kg;\n  (void)k;\n";
    const int lo = p == 0 ? 0 : nb, n = p == 0 ? nb : ne;
    if (n == 0) {
      src += "  return V3{0.0, 0.0, 0.0};\n";
    } else if (!emit_body(code, lo, n, src)) {
      return false;
    }
    src += "}\n";
  }
  src += "}  // namespace rapt\n";
  src += slot < kFirstGroupSlot ? "#include \"rapt_trace_kernel.hpp\"\n"
         : slot < kFirstLineSlot ? "#include \"rapt_gcn_kernel.hpp\"\n"
         : slot < kFirstEvalSlot ? "#include \"rapt_bounce_kernel.hpp\"\n"
         : slot == RAPT_JIT_PARTICLE_GC ? "#include \"rapt_diag_kernel.hpp\"\n"
                                        : "#include \"rapt_trace_kernel.hpp\"\n";
  return true;
}

struct JitCode {
  std::vector<char> obj;
  std::string lowered;  // the kernel's symbol
};
struct JitModule {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};

std::mutex g_mu;
std::map<std::string, JitCode> g_code;                    // key: arch, kernel, source
std::map<std::pair<int, std::string>, JitModule> g_mods;  // (device, key)

std::string cache_key(const std::string& arch, int slot, const std::string& src) {
  return arch + "\n" + kKernelNames[slot] + "\n" + src;
}

void put_log(const std::string& text, char* log, int64_t size) {
  if (!log || size <= 0) return;
  const size_t n = text.size() < (size_t)(size - 1) ? text.size() : (size_t)(size - 1);
  memcpy(log, text.data(), n);
  log[n] = 0;
}

// Compiles kernel `slot` of src for arch (cached); the caller holds g_mu.
// On failure msg holds the reason and the compiler's log.
int compile_locked(const std::string& arch, int slot, const std::string& src,
                   const JitCode** out, std::string& msg) {
  const std::string key = cache_key(arch, slot, src);
  auto it = g_code.find(key);
  if (it != g_code.end()) {
    *out = &it->second;
    return RAPT_OK;
  }
  hiprtcProgram prog = nullptr;
  if (hiprtcCreateProgram(&prog, src.c_str(), "rapt_user_field.hip", kJitHeaderCount,
                          kJitHeaderTexts, kJitHeaderNames) != HIPRTC_SUCCESS) {
    msg = "hiprtcCreateProgram failed";
    return RAPT_E_HIP;
  }
  hiprtcAddNameExpression(prog, kKernelNames[slot]);
  // the library's own device flags: no FMA contraction, no packed FP32
  const std::string archopt = "--offload-arch=" + arch;
  const char* opts[] = {archopt.c_str(), "-O3", "-std=c++17", "-ffp-contract=off", "-w",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  const hiprtcResult rc =
      hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  size_t ls = 0;
  std::string log;
  if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
    log.resize(ls);
    hiprtcGetProgramLog(prog, &log[0]);
  }
  if (rc != HIPRTC_SUCCESS) {
    msg = std::string("user-written field: run-time compilation failed (") +
          hiprtcGetErrorString(rc) + "): " + log;
    hiprtcDestroyProgram(&prog);
    return RAPT_E_HIP;
  }
  JitCode jc;
  size_t cs = 0;
  bool ok = hiprtcGetCodeSize(prog, &cs) == HIPRTC_SUCCESS && cs > 0;
  if (ok) {
    jc.obj.resize(cs);
    ok = hiprtcGetCode(prog, jc.obj.data()) == HIPRTC_SUCCESS;
  }
  if (ok) {
    const char* low = nullptr;
    ok = hiprtcGetLoweredName(prog, kKernelNames[slot], &low) == HIPRTC_SUCCESS && low;
    if (ok) jc.lowered = low;
  }
  hiprtcDestroyProgram(&prog);
  if (!ok) {
    msg = "user-written field: no code object or kernel names from hiprtc";
    return RAPT_E_HIP;
  }
  *out = &g_code.emplace(key, std::move(jc)).first->second;
  return RAPT_OK;
}

}  // namespace

// Compiles (or finds) and loads the tracer kernel of the context's program
// for a kind on its device.  Called where a launch may allocate (the module
// load does); the allocation-free entry point uses a kernel only once this
// has run.
static int prepare_slot(rapt_ctx* c, int slot) {
  if (c->jit_fn[slot]) return RAPT_OK;
  if (!c->expr_set || c->expr_code.empty())
    return set_err(c, RAPT_E_STATE, "user-written field: no program uploaded");
  hipDeviceProp_t prop;
  HIP_OK(c, hipGetDeviceProperties(&prop, c->device));
  // the processor without its feature flags (gfx950, not gfx950:sramecc+:...),
  // as the library itself is built: the code runs whatever the flags, and a
  // device-free rapt_field_expr_compile for "gfx950" fills the same cache entry
  std::string arch = prop.gcnArchName;
  arch = arch.substr(0, arch.find(':'));
  std::string src;
  if (!make_source(c->expr_nb, c->expr_ne, c->expr_code.data(), slot, src))
    return set_err(c, RAPT_E_ARG, "user-written field: unknown opcode");
  std::lock_guard<std::mutex> lock(g_mu);
  const JitCode* jc = nullptr;
  std::string msg;
  const int rc = compile_locked(arch, slot, src, &jc, msg);
  if (rc) return set_err(c, rc, "%s", msg.substr(0, 480).c_str());
  const auto key = std::make_pair(c->device, cache_key(arch, slot, src));
  auto it = g_mods.find(key);
  if (it == g_mods.end()) {
    HIP_OK(c, hipSetDevice(c->device));
    JitModule jm;
    HIP_OK(c, hipModuleLoadData(&jm.mod, jc->obj.data()));
    const hipError_t e = hipModuleGetFunction(&jm.fn, jm.mod, jc->lowered.c_str());
    if (e != hipSuccess) {
      (void)hipModuleUnload(jm.mod);
      return set_err(c, RAPT_E_HIP, "user-written field: kernel %s not in the module: %s",
                     kKernelNames[slot], hipGetErrorString(e));
    }
    it = g_mods.emplace(key, jm).first;
  }
  c->jit_fn[slot] = it->second.fn;
  return RAPT_OK;
}

int rapt_jit_prepare(rapt_ctx* c, int kind, bool sp) { return prepare_slot(c, jit_slot(kind, sp)); }

int rapt_jit_prepare_gcn(rapt_ctx* c, int lanes_per_particle, bool sp) {
  return prepare_slot(c, jit_group_slot(lanes_per_particle, sp));
}

// One-lane tracer launch of a compiled kernel, with the same grid as the
// built-in kernels' launch; false (nothing launched) when rapt_jit_prepare
// has not loaded it.  Errors surface through hipGetLastError like those of
// a built-in launch.
bool rapt_jit_launch(rapt_ctx* c, int kind, bool sp, const KArgs& A, int block, int64_t lanes) {
  hipFunction_t f = c->jit_fn[jit_slot(kind, sp)];
  if (!f) return false;
  const int64_t g = (lanes + block - 1) / block;
  void* args[] = {const_cast<KArgs*>(&A)};
  (void)hipModuleLaunchKernel(f, (unsigned)g, 1, 1, (unsigned)block, 1, 1, 0, c->stream, args,
                              nullptr);
  return true;
}

// The lane-group launch (rapt_launch_gcn's grids) of a compiled kernel; false
// when it is not loaded.
bool rapt_jit_launch_gcn(rapt_ctx* c, int lanes_per_particle, bool sp, const KArgs& A,
                         int block, int64_t groups, int64_t quads, int64_t pairs) {
  hipFunction_t f = c->jit_fn[jit_group_slot(lanes_per_particle, sp)];
  if (!f) return false;
  const GcnGrid g = gcn_grid(lanes_per_particle, block, groups, quads, pairs);
  if (lanes_per_particle == 3 || lanes_per_particle == 5) {
    int32_t ob = g.oct_blocks, wb = g.wide_blocks;
    void* args[] = {const_cast<KArgs*>(&A), &ob, &wb};
    (void)hipModuleLaunchKernel(f, g.blocks, 1, 1, (unsigned)block, 1, 1, 0, c->stream, args,
                                nullptr);
  } else {
    void* args[] = {const_cast<KArgs*>(&A)};
    (void)hipModuleLaunchKernel(f, g.blocks, 1, 1, (unsigned)block, 1, 1, 0, c->stream, args,
                                nullptr);
  }
  return true;
}

// A kernel of the context's program outside the tracer launches (the
// field-line kernels, field evaluation, right-hand sides, row diagnostics:
// slots RAPT_JIT_HALFBOUNCE onward): compiled and loaded when first needed,
// then launched with the built-in launch's grid and arguments.
int rapt_jit_run_slot(rapt_ctx* c, int slot, unsigned grid, unsigned block, void** args) {
  const int rc = prepare_slot(c, slot);
  if (rc) return rc;
  HIP_OK(c, hipModuleLaunchKernel(c->jit_fn[slot], grid, 1, 1, block, 1, 1, 0, c->stream, args,
                                  nullptr));
  return RAPT_OK;
}

extern "C" {

int rapt_field_expr_mode(rapt_ctx* c, int32_t mode) {
  if (!c) return RAPT_E_ARG;
  if (mode != RAPT_EXPR_INTERPRETED && mode != RAPT_EXPR_COMPILED)
    return set_err(c, RAPT_E_ARG, "expression mode must be RAPT_EXPR_INTERPRETED or "
                                  "RAPT_EXPR_COMPILED");
  c->expr_mode = mode;
  return RAPT_OK;
}

int rapt_field_expr_compile(int32_t nb, int32_t ne, const int32_t* code, const char* arch,
                            int32_t kernels, char* log, int64_t log_size) {
  put_log("", log, log_size);
  if (!code || nb < 1 || ne < 0 || nb + ne > RAPT_EXPR_MAXCODE) return RAPT_E_ARG;
  for (int32_t i = 0; i < nb + ne; ++i) {
    const int32_t op = code[4 * i], d = code[4 * i + 1], a = code[4 * i + 2],
                  b = code[4 * i + 3];
    if (op < 0 || op > 26 || d < 4 || d >= EXPR_NREG || a < 0 || b < 0 ||
        (op != 0 && a >= EXPR_NREG) || b >= EXPR_NREG) {
      put_log("bad instruction " + std::to_string(i), log, log_size);
      return RAPT_E_ARG;
    }
  }
  std::lock_guard<std::mutex> lock(g_mu);
  size_t bytes = 0;
  int count = 0;
  for (int slot = 0; slot < kSlots; ++slot) {
    if (kernels != 0 && !((kernels >> slot) & 1)) continue;
    ++count;
    std::string src;
    if (!make_source(nb, ne, code, slot, src)) return RAPT_E_ARG;
    const JitCode* jc = nullptr;
    std::string msg;
    const int rc = compile_locked(arch && *arch ? arch : "gfx950", slot, src, &jc, msg);
    if (rc) {
      put_log(msg, log, log_size);
      return rc;
    }
    bytes += jc->obj.size();
  }
  put_log(std::to_string(bytes) + " bytes of code in " + std::to_string(count) + " kernels",
          log, log_size);
  return RAPT_OK;
}

}  // extern "C"
