// rapt_mlp.hip -- the bounce-drift surrogate: an ensemble of MLPs
// 3 -> 256 -> 256 -> 256 -> 3 with SiLU, BF16 weights and FP32 accumulation,
// trained and evaluated by the kernels below on gfx950.  It has no
// counterpart in the reference: it replaces the traced bounce quantities of
// rapt_bounce_labels (BounceCenter's drift, rapt/BounceCenter.py:235-244,
// and flutils.bounceperiod, rapt/flutils.py:232-252) where the ensemble
// agrees, and the traced labels are the fallback where it does not.
//
// Model.  Query q = (L [Re], alpha_eq [deg], log10 K [eV]).  Features
// x = ((L - 5)/3, (alpha - 49.5)/39.5, (log10 K - 5.5)/1.5).  Each output is a
// log-residual against a dipole closed form b(q) (Walt's bounce and drift
// period approximations, rapt/utils.py:435-513, and the exact dipole mirror
// latitude): y = ln|label| - b(q), normalised by per-output shift/scale from
// the training set.  The network predicts the normalised residual.
//
// Parameters per member, one FP32 master array theta (offsets below) with an
// Adam state and a BF16 copy used by every forward pass.  Layer 1 (K = 3) and
// the output layer (N = 3) run on the vector ALU in FP32 with BF16 weights;
// the two 256 x 256 layers run on MFMA (v_mfma_f32_32x32x16_bf16) with BF16
// activations.  The third hidden activation stays FP32 into the output layer.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <new>

#include "../../include/rapt_hip.h"
#include "rapt_fields.hpp"
#include "rapt_internal.hpp"

namespace rapt {
namespace mlp {

constexpr int H = 256;
constexpr int OFF_W1 = 0;                 // [256][3]
constexpr int OFF_B1 = OFF_W1 + H * 3;    // [256]
constexpr int OFF_W2 = OFF_B1 + H;        // [256][256] (out, in)
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;
constexpr int OFF_B3 = OFF_W3 + H * H;
constexpr int OFF_W4 = OFF_B3 + H;        // [3][256]
constexpr int OFF_B4 = OFF_W4 + 3 * H;    // [3]
constexpr int NPARAM = OFF_B4 + 3;        // 133,379
constexpr int NPAD = (NPARAM + 255) / 256 * 256;
constexpr int PACKED = 8 * 16 * 64 * 8;   // bf16 elements of one packed 256x256 layer

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// The network is a SiLU MLP whose parameters are kept in base-2 units.  With
// the plain network's (W1, b1, W2, b2, W3, b3, W4, b4) the stored parameters
// are (log2(e) W1, log2(e) b1, W2, log2(e) b2, W3, log2(e) b3, ln(2) W4, b4):
// every hidden pre-activation is then a = log2(e) z and every hidden output
// g = log2(e) silu(z) = a / (1 + 2^-a), the same function, while the logistic
// factor is rcp(1 + exp2(-a)), one v_exp_f32 and one v_rcp_f32 (about 1 ulp
// each) with no scaling multiply in front.  Training and inference use these
// same definitions, so both round the same hidden values to bf16.
constexpr float LN2 = 0.6931471805599453f;
__device__ __forceinline__ float sig2(float a) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-a));
}
__device__ __forceinline__ float act2(float a) { return a * sig2(a); }
// d act2 / da = s (1 + ln2 a (1 - s)), s = sig2(a)
__device__ __forceinline__ float dact2(float a) {
  const float s = sig2(a);
  return s * (1.0f + LN2 * a * (1.0f - s));
}

// ---- dipole baselines (the residual targets' reference point) -----------------------
struct Base {
  float lnT, lnW, lnLat;
};
// speedfromKE (rapt/utils.py:218-249) then Walt's approximations
// (rapt/utils.py:435-513): T_b = 0.117 L / beta (1 - 0.4635 sin^0.75 a),
// T_d = C / (L gamma beta^2) (1 - sin^0.62 a / 3) with C = 8.481 (protons,
// scaled by m / m_p for other masses); the dipole mirror latitude solves
// (1 - x)^3 = sin^2 a sqrt(1 + 3 x), x = sin^2 lambda, by Newton's method.
__device__ __forceinline__ Base baseline(float L, float adeg, float lgk, double mass,
                                         double eV) {
  const double c = 299792458.0;
  const double K = pow(10.0, (double)lgk) * eV;
  const double mc2 = mass * c * c;
  double beta, gam;
  if (K / mc2 < 1e-6) {
    beta = sqrt(2.0 * K / mass) / c;
    gam = 1.0 / sqrt(1.0 - beta * beta);
  } else {
    gam = 1.0 + K / mc2;
    beta = sqrt(1.0 - 1.0 / (gam * gam));
  }
  const double a = (double)adeg * (3.141592653589793 / 180.0);
  const double sa = sin(a);
  const double Tb = 0.117 * L / beta * (1.0 - 0.4635 * pow(sa, 0.75));
  const double Cd = 8.481 * (mass / 1.672621777e-27);
  const double Td = Cd / (L * gam * beta * beta) * (1.0 - pow(sa, 0.62) / 3.0);
  const double s2 = sa * sa;
  double x = 0.5 * (1.0 - s2);  // sin^2 lambda
  for (int it = 0; it < 40; ++it) {
    const double u = 1.0 - x, r = sqrt(1.0 + 3.0 * x);
    const double g = u * u * u - s2 * r;
    const double dg = -3.0 * u * u - 1.5 * s2 / r;
    double xn = x - g / dg;
    xn = fmin(fmax(xn, 0.0), 1.0);
    if (fabs(xn - x) < 1e-15) { x = xn; break; }
    x = xn;
  }
  const double lat = asin(sqrt(fmax(x, 1e-300))) * (180.0 / 3.141592653589793);
  return Base{(float)log(Tb), (float)log(6.283185307179586 / Td), (float)log(lat)};
}

// ---- the two regimes of the drift label ------------------------------------------------
// The reference's gradI (rapt/flutils.py:196-227) differences I on lines
// eyegradientstep d either side of the query point; on the query's x axis the
// inner line starts at (L Re - d, 0, 0).  When the equatorial |B| of that line
// reaches the mirror field (I = 0 there, flutils.eye's early return,
// rapt/flutils.py:97-100), the label takes a one-sided difference instead of a
// central one: a jump of O(d / L) across the curve B(L Re - d) = B_m, which an
// MLP smooths over.  With the split on, the query's side of that curve moves
// the pitch-angle feature by REGIME_SHIFT, so the two smooth pieces sit apart in
// input space; queries within 1e-5 (relative) of the curve, where the float
// query may fall on the other side than the traced label's double one, go to
// the fallback.
struct SplitC {
  int32_t on;
  double Re, d;
  FieldC F;
};
constexpr float REGIME_SHIFT = 2.5f;

template <int M>
__device__ __forceinline__ double eq_magB(const FieldC& F, double x) {
  return magB<M>(F, 0.0, x, 0.0, 0.0);
}

// 1 for the one-sided regime; *near when the query is within 1e-5 of the curve
__device__ int regime(const SplitC& S, float L, float adeg, bool* near) {
  *near = false;
  if (!S.on) return 0;
  const double x = (double)L * S.Re;
  double Bq, Bin;
  switch (S.F.model) {
    case 1: Bq = eq_magB<1>(S.F, x); Bin = eq_magB<1>(S.F, x - S.d); break;
    case 2: Bq = eq_magB<2>(S.F, x); Bin = eq_magB<2>(S.F, x - S.d); break;
    case 5: Bq = eq_magB<5>(S.F, x); Bin = eq_magB<5>(S.F, x - S.d); break;
    case 8: Bq = eq_magB<8>(S.F, x); Bin = eq_magB<8>(S.F, x - S.d); break;
    default: return 0;
  }
  const double sa = sin((double)adeg * (3.141592653589793 / 180.0));
  const double Bm = Bq / (sa * sa);
  *near = fabs(Bin - Bm) < 1e-5 * Bm;
  return (Bin > Bm || fabs(Bin - Bm) / Bm < 1e-12) ? 1 : 0;
}

__device__ __forceinline__ void features(const float* q, float x[3], int side) {
  x[0] = (q[0] - 5.0f) * (1.0f / 3.0f);
  x[1] = (q[1] - 49.5f) * (1.0f / 39.5f) + (side ? REGIME_SHIFT : 0.0f);
  x[2] = (q[2] - 5.5f) * (1.0f / 1.5f);
}

// ---- training kernels -----------------------------------------------------------------
// Every kernel below runs all members at once: blockIdx.z (or y) = member,
// with per-member strides.

// C[M][N] = op(A)[M][K] * op(B)[K][N], BF16 in, FP32 out.  op(A)[m][k] =
// A[m*lda+k] (TA false) or A[k*lda+m]; op(B)[k][n] = B[k*ldb+n] (TB false)
// or B[n*ldb+k].  M, N multiples of 64, K a multiple of 16.  Block: 4 waves,
// 64 x 64 tile, wave (w/2, w%2) owns a 32 x 32 quarter.
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const bf16* A, int lda,
                                              int64_t sA, const bf16* B, int ldb, int64_t sB,
                                              float* C, int ldc, int64_t sC) {
  __shared__ bf16 As[64][16 + 8];
  __shared__ bf16 Bs[16][64 + 8];
  const int mem = blockIdx.z;
  A += mem * sA;
  B += mem * sB;
  C += mem * sC;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  const int wm = w >> 1, wn = w & 1;
  f32x16 acc = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    if (!TA) {
      const int i = t >> 2, kk = (t & 3) * 4;
      for (int e = 0; e < 4; ++e) As[i][kk + e] = A[(int64_t)(m0 + i) * lda + k0 + kk + e];
    } else {
      const int kk = t >> 4, i = (t & 15) * 4;
      for (int e = 0; e < 4; ++e) As[i + e][kk] = A[(int64_t)(k0 + kk) * lda + m0 + i + e];
    }
    if (!TB) {
      const int kk = t >> 4, j = (t & 15) * 4;
      for (int e = 0; e < 4; ++e) Bs[kk][j + e] = B[(int64_t)(k0 + kk) * ldb + n0 + j + e];
    } else {
      const int j = t >> 2, kk = (t & 3) * 4;
      for (int e = 0; e < 4; ++e) Bs[kk + e][j] = B[(int64_t)(n0 + j) * ldb + k0 + kk + e];
    }
    __syncthreads();
    bf16x8 a, b;
    for (int j = 0; j < 8; ++j) {
      a[j] = As[wm * 32 + (l & 31)][8 * (l >> 5) + j];
      b[j] = Bs[8 * (l >> 5) + j][wn * 32 + (l & 31)];
    }
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    __syncthreads();
  }
  for (int r = 0; r < 16; ++r) {
    const int row = wm * 32 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
    const int col = wn * 32 + (l & 31);
    C[(int64_t)(m0 + row) * ldc + n0 + col] = acc[r];
  }
}

// layer 1: Z1 = W1 x + b1, H1 = act2(Z1) (bf16).  grid (Bt*256/256, 1, members)
__global__ void k_in_fwd(int Bt, const float* X, const bf16* wb, float* Z1, bf16* H1,
                         int64_t sZ) {
  const int mem = blockIdx.z;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)Bt * H) return;
  const int s = (int)(idx / H), n = (int)(idx % H);
  const bf16* w = wb + (int64_t)mem * NPAD;
  float z = (float)w[OFF_B1 + n];
  for (int i = 0; i < 3; ++i) z += (float)w[OFF_W1 + n * 3 + i] * X[s * 3 + i];
  Z1[mem * sZ + idx] = z;
  H1[mem * sZ + idx] = (bf16)act2(z);
}

// Z += b (layer bias at boff), H = act2(Z), as bf16 (Hb) or fp32 (Hf)
__global__ void k_bias_act(int Bt, const bf16* wb, int boff, float* Z, bf16* Hb, float* Hf,
                           int64_t sZ) {
  const int mem = blockIdx.z;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)Bt * H) return;
  const int n = (int)(idx % H);
  const float z = Z[mem * sZ + idx] + (float)wb[(int64_t)mem * NPAD + boff + n];
  Z[mem * sZ + idx] = z;
  if (Hb) Hb[mem * sZ + idx] = (bf16)act2(z);
  if (Hf) Hf[mem * sZ + idx] = act2(z);
}

// output layer and loss: Y = W4 H3 + b4; dY = 2 (Y - T) / (3 Bt); loss sum
__global__ void k_out_fwd(int Bt, const float* H3, const bf16* wb, const float* T, float* Y,
                          float* dY, double* loss, int64_t sZ) {
  const int mem = blockIdx.z;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double part = 0.0;
  if (s < Bt) {
    const bf16* w = wb + (int64_t)mem * NPAD;
    const float* h = H3 + mem * sZ + (int64_t)s * H;
    for (int k = 0; k < 3; ++k) {
      float y = (float)w[OFF_B4 + k];
      for (int n = 0; n < H; ++n) y += (float)w[OFF_W4 + k * H + n] * h[n];
      const int64_t o = (int64_t)mem * Bt * 3 + s * 3 + k;
      Y[o] = y;
      if (T) {
        const float d = y - T[s * 3 + k];
        dY[o] = 2.0f * d / (3.0f * Bt);
        part += (double)d * d;
      }
    }
  }
  if (T && loss) {
    for (int off = 32; off > 0; off >>= 1) part += __shfl_down(part, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(&loss[mem], part);
  }
}

// dH3 = dY W4, dZ3 = dH3 act2'(Z3) -> bf16
__global__ void k_out_bwd(int Bt, const float* dY, const bf16* wb, const float* Z3, bf16* dZ3,
                          int64_t sZ) {
  const int mem = blockIdx.z;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)Bt * H) return;
  const int s = (int)(idx / H), n = (int)(idx % H);
  const bf16* w = wb + (int64_t)mem * NPAD;
  const float* dy = dY + (int64_t)mem * Bt * 3 + s * 3;
  const float g = dy[0] * (float)w[OFF_W4 + n] + dy[1] * (float)w[OFF_W4 + H + n] +
                  dy[2] * (float)w[OFF_W4 + 2 * H + n];
  dZ3[mem * sZ + idx] = (bf16)(g * dact2(Z3[mem * sZ + idx]));
}

// dZ = dH act2'(Z) -> bf16
__global__ void k_act_bwd(int Bt, const float* dH, const float* Z, bf16* dZ, int64_t sZ) {
  const int mem = blockIdx.z;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (int64_t)Bt * H) return;
  dZ[mem * sZ + idx] = (bf16)(dH[mem * sZ + idx] * dact2(Z[mem * sZ + idx]));
}

// Batch reductions in two deterministic stages: per chunk of RCH samples
// (grid x = chunk), then a sum over chunks in chunk order.
constexpr int RCH = 64;
constexpr int NPART = 8;  // partial slots per column

// bias gradient partials: part[mem][c][n][slot] = sum over chunk of dZ[s][n]
__global__ void k_colsum_part(int Bt, const bf16* dZ, float* part, int slot, int64_t sZ) {
  const int mem = blockIdx.z, c = blockIdx.x, n = threadIdx.x;
  const int chunks = Bt / RCH;
  const bf16* p = dZ + mem * sZ + (int64_t)c * RCH * H;
  float acc = 0.0f;
  for (int s = 0; s < RCH; ++s) acc += (float)p[(int64_t)s * H + n];
  part[(((int64_t)mem * chunks + c) * H + n) * NPART + slot] = acc;
}

// partials of dW4[k][n] = sum_s dY[s][k] H3[s][n] (slots 0..2), dW1[n][i] =
// sum_s dZ1[s][n] X[s][i] (slots 3..5, dZ1 = dH1 act2'(Z1) in fp32), db1
// (slot 6) and db4 (slot 7, columns 0..2)
__global__ void k_small_part(int Bt, const float* dY, const float* H3, const float* dH1,
                             const float* Z1, const float* X, float* part, int64_t sZ) {
  const int mem = blockIdx.z, c = blockIdx.x, n = threadIdx.x;
  const int chunks = Bt / RCH;
  const float* h3 = H3 + mem * sZ;
  const float* dh1 = dH1 + mem * sZ;
  const float* z1 = Z1 + mem * sZ;
  const float* dy = dY + (int64_t)mem * Bt * 3;
  float w4[3] = {0, 0, 0}, w1[3] = {0, 0, 0}, b1 = 0.0f, b4 = 0.0f;
  for (int s = c * RCH; s < (c + 1) * RCH; ++s) {
    const float hv = h3[(int64_t)s * H + n];
    for (int k = 0; k < 3; ++k) w4[k] += dy[s * 3 + k] * hv;
    const float dz = dh1[(int64_t)s * H + n] * dact2(z1[(int64_t)s * H + n]);
    b1 += dz;
    for (int i = 0; i < 3; ++i) w1[i] += dz * X[s * 3 + i];
    if (n < 3) b4 += dy[s * 3 + n];
  }
  float* q = part + (((int64_t)mem * chunks + c) * H + n) * NPART;
  q[0] = w4[0]; q[1] = w4[1]; q[2] = w4[2];
  q[3] = w1[0]; q[4] = w1[1]; q[5] = w1[2];
  q[6] = b1;
  q[7] = b4;
}

// second stages: sum the partials over chunks and scatter the gradients
// (k_small_part's slots; k_colsum_part's slot 0 = b3, slot 1 = b2)
__global__ void k_sum_small(int chunks, const float* part, float* grad) {
  const int mem = blockIdx.z, n = threadIdx.x;
  float v[NPART] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int c = 0; c < chunks; ++c) {
    const float* q = part + (((int64_t)mem * chunks + c) * H + n) * NPART;
    for (int k = 0; k < NPART; ++k) v[k] += q[k];
  }
  float* g = grad + (int64_t)mem * NPAD;
  for (int k = 0; k < 3; ++k) g[OFF_W4 + k * H + n] = v[k];
  for (int i = 0; i < 3; ++i) g[OFF_W1 + n * 3 + i] = v[3 + i];
  g[OFF_B1 + n] = v[6];
  if (n < 3) g[OFF_B4 + n] = v[7];
}

__global__ void k_sum_bias(int chunks, const float* part, float* grad) {
  const int mem = blockIdx.z, n = threadIdx.x;
  float b3 = 0.0f, b2 = 0.0f;
  for (int c = 0; c < chunks; ++c) {
    const float* q = part + (((int64_t)mem * chunks + c) * H + n) * NPART;
    b3 += q[0];
    b2 += q[1];
  }
  float* g = grad + (int64_t)mem * NPAD;
  g[OFF_B3 + n] = b3;
  g[OFF_B2 + n] = b2;
}

// Adam (Kingma & Ba 2015) on the FP32 masters; refreshes the BF16 copy
__global__ void k_adam(float* theta, const float* grad, float* m1, float* m2, bf16* wb,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
  const int mem = blockIdx.z;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= NPARAM) return;
  const int64_t o = (int64_t)mem * NPAD + i;
  const float g = grad[o];
  const float a = b1 * m1[o] + (1.0f - b1) * g;
  const float v = b2 * m2[o] + (1.0f - b2) * g * g;
  m1[o] = a;
  m2[o] = v;
  const float p = theta[o] - lr * (a / c1) / (sqrtf(v / c2) + eps);
  theta[o] = p;
  wb[o] = (bf16)p;
}

__global__ void k_gather(int bt, const int64_t* ord, const float* Xa, const float* Ta, float* Xb,
                         float* Tb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= bt) return;
  const int64_t j = ord[i];
  for (int k = 0; k < 3; ++k) {
    Xb[3 * i + k] = Xa[3 * j + k];
    Tb[3 * i + k] = Ta[3 * j + k];
  }
}

__global__ void k_zero_d(double* p, int k) {
  if ((int)threadIdx.x < k) p[threadIdx.x] = 0.0;
}

__global__ void k_zero_f(float* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0f;
}

__global__ void k_to_bf16(const float* theta, bf16* wb, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) wb[i] = (bf16)theta[i];
}

// ---- fused inference -------------------------------------------------------------------
// One kernel runs the whole ensemble forward pass, the disagreement test and
// the baseline for a batch of queries; activations never leave registers.
//
// v_mfma_f32_32x32x16_bf16, 32 queries per wave: lane l holds query l%32 and
// half h = l/32 of each activation.  A 32x32 output tile t leaves register r
// of lane l holding neuron 32t + row32(r, h), row32(r, h) = 8(r/4) + 4h + r%4
// (the D layout of the instruction: D[8(r/4)+4(l/32)+r%4][l%32]; A operand:
// lane l holds row l%32, K indices 8(l/32)+j; B operand: column l%32, K
// indices 8(l/32)+j).  Registers 8g..8g+7 of tile t form the B operand of
// K-block kb = 2t+g of the next layer, so K index 8h + j of block kb is input
// neuron pin32(kb, h, j) = 32(kb/2) + 16(kb%2) + 8(j/4) + 4h + j%4, and the
// packed weights are P[t][kb][lane][j] = W[32t + lane%32][pin32(kb, lane/32,
// j)].  A block of IW waves stages one member's packed 256x256 layer (128
// KB) in LDS and its waves (IW x 32 queries) read the A fragments from
// there, so L2 supplies each weight once per block instead of once per wave;
// each fragment read from LDS feeds a 32-query MFMA.
constexpr int MAXM = 4;   // largest ensemble

struct InferArgs {
  int64_t n;
  int members;
  const bf16* wb;         // [members][NPAD]
  const bf16* packed;     // [members][2][PACKED]: layers 2 and 3
  const bf16* p1;         // [members][16*64*8]: layer-1 A operands
  const float* norm;      // shift[3], scale[3]
  float thresh;           // relative member deviation that sends a query to the fallback
  const float* prep;      // [n][8] k_prep's features and baselines
  float* out;             // [n][3] physical outputs (T_b, |omega_d|, latitude)
  float* raw;             // optional [members][n][3] normalised outputs
  int32_t* flag;          // [n] in: k_prep's side | near << 1; out: 1 = fallback
};

// one member's small parameters, staged in LDS: b1, b2, b3, w4 (one float4
// of the three output weights per layer-3 neuron), b4
struct SmallW {
  float b1[H], b2[H], b3[H];
  float4 w4[H];
  float b4[4];
};
typedef float f32x4 __attribute__((ext_vector_type(4)));

// the biases of one 32x32 output tile in a lane's accumulator order: register
// r holds neuron row32(r, h), so b points at the tile's bias + 4h and the 16
// values are four float4 runs at 8-element strides
__device__ __forceinline__ f32x16 bias16(const float* b) {
  f32x16 acc;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 v = *(const float4*)(b + 8 * q);
    acc[4 * q] = v.x;
    acc[4 * q + 1] = v.y;
    acc[4 * q + 2] = v.z;
    acc[4 * q + 3] = v.w;
  }
  return acc;
}

__device__ __forceinline__ int pin32(int kb, int h, int j) {
  return 32 * (kb >> 1) + 16 * (kb & 1) + 8 * (j >> 2) + 4 * h + (j & 3);
}
__device__ __forceinline__ int row32(int r, int h) { return 8 * (r >> 2) + 4 * h + (r & 3); }

__global__ void k_pack(const bf16* wb, bf16* packed, int woff) {
  const int mem = blockIdx.z;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= PACKED) return;
  const int j = idx & 7, lane = (idx >> 3) & 63, kb = (idx >> 9) & 15, tp = idx >> 13;
  const int row = 32 * tp + (lane & 31), col = pin32(kb, lane >> 5, j);
  packed[(int64_t)mem * 2 * PACKED + idx] = wb[(int64_t)mem * NPAD + woff + row * H + col];
}

// Layer 1 as an MFMA: the three features are split into bf16 hi and lo parts
// (x = hi + lo to 2^-17 relative), so K index j < 3 multiplies W1[.][j] by
// hi_j and 3 <= j < 6 multiplies W1[.][j-3] by lo_{j-3}: the product with the
// bf16 weights is as accurate as the FP32 vector form.  A operand of tile t
// (32 neurons), lane (h, j): W1[32t + l%32][j%3] for h = 0, j < 6, else 0.
constexpr int P1SIZE = 8 * 64 * 8;  // bf16 elements of one member's layer-1 operands
__global__ void k_pack1(const bf16* wb, bf16* p1) {
  const int mem = blockIdx.z;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // 8 tiles * 64 lanes * 8
  if (idx >= P1SIZE) return;
  const int j = idx & 7, lane = (idx >> 3) & 63, t = idx >> 9;
  const int n = 32 * t + (lane & 31);
  bf16 v = (bf16)0.0f;
  if ((lane >> 5) == 0 && j < 6) v = wb[(int64_t)mem * NPAD + OFF_W1 + n * 3 + (j % 3)];
  p1[(int64_t)mem * P1SIZE + idx] = v;
}

// Per-query pre-pass of the inference kernel: the regime flag (side, bit 0;
// near, bit 1; 0 when the split is off), the three input features and the
// three dipole baselines, prep[8i .. 8i+5] = (x0, x1, x2, lnT, lnW, lnLat).
// The double-precision work (field evaluations, Walt's forms, the
// mirror-latitude Newton loop) runs here on every lane of a wave, instead of
// on the 16 query lanes of each inference wave inside its register budget.
__global__ void k_prep(int64_t n, const float* query, SplitC S, double mass, double eV,
                       float* prep, int32_t* flag) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float q[3] = {query[3 * i], query[3 * i + 1], query[3 * i + 2]};
  bool near;
  const int side = regime(S, q[0], q[1], &near);
  flag[i] = side | (near ? 2 : 0);
  float x[3];
  features(q, x, side);
  const Base b = baseline(q[0], q[1], q[2], mass, eV);
  float4* o = (float4*)(prep + 8 * i);
  o[0] = make_float4(x[0], x[1], x[2], b.lnT);
  o[1] = make_float4(b.lnW, b.lnLat, 0.0f, 0.0f);
}

// IW = 8 waves per block (two per SIMD), 32 queries per wave.  The biases
// start the accumulators.  Staging is pipelined: each member's two 256x256
// layers pass through LDS as four 64 KB halves (layer 2 tiles 0-3, 4-7, then
// layer 3), alternating between two buffers; while the waves compute with
// one half, the next half (the next member's, or member 0's for the next
// chunk) streams into the other buffer with direct global-to-LDS loads, so
// the loads never wait on a barrier with the matrix cores idle.  A half is
// 64 fragments of 1 KB; wave w loads fragments 8i + w.
constexpr int IW = 8;
constexpr int QW = 32;  // queries per wave
constexpr int HALF = PACKED / 16;  // bf16x8 elements of half a packed layer (64 KB)

// stage s of a member's four: layer 2 + (s >> 1), fragments 64 (s & 1) onward.
// Wave wv loads fragments 8i + wv, starting at i = rot (the block's rotation,
// so that the CUs of an XCD do not all ask the L2 for the same lines at the
// same moment).  The loads go through a buffer resource
// (buffer_load_dwordx4 ... lds), which counts on vmcnt alone: with the FLAT
// form (global_load_lds) in flight the compiler must treat the LDS counter as
// unordered, and every LDS read of the half being computed would then wait
// for all reads issued before it.
__device__ __forceinline__ void stage_load(__amdgpu_buffer_rsrc_t rs, int m, int s, bf16x8* buf,
                                           int wv, int l, int rot) {
  const int base = (m * 2 + (s >> 1)) * PACKED * 2 + (s & 1) * HALF * 16;  // bytes
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int f = 8 * ((i + rot) & 7) + wv;
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        rs, (__attribute__((address_space(3))) void*)(buf + 64 * f), 16,
        base + (64 * f + l) * 16, 0, 0, 0);
  }
}

__device__ __forceinline__ void wait_loads() { __builtin_amdgcn_s_waitcnt(0); }

// one 32x32 output tile of a hidden layer from the fragments of buf (tile
// t of the half: fragments 16 t .. 16 t + 15).  The A fragments are read
// PF ahead into a ring of registers, so each MFMA's LDS read has PF - 1
// matrix operations of cover (the compiler's own schedule gives one).
constexpr int PF = 4;
__device__ __forceinline__ f32x16 tile32(const bf16x8* buf, int t, const bf16x8* in,
                                         f32x16 acc, int l) {
  const bf16x8* p = buf + 16 * t * 64 + l;
  bf16x8 a[PF];
#pragma unroll
  for (int i = 0; i < PF; ++i) a[i] = p[i * 64];
#pragma unroll
  for (int kb = 0; kb < 16; ++kb) {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[kb % PF], in[kb], acc, 0, 0, 0);
    if (kb + PF < 16) a[kb % PF] = p[(kb + PF) * 64];
  }
  // keep that order in the schedule (left alone, the scheduler sinks each
  // read next to its MFMA to save registers): PF LDS reads (mask 0x100),
  // then per step one MFMA (0x008), one LDS read and up to five other
  // vector instructions (0x002)
  __builtin_amdgcn_sched_group_barrier(0x100, PF, 0);
#pragma unroll
  for (int kb = 0; kb < 16; ++kb) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
    if (kb + PF < 16) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
    __builtin_amdgcn_sched_group_barrier(0x002, 5, 0);
  }
  return acc;
}

__global__ __launch_bounds__(IW * 64) void k_infer(InferArgs A) {
  __shared__ bf16x8 buf0[HALF], buf1[HALF];
  __shared__ SmallW sw[MAXM];
  const int l = threadIdx.x & 63, wv = threadIdx.x >> 6, h = l >> 5;
  const int64_t groups = (A.n + QW - 1) / QW;
  const int64_t chunks = (groups + IW - 1) / IW;
  if ((int64_t)blockIdx.x >= chunks) return;  // no work: no barrier is reached
  for (int m = 0; m < A.members; ++m) {
    const bf16* w = A.wb + (int64_t)m * NPAD;
    for (int i = threadIdx.x; i < H; i += IW * 64) {
      sw[m].b1[i] = (float)w[OFF_B1 + i];
      sw[m].b2[i] = (float)w[OFF_B2 + i];
      sw[m].b3[i] = (float)w[OFF_B3 + i];
      sw[m].w4[i] = make_float4((float)w[OFF_W4 + i], (float)w[OFF_W4 + H + i],
                                (float)w[OFF_W4 + 2 * H + i], 0.0f);
    }
    if (threadIdx.x < 3) sw[m].b4[threadIdx.x] = (float)w[OFF_B4 + threadIdx.x];
  }
  // the packed layers 2 and 3 of every member as a raw buffer (stride 0,
  // the gfx9 word-3 format bits); reads past the end return 0
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void*)A.packed, (short)0, A.members * 2 * PACKED * (int)sizeof(bf16), 0x00020000);
  const int rot = blockIdx.x & 7;
  stage_load(rs, 0, 0, buf0, wv, l, rot);
  for (int64_t ch = blockIdx.x; ch < chunks; ch += gridDim.x) {
    const int64_t g = ch * IW + wv;
    const int64_t qi = g * QW + (l & 31);
    const bool valid = g < groups && qi < A.n;
    const float4* pq = (const float4*)(A.prep + 8 * (valid ? qi : 0));
    const float4 p0 = pq[0];
    const float xin[3] = {p0.x, p0.y, p0.z};
    const int rg = valid ? A.flag[qi] : 0;  // k_prep's side | near << 1
    bf16x8 xb;
#pragma unroll
    for (int j = 0; j < 8; ++j) xb[j] = (bf16)0.0f;
    if (h == 0) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const bf16 hv = (bf16)xin[j];
        xb[j] = hv;
        xb[3 + j] = (bf16)(xin[j] - (float)hv);
      }
    }
    float mean[3] = {0.0f, 0.0f, 0.0f}, lo[3], hi[3];
    for (int c = 0; c < 3; ++c) { lo[c] = 3.0e38f; hi[c] = -3.0e38f; }
    for (int m = 0; m < A.members; ++m) {
      const SmallW& W = sw[m];
      const bf16x8* P1 = (const bf16x8*)(A.p1 + (int64_t)m * P1SIZE);
      const int mn = m + 1 < A.members ? m + 1 : 0;  // the member whose layer 2 comes next
      // layer 1 (its operands come from global memory)
      bf16x8 act[16];
      wait_loads();
      __syncthreads();  // first pass: sw written; every pass: this member's half 0 landed
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        f32x16 acc = bias16(&W.b1[32 * t + 4 * h]);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(P1[t * 64 + l], xb, acc, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 16; ++r) act[2 * t + (r >> 3)][r & 7] = (bf16)act2(acc[r]);
      }
      // layer 2, tiles 0-3 from buf0 while tiles 4-7 stream into buf1
      stage_load(rs, m, 1, buf1, wv, l, rot);
      bf16x8 nxt[16];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        f32x16 acc = tile32(buf0, t, act, bias16(&W.b2[32 * t + 4 * h]), l);
#pragma unroll
        for (int r = 0; r < 16; ++r) nxt[2 * t + (r >> 3)][r & 7] = (bf16)act2(acc[r]);
      }
      wait_loads();
      __syncthreads();  // buf1 landed; buf0 free
      stage_load(rs, m, 2, buf0, wv, l, rot);
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        f32x16 acc = tile32(buf1, t, act, bias16(&W.b2[32 * (t + 4) + 4 * h]), l);
#pragma unroll
        for (int r = 0; r < 16; ++r) nxt[2 * (t + 4) + (r >> 3)][r & 7] = (bf16)act2(acc[r]);
      }
      wait_loads();
      __syncthreads();  // layer 3 half 0 in buf0; buf1 free
      stage_load(rs, m, 3, buf1, wv, l, rot);
      // layer 3 and the output layer (FP32)
      float o0 = 0.0f, o1 = 0.0f, o2 = 0.0f;
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        if (half == 1) {
          wait_loads();
          __syncthreads();  // layer 3 half 1 in buf1; buf0 free
          stage_load(rs, mn, 0, buf0, wv, l, rot);
        }
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int tt = 4 * half + t;
          f32x16 acc = tile32(half ? buf1 : buf0, t, nxt, bias16(&W.b3[32 * tt + 4 * h]), l);
#pragma unroll
          for (int r = 0; r < 16; ++r) {
            const float4 c = W.w4[32 * tt + row32(r, h)];
            const float v = act2(acc[r]);
            // fused multiply-adds (the library compiles without contraction)
            o0 = __builtin_fmaf(c.x, v, o0);
            o1 = __builtin_fmaf(c.y, v, o1);
            o2 = __builtin_fmaf(c.z, v, o2);
          }
        }
      }
      o0 += __shfl_xor(o0, 32);
      o1 += __shfl_xor(o1, 32);
      o2 += __shfl_xor(o2, 32);
      const float o[3] = {o0 + W.b4[0], o1 + W.b4[1], o2 + W.b4[2]};
      for (int c = 0; c < 3; ++c) {
        mean[c] += o[c];
        lo[c] = fminf(lo[c], o[c]);
        hi[c] = fmaxf(hi[c], o[c]);
      }
      if (A.raw && valid && l < 32)
        for (int c = 0; c < 3; ++c) A.raw[((int64_t)m * A.n + qi) * 3 + c] = o[c];
    }
    if (valid && l < 32) {
      const float4 p1 = pq[1];
      const float bb[3] = {p0.w, p1.x, p1.y};
      int bad = (rg & 2) ? 1 : 0;
      for (int k = 0; k < 3; ++k) {
        const float sc = A.norm[3 + k];
        const float mu = mean[k] / (float)A.members;
        if (expm1f(sc * fmaxf(hi[k] - mu, mu - lo[k])) > A.thresh) bad = 1;
        A.out[qi * 3 + k] = expf(bb[k] + A.norm[k] + sc * mu);
      }
      A.flag[qi] = bad;
    }
  }
  wait_loads();  // the prefetch issued for a chunk this block does not run
}

// targets: normalised log-residuals from traced labels (T_b, omega_d, lat)
__global__ void k_targets(int64_t n, const float* query, const double* labels,
                          const float* norm, double mass, double eV, SplitC S, float* X,
                          float* T) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float x[3];
  bool near;
  features(query + 3 * i, x, regime(S, query[3 * i], query[3 * i + 1], &near));
  const Base b = baseline(query[3 * i], query[3 * i + 1], query[3 * i + 2], mass, eV);
  const float bb[3] = {b.lnT, b.lnW, b.lnLat};
  for (int k = 0; k < 3; ++k) {
    X[3 * i + k] = x[k];
    const double lab = fabs(labels[3 * i + k]);
    T[3 * i + k] = norm ? ((float)(log(lab) - (double)bb[k]) - norm[k]) / norm[3 + k]
                        : (float)(log(lab) - (double)bb[k]);
  }
}

}  // namespace mlp
}  // namespace rapt

using namespace rapt;
using namespace rapt::mlp;

struct rapt_mlp {
  rapt_ctx* ctx = nullptr;
  int members = 0;
  int batch = 0;
  double mass = 0.0, eV = 0.0;
  int64_t step = 0;
  float norm_h[6] = {0, 0, 0, 1, 1, 1};
  float* theta = nullptr;
  float* grad = nullptr;
  float* m1 = nullptr;
  float* m2 = nullptr;
  bf16* wb = nullptr;
  bf16* packed = nullptr;
  bf16* p1 = nullptr;
  float* norm = nullptr;
  bool packed_ok = false;
  SplitC split{};  // regime of the drift label (rapt_mlp_set_split)
  // training workspace (batch-sized)
  float *X = nullptr, *T = nullptr, *Z1 = nullptr, *Z2 = nullptr, *Z3 = nullptr;
  float *H3 = nullptr, *dH = nullptr, *Y = nullptr, *dY = nullptr;
  bf16 *H1 = nullptr, *H2 = nullptr, *dZb = nullptr;
  float *partS = nullptr, *partB = nullptr;
  double* loss = nullptr;
};

static void mlp_free_work(rapt_mlp* m) {
  void* p[] = {m->X, m->T, m->Z1, m->Z2, m->Z3, m->H3, m->dH, m->Y, m->dY,
               m->H1, m->H2, m->dZb, m->partS, m->partB, m->loss};
  for (void* q : p) (void)hipFree(q);
  m->X = m->T = m->Z1 = m->Z2 = m->Z3 = m->H3 = m->dH = m->Y = m->dY = nullptr;
  m->partS = m->partB = nullptr;
  m->H1 = m->H2 = m->dZb = nullptr;
  m->loss = nullptr;
  m->batch = 0;
}

static int mlp_work(rapt_mlp* m, int bt) {
  if (m->batch == bt) return RAPT_OK;
  mlp_free_work(m);
  rapt_ctx* c = m->ctx;
  const size_t act = (size_t)m->members * bt * H;
  bool ok = hipMalloc(&m->X, sizeof(float) * 3 * bt) == hipSuccess &&
            hipMalloc(&m->T, sizeof(float) * 3 * bt) == hipSuccess &&
            hipMalloc(&m->Z1, sizeof(float) * act) == hipSuccess &&
            hipMalloc(&m->Z2, sizeof(float) * act) == hipSuccess &&
            hipMalloc(&m->Z3, sizeof(float) * act) == hipSuccess &&
            hipMalloc(&m->H3, sizeof(float) * act) == hipSuccess &&
            hipMalloc(&m->dH, sizeof(float) * act) == hipSuccess &&
            hipMalloc(&m->Y, sizeof(float) * 3 * bt * m->members) == hipSuccess &&
            hipMalloc(&m->dY, sizeof(float) * 3 * bt * m->members) == hipSuccess &&
            hipMalloc(&m->H1, sizeof(bf16) * act) == hipSuccess &&
            hipMalloc(&m->H2, sizeof(bf16) * act) == hipSuccess &&
            hipMalloc(&m->dZb, sizeof(bf16) * act) == hipSuccess &&
            hipMalloc(&m->partS, sizeof(float) * (act / RCH) * NPART) == hipSuccess &&
            hipMalloc(&m->partB, sizeof(float) * (act / RCH) * NPART) == hipSuccess &&
            hipMalloc(&m->loss, sizeof(double) * m->members) == hipSuccess;
  if (!ok) {
    mlp_free_work(m);
    return set_err(c, RAPT_E_OOM, "MLP workspace allocation failed");
  }
  m->batch = bt;
  return RAPT_OK;
}

// forward of a training batch already in m->X; T may be null (no loss)
static void mlp_forward(rapt_mlp* m, int bt, bool with_loss) {
  hipStream_t st = m->ctx->stream;
  const int mem = m->members;
  const int64_t sZ = (int64_t)bt * H;
  const dim3 ew((unsigned)((sZ + 255) / 256), 1, mem);
  const dim3 gg(H / 64, bt / 64, mem);
  hipLaunchKernelGGL(k_in_fwd, ew, dim3(256), 0, st, bt, m->X, m->wb, m->Z1, m->H1, sZ);
  // Z2 = H1 W2^T
  hipLaunchKernelGGL((k_gemm<false, true>), gg, dim3(256), 0, st, bt, H, H, m->H1, H, sZ,
                     m->wb + OFF_W2, H, (int64_t)NPAD, m->Z2, H, sZ);
  hipLaunchKernelGGL(k_bias_act, ew, dim3(256), 0, st, bt, m->wb, OFF_B2, m->Z2, m->H2,
                     (float*)nullptr, sZ);
  hipLaunchKernelGGL((k_gemm<false, true>), gg, dim3(256), 0, st, bt, H, H, m->H2, H, sZ,
                     m->wb + OFF_W3, H, (int64_t)NPAD, m->Z3, H, sZ);
  hipLaunchKernelGGL(k_bias_act, ew, dim3(256), 0, st, bt, m->wb, OFF_B3, m->Z3,
                     (bf16*)nullptr, m->H3, sZ);
  hipLaunchKernelGGL(k_out_fwd, dim3((unsigned)((bt + 255) / 256), 1, mem), dim3(256), 0, st,
                     bt, m->H3, m->wb, with_loss ? m->T : (const float*)nullptr, m->Y, m->dY,
                     m->loss, sZ);
}

static void mlp_backward(rapt_mlp* m, int bt) {
  hipStream_t st = m->ctx->stream;
  const int mem = m->members;
  const int64_t sZ = (int64_t)bt * H;
  const int chunks = bt / RCH;
  const dim3 ew((unsigned)((sZ + 255) / 256), 1, mem);
  const dim3 gw(H / 64, H / 64, mem);   // 256 x 256 weight gradients
  const dim3 ga(H / 64, bt / 64, mem);  // batch x 256 activation gradients
  const dim3 gr(chunks, 1, mem);        // per-chunk partial reductions
  // layer 3
  hipLaunchKernelGGL(k_out_bwd, ew, dim3(256), 0, st, bt, m->dY, m->wb, m->Z3, m->dZb, sZ);
  hipLaunchKernelGGL((k_gemm<true, false>), gw, dim3(256), 0, st, H, H, bt, m->dZb, H, sZ,
                     m->H2, H, sZ, m->grad + OFF_W3, H, (int64_t)NPAD);
  hipLaunchKernelGGL(k_colsum_part, gr, dim3(256), 0, st, bt, m->dZb, m->partB, 0, sZ);
  hipLaunchKernelGGL((k_gemm<false, false>), ga, dim3(256), 0, st, bt, H, H, m->dZb, H, sZ,
                     m->wb + OFF_W3, H, (int64_t)NPAD, m->dH, H, sZ);
  // layer 2
  hipLaunchKernelGGL(k_act_bwd, ew, dim3(256), 0, st, bt, m->dH, m->Z2, m->dZb, sZ);
  hipLaunchKernelGGL((k_gemm<true, false>), gw, dim3(256), 0, st, H, H, bt, m->dZb, H, sZ,
                     m->H1, H, sZ, m->grad + OFF_W2, H, (int64_t)NPAD);
  hipLaunchKernelGGL(k_colsum_part, gr, dim3(256), 0, st, bt, m->dZb, m->partB, 1, sZ);
  hipLaunchKernelGGL(k_sum_bias, dim3(1, 1, mem), dim3(256), 0, st, chunks, m->partB, m->grad);
  hipLaunchKernelGGL((k_gemm<false, false>), ga, dim3(256), 0, st, bt, H, H, m->dZb, H, sZ,
                     m->wb + OFF_W2, H, (int64_t)NPAD, m->dH, H, sZ);
  // layer 1 and the output layer
  hipLaunchKernelGGL(k_small_part, gr, dim3(256), 0, st, bt, m->dY, m->H3, m->dH, m->Z1, m->X,
                     m->partS, sZ);
  hipLaunchKernelGGL(k_sum_small, dim3(1, 1, mem), dim3(256), 0, st, chunks, m->partS, m->grad);
}

extern "C" {

int rapt_mlp_create(rapt_ctx* c, int32_t members, const float* theta0, double mass, double eV,
                    rapt_mlp** out) {
  if (!c || !out || members < 1 || members > MAXM || !theta0 || !(mass > 0.0) || !(eV > 0.0))
    return c ? set_err(c, RAPT_E_ARG, "bad MLP arguments") : RAPT_E_ARG;
  HIP_OK(c, hipSetDevice(c->device));
  rapt_mlp* m = new (std::nothrow) rapt_mlp();
  if (!m) return RAPT_E_OOM;
  m->ctx = c;
  m->members = members;
  m->mass = mass;
  m->eV = eV;
  const size_t np = (size_t)members * NPAD;
  bool ok = hipMalloc(&m->theta, sizeof(float) * np) == hipSuccess &&
            hipMalloc(&m->grad, sizeof(float) * np) == hipSuccess &&
            hipMalloc(&m->m1, sizeof(float) * np) == hipSuccess &&
            hipMalloc(&m->m2, sizeof(float) * np) == hipSuccess &&
            hipMalloc(&m->wb, sizeof(bf16) * np) == hipSuccess &&
            hipMalloc(&m->packed, sizeof(bf16) * members * 2 * PACKED) == hipSuccess &&
            hipMalloc(&m->p1, sizeof(bf16) * members * P1SIZE) == hipSuccess &&
            hipMalloc(&m->norm, sizeof(float) * 6) == hipSuccess;
  if (!ok) {
    rapt_mlp_destroy(m);
    return set_err(c, RAPT_E_OOM, "MLP allocation failed");
  }
  hipStream_t st = c->stream;
  hipLaunchKernelGGL(k_zero_f, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, m->theta,
                     (int64_t)np);
  for (int k = 0; k < members; ++k)
    if (hipMemcpyAsync(m->theta + (size_t)k * NPAD, theta0 + (size_t)k * NPARAM,
                       sizeof(float) * NPARAM, hipMemcpyHostToDevice, st) != hipSuccess) {
      rapt_mlp_destroy(m);
      return set_err(c, RAPT_E_HIP, "MLP upload failed");
    }
  hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, m->theta,
                     m->wb, (int64_t)np);
  hipLaunchKernelGGL(k_zero_f, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, m->m1,
                     (int64_t)np);
  hipLaunchKernelGGL(k_zero_f, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, m->m2,
                     (int64_t)np);
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(m->norm, m->norm_h, sizeof(float) * 6, hipMemcpyHostToDevice, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess) {
    rapt_mlp_destroy(m);
    return set_err(c, RAPT_E_HIP, "MLP initialisation failed");
  }
  *out = m;
  return RAPT_OK;
}

int rapt_mlp_destroy(rapt_mlp* m) {
  if (!m) return RAPT_E_ARG;
  if (m->ctx) (void)hipSetDevice(m->ctx->device);
  mlp_free_work(m);
  void* p[] = {m->theta, m->grad, m->m1, m->m2, m->wb, m->packed, m->p1, m->norm};
  for (void* q : p) (void)hipFree(q);
  delete m;
  return RAPT_OK;
}

int rapt_mlp_set_norm(rapt_mlp* m, const float* shift_scale) {
  if (!m || !shift_scale) return RAPT_E_ARG;
  rapt_ctx* c = m->ctx;
  for (int k = 0; k < 3; ++k)
    if (!(shift_scale[3 + k] > 0.0f)) return set_err(c, RAPT_E_ARG, "scales must be > 0");
  memcpy(m->norm_h, shift_scale, sizeof m->norm_h);
  HIP_OK(c, hipMemcpyAsync(m->norm, m->norm_h, sizeof(float) * 6, hipMemcpyHostToDevice,
                           c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return RAPT_OK;
}

int rapt_mlp_set_split(rapt_mlp* m, double eyegradientstep, double Re) {
  if (!m) return RAPT_E_ARG;
  rapt_ctx* c = m->ctx;
  if (eyegradientstep <= 0.0) {
    m->split.on = 0;
    return RAPT_OK;
  }
  if (!(Re > 0.0)) return set_err(c, RAPT_E_ARG, "Re must be > 0");
  if (!c->field_set) return set_err(c, RAPT_E_STATE, "no field set");
  const int md = c->field.model;
  if (md != 1 && md != 2 && md != 5 && md != 8)
    return set_err(c, RAPT_E_ARG, "the regime split needs a dipole-type field (model %d)", md);
  m->split.on = 1;
  m->split.Re = Re;
  m->split.d = eyegradientstep;
  m->split.F = c->field;
  return RAPT_OK;
}

int rapt_mlp_targets(rapt_mlp* m, int64_t n, const float* query, const double* labels,
                     int normalised, float* X, float* T) {
  if (!m || n <= 0 || !query || !labels || !X || !T) return m ? set_err(m->ctx, RAPT_E_ARG,
                                                                        "bad targets call")
                                                              : RAPT_E_ARG;
  rapt_ctx* c = m->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  float *dq = nullptr, *dX = nullptr, *dT = nullptr;
  double* dl = nullptr;
  int rc = RAPT_OK;
  if (hipMalloc(&dq, sizeof(float) * 3 * n) != hipSuccess ||
      hipMalloc(&dl, sizeof(double) * 3 * n) != hipSuccess ||
      hipMalloc(&dX, sizeof(float) * 3 * n) != hipSuccess ||
      hipMalloc(&dT, sizeof(float) * 3 * n) != hipSuccess) {
    rc = set_err(c, RAPT_E_OOM, "targets allocation failed");
  } else if (hipMemcpyAsync(dq, query, sizeof(float) * 3 * n, hipMemcpyHostToDevice,
                            c->stream) != hipSuccess ||
             hipMemcpyAsync(dl, labels, sizeof(double) * 3 * n, hipMemcpyHostToDevice,
                            c->stream) != hipSuccess) {
    rc = set_err(c, RAPT_E_HIP, "targets copy failed");
  } else {
    hipLaunchKernelGGL(k_targets, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream,
                       n, dq, dl, normalised ? m->norm : (const float*)nullptr, m->mass, m->eV,
                       m->split, dX, dT);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(X, dX, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipMemcpyAsync(T, dT, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = set_err(c, RAPT_E_HIP, "targets kernel failed");
  }
  (void)hipFree(dq);
  (void)hipFree(dl);
  (void)hipFree(dX);
  (void)hipFree(dT);
  return rc;
}

int rapt_mlp_train(rapt_mlp* m, int64_t n, const float* X, const float* T, int32_t batch,
                   int32_t steps, const int64_t* order, float lr0, float lr1, double* loss_out,
                   float* grad_out) {
  if (!m) return RAPT_E_ARG;
  rapt_ctx* c = m->ctx;
  if (n <= 0 || !X || !T || batch < 64 || batch % 64 != 0 || steps < 0 || !order)
    return set_err(c, RAPT_E_ARG, "train: batch must be a positive multiple of 64");
  HIP_OK(c, hipSetDevice(c->device));
  int rc = mlp_work(m, batch);
  if (rc) return rc;
  hipStream_t st = c->stream;
  float *dXall = nullptr, *dTall = nullptr;
  int64_t* dorder = nullptr;
  const int64_t nidx = (int64_t)steps * batch;
  if (hipMalloc(&dXall, sizeof(float) * 3 * n) != hipSuccess ||
      hipMalloc(&dTall, sizeof(float) * 3 * n) != hipSuccess ||
      hipMalloc(&dorder, sizeof(int64_t) * (nidx > 0 ? nidx : 1)) != hipSuccess) {
    (void)hipFree(dXall);
    (void)hipFree(dTall);
    (void)hipFree(dorder);
    return set_err(c, RAPT_E_OOM, "training data allocation failed");
  }
  for (int64_t i = 0; i < nidx; ++i)
    if (order[i] < 0 || order[i] >= n) {
      (void)hipFree(dXall);
      (void)hipFree(dTall);
      (void)hipFree(dorder);
      return set_err(c, RAPT_E_ARG, "sample index out of range");
    }
  rc = RAPT_OK;
  if (hipMemcpyAsync(dXall, X, sizeof(float) * 3 * n, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(dTall, T, sizeof(float) * 3 * n, hipMemcpyHostToDevice, st) != hipSuccess ||
      (nidx > 0 && hipMemcpyAsync(dorder, order, sizeof(int64_t) * nidx, hipMemcpyHostToDevice,
                                  st) != hipSuccess))
    rc = set_err(c, RAPT_E_HIP, "training data copy failed");
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int s = 0; s < steps && rc == RAPT_OK; ++s) {
    // gather the batch
    hipLaunchKernelGGL(k_gather, dim3((batch + 255) / 256), dim3(256), 0, st, batch,
                       dorder + (int64_t)s * batch, dXall, dTall, m->X, m->T);
    const bool last = s == steps - 1;
    if (last && loss_out)
      hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, st, m->loss, m->members);
    mlp_forward(m, batch, true);
    mlp_backward(m, batch);
    if (last && grad_out)
      for (int k = 0; k < m->members; ++k)
        if (hipMemcpyAsync(grad_out + (size_t)k * NPARAM, m->grad + (size_t)k * NPAD,
                           sizeof(float) * NPARAM, hipMemcpyDeviceToHost, st) != hipSuccess)
          rc = set_err(c, RAPT_E_HIP, "gradient copy failed");
    // cosine learning-rate schedule from lr0 to lr1 over this call
    const float frac = steps > 1 ? (float)s / (float)(steps - 1) : 0.0f;
    const float lr = lr1 + 0.5f * (lr0 - lr1) * (1.0f + cosf(3.14159265f * frac));
    ++m->step;
    const float c1 = 1.0f - powf(b1, (float)m->step), c2 = 1.0f - powf(b2, (float)m->step);
    if (lr0 > 0.0f)
      hipLaunchKernelGGL(k_adam, dim3((NPARAM + 255) / 256, 1, m->members), dim3(256), 0, st,
                         m->theta, m->grad, m->m1, m->m2, m->wb, lr, b1, b2, eps, c1, c2);
  }
  if (rc == RAPT_OK && loss_out) {
    double lh[16];
    if (hipMemcpyAsync(lh, m->loss, sizeof(double) * m->members, hipMemcpyDeviceToHost, st) !=
        hipSuccess || hipStreamSynchronize(st) != hipSuccess)
      rc = set_err(c, RAPT_E_HIP, "loss copy failed");
    for (int k = 0; k < m->members && rc == RAPT_OK; ++k) loss_out[k] = lh[k] / (3.0 * batch);
  }
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
    rc = rc ? rc : set_err(c, RAPT_E_HIP, "training kernels failed");
  m->packed_ok = false;
  (void)hipFree(dXall);
  (void)hipFree(dTall);
  (void)hipFree(dorder);
  return rc;
}

int rapt_mlp_get_params(rapt_mlp* m, float* theta) {
  if (!m || !theta) return RAPT_E_ARG;
  rapt_ctx* c = m->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  for (int k = 0; k < m->members; ++k)
    HIP_OK(c, hipMemcpyAsync(theta + (size_t)k * NPARAM, m->theta + (size_t)k * NPAD,
                             sizeof(float) * NPARAM, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return RAPT_OK;
}

// Forward pass through the training path (GEMM kernels) for n <= workspace
// batch rows of features X; Y out [members][n][3].
int rapt_mlp_forward(rapt_mlp* m, int64_t n, const float* X, float* Y) {
  if (!m || n <= 0 || !X || !Y) return m ? set_err(m->ctx, RAPT_E_ARG, "bad forward call")
                                         : RAPT_E_ARG;
  rapt_ctx* c = m->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  const int bt = (int)((n + 63) / 64 * 64);
  int rc = mlp_work(m, bt);
  if (rc) return rc;
  hipStream_t st = c->stream;
  hipLaunchKernelGGL(k_zero_f, dim3((unsigned)((3 * bt + 255) / 256)), dim3(256), 0, st, m->X,
                     (int64_t)3 * bt);
  HIP_OK(c, hipMemcpyAsync(m->X, X, sizeof(float) * 3 * n, hipMemcpyHostToDevice, st));
  mlp_forward(m, bt, false);
  HIP_OK(c, hipGetLastError());
  for (int k = 0; k < m->members; ++k)
    HIP_OK(c, hipMemcpyAsync(Y + (size_t)k * n * 3, m->Y + (size_t)k * bt * 3,
                             sizeof(float) * 3 * n, hipMemcpyDeviceToHost, st));
  HIP_OK(c, hipStreamSynchronize(st));
  return RAPT_OK;
}

// Ensemble inference with the fused kernel.  query [n][3] (host); out [n][3]
// physical (T_b [s], |omega_d| [rad/s], mirror latitude [deg]); flag [n]:
// 1 where the members' relative spread exceeds thresh on any output; raw
// (optional) [members][n][3] normalised member outputs.  kernel_ms: time of
// the inference kernel alone.
int rapt_mlp_infer(rapt_mlp* m, int64_t n, const float* query, float thresh, float* out,
                   int32_t* flag, float* raw, double* kernel_ms) {
  if (!m || n <= 0 || !query || !out || !flag)
    return m ? set_err(m->ctx, RAPT_E_ARG, "bad infer call") : RAPT_E_ARG;
  rapt_ctx* c = m->ctx;
  HIP_OK(c, hipSetDevice(c->device));
  hipStream_t st = c->stream;
  if (!m->packed_ok) {
    for (int layer = 0; layer < 2; ++layer)
      hipLaunchKernelGGL(k_pack, dim3(PACKED / 256, 1, m->members), dim3(256), 0, st, m->wb,
                         m->packed + (int64_t)layer * PACKED, layer == 0 ? OFF_W2 : OFF_W3);
    hipLaunchKernelGGL(k_pack1, dim3(P1SIZE / 256, 1, m->members), dim3(256), 0, st, m->wb,
                       m->p1);
    HIP_OK(c, hipGetLastError());
    m->packed_ok = true;
  }
  float *dq = nullptr, *dout = nullptr, *draw = nullptr, *dprep = nullptr;
  int32_t* dflag = nullptr;
  int rc = RAPT_OK;
  if (hipMalloc(&dq, sizeof(float) * 3 * n) != hipSuccess ||
      hipMalloc(&dprep, sizeof(float) * 8 * n) != hipSuccess ||
      hipMalloc(&dout, sizeof(float) * 3 * n) != hipSuccess ||
      hipMalloc(&dflag, sizeof(int32_t) * n) != hipSuccess ||
      (raw && hipMalloc(&draw, sizeof(float) * 3 * n * m->members) != hipSuccess)) {
    rc = set_err(c, RAPT_E_OOM, "inference allocation failed");
  } else if (hipMemcpyAsync(dq, query, sizeof(float) * 3 * n, hipMemcpyHostToDevice, st) !=
             hipSuccess) {
    rc = set_err(c, RAPT_E_HIP, "query copy failed");
  } else {
    InferArgs A;
    A.n = n;
    A.members = m->members;
    A.wb = m->wb;
    A.packed = m->packed;
    A.p1 = m->p1;
    A.norm = m->norm;
    A.thresh = thresh;
    A.prep = dprep;
    A.out = dout;
    A.raw = draw;
    A.flag = dflag;
    int dev_cus = 256;
    (void)hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, c->device);
    // persistent: one block per CU, each taking chunks of IW x QW queries
    const int64_t groups = (n + QW - 1) / QW;
    int64_t blocks = (groups + IW - 1) / IW;
    if (blocks > dev_cus) blocks = dev_cus;
    HIP_OK(c, hipEventRecord(c->ev0, st));
    hipLaunchKernelGGL(k_prep, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, dq,
                       m->split, m->mass, m->eV, dprep, dflag);
    hipLaunchKernelGGL(k_infer, dim3((unsigned)blocks), dim3(IW * 64), 0, st, A);
    HIP_OK(c, hipEventRecord(c->ev1, st));
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(out, dout, sizeof(float) * 3 * n, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipMemcpyAsync(flag, dflag, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        (raw && hipMemcpyAsync(raw, draw, sizeof(float) * 3 * n * m->members,
                               hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess) {
      rc = set_err(c, RAPT_E_HIP, "inference kernel failed");
    } else if (kernel_ms) {
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
      *kernel_ms = ms;
    }
  }
  (void)hipFree(dq);
  (void)hipFree(dprep);
  (void)hipFree(dout);
  (void)hipFree(dflag);
  (void)hipFree(draw);
  return rc;
}

}  // extern "C"
