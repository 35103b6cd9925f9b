// Learned quadrature-rule selection: a small MLP that maps the features of a (collocation point,
// element) pair to the Gauss-Legendre order the pair needs, trained on the GPU.
//
// Network (FP32): kQIn inputs -> 3 hidden layers of kQHid units (tanh) -> kQOut logits.  The
// logits are the classes "precalculated rule of order 2 + c" (c = 0..7) and "adaptive" (c = 8),
// ordered by cost; a pair's rule is the cheapest class whose cumulative softmax probability
// reaches the confidence level, i.e. a rule chosen with a safety margin.  It replaces the
// reference's fitted estimator fbem_qs_n_estimation_standard
// (lib/fbem/src/quasisingular_integration.f90:294-715) in the rule choice of sbie_auto
// (lib/fbem/src/bem_staela3d.f90:549-587, bem_harela3d.f90:1498-1536).
//
// Parameter layout (one contiguous FP32 array): for each layer l = 0..3, W_l stored transposed as
// Wt_l[i][o] (n_in_l x n_out_l, o fastest), then b_l[o].
//
// Kernels: k_qnet_fwdbwd (one workgroup of kQB samples: forward, softmax cross-entropy, backward,
// per-workgroup gradient partial), k_qnet_reduce (partials summed in a fixed order), k_qnet_adam,
// k_qnet_infer (forward + class choice).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "mfbem.h"
#include "mfbem_internal.h"

namespace mfb {

constexpr int kQIn = MFB_QNET_INPUTS;
constexpr int kQHid = 64;
constexpr int kQOut = MFB_QNET_CLASSES;
constexpr int kQB = 64;  // samples per workgroup (one wave)

constexpr int kQLayerIn[4] = {kQIn, kQHid, kQHid, kQHid};
constexpr int kQLayerOut[4] = {kQHid, kQHid, kQHid, kQOut};
__host__ __device__ constexpr int qoff_w(int l) {
  return l == 0 ? 0
                : qoff_w(l - 1) + kQLayerIn[l - 1] * kQLayerOut[l - 1] + kQLayerOut[l - 1];
}
__host__ __device__ constexpr int qoff_b(int l) { return qoff_w(l) + kQLayerIn[l] * kQLayerOut[l]; }
constexpr int kQParams = qoff_w(4);

// ------------------------------------------------------------------------------------------------
// forward of one layer for the workgroup's samples: out[o][s] = act(sum_i Wt[i][o] in[i][s] + b[o])
// (activations stored feature-major, sample fastest); thread t owns sample t
// ------------------------------------------------------------------------------------------------
template <int NI, int NO, bool TANH>
__device__ inline void q_layer(const float* __restrict__ P, int l, const float (*in)[kQB],
                               float (*out)[kQB], int t) {
  const float* Wt = P + qoff_w(l);
  const float* b = P + qoff_b(l);
  constexpr int CH = NO < 16 ? NO : 16;  // outputs per pass (NO is a multiple of CH)
  for (int o0 = 0; o0 < NO; o0 += CH) {
    float acc[CH];
#pragma unroll
    for (int o = 0; o < CH; o++) acc[o] = b[o0 + o];
    for (int i = 0; i < NI; i++) {
      const float x = in[i][t];
#pragma unroll
      for (int o = 0; o < CH; o++) acc[o] = fmaf(Wt[i * NO + o0 + o], x, acc[o]);
    }
#pragma unroll
    for (int o = 0; o < CH; o++) out[o0 + o][t] = TANH ? tanhf(acc[o]) : acc[o];
  }
}

// logits -> probabilities in place (thread t, sample t)
__device__ inline void q_softmax(float (*z)[kQB], int t) {
  float mx = z[0][t];
  for (int o = 1; o < kQOut; o++) mx = fmaxf(mx, z[o][t]);
  float s = 0.f;
  for (int o = 0; o < kQOut; o++) {
    const float e = expf(z[o][t] - mx);
    z[o][t] = e;
    s += e;
  }
  const float inv = 1.f / s;
  for (int o = 0; o < kQOut; o++) z[o][t] *= inv;
}

// Weights are read from LDS by every thread at the same address (broadcast).
struct QShared {
  float P[kQParams];
  float x[kQIn][kQB];
  float h[3][kQHid][kQB];
  float z[kQOut][kQB];
  float d[2][kQHid][kQB];
};

// Forward + softmax cross-entropy + backward for the samples [blockIdx.x*kQB, +kQB) of the
// batch (idx maps batch positions to rows of X / y; rows beyond n are padding with zero weight).
// part[blockIdx.x][p] = sum over the workgroup's samples of dLoss/dparam_p, with
// Loss = (1/nb) sum_s CE(s); lossp[blockIdx.x] = sum of CE over the workgroup's samples.
__global__ void __launch_bounds__(kQB) k_qnet_fwdbwd(const float* __restrict__ params,
                                                     const float* __restrict__ X,
                                                     const int* __restrict__ y,
                                                     const int* __restrict__ idx, int nb,
                                                     float* __restrict__ part,
                                                     float* __restrict__ lossp) {
  extern __shared__ float smem[];
  QShared& S = *reinterpret_cast<QShared*>(smem);
  const int t = threadIdx.x;
  for (int p = t; p < kQParams; p += kQB) S.P[p] = params[p];
  const int s = blockIdx.x * kQB + t;
  const bool valid = s < nb;
  const int row = valid ? (idx ? idx[s] : s) : 0;
  for (int i = 0; i < kQIn; i++) S.x[i][t] = valid ? X[(long long)row * kQIn + i] : 0.f;
  __syncthreads();
  q_layer<kQIn, kQHid, true>(S.P, 0, S.x, S.h[0], t);
  q_layer<kQHid, kQHid, true>(S.P, 1, S.h[0], S.h[1], t);
  q_layer<kQHid, kQHid, true>(S.P, 2, S.h[1], S.h[2], t);
  q_layer<kQHid, kQOut, false>(S.P, 3, S.h[2], S.z, t);
  q_softmax(S.z, t);
  const int lab = valid ? y[row] : 0;
  float ce = valid ? -logf(fmaxf(S.z[lab][t], 1e-30f)) : 0.f;
  // dL/dz = (p - onehot) / nb for valid samples
  const float sc = valid ? 1.f / (float)nb : 0.f;
  for (int o = 0; o < kQOut; o++) S.z[o][t] = (S.z[o][t] - (o == lab ? 1.f : 0.f)) * sc;
  // back through layer 3 into d[0] (w.r.t. h[2]), then layers 2, 1
  {
    const float* Wt = S.P + qoff_w(3);
    for (int i = 0; i < kQHid; i++) {
      float a = 0.f;
      for (int o = 0; o < kQOut; o++) a = fmaf(Wt[i * kQOut + o], S.z[o][t], a);
      const float hv = S.h[2][i][t];
      S.d[0][i][t] = a * (1.f - hv * hv);
    }
  }
  // layer-1 delta (w.r.t. the pre-activation of h[1]) from the layer-2 delta d[0], into d[1]
  {
    const float* Wt = S.P + qoff_w(2);
    for (int i = 0; i < kQHid; i++) {
      float a = 0.f;
      for (int o = 0; o < kQHid; o++) a = fmaf(Wt[i * kQHid + o], S.d[0][o][t], a);
      const float hv = S.h[1][i][t];
      S.d[1][i][t] = a * (1.f - hv * hv);
    }
  }
  __syncthreads();
  // deltas: layer 3 -> S.z, layer 2 -> d[0], layer 1 -> d[1], layer 0 -> d[0] (computed below)
  // gradient partials of layers 3, 2, 1 (their deltas are all available now)
  float* out = part + (long long)blockIdx.x * kQParams;
  auto wgrad = [&](int l, int ni, int no, const float (*inp)[kQB], const float (*del)[kQB]) {
    for (int q = t; q < ni * no; q += kQB) {
      const int i = q / no, o = q % no;
      float a = 0.f;
      for (int k = 0; k < kQB; k++) a = fmaf(inp[i][k], del[o][k], a);
      out[qoff_w(l) + q] = a;
    }
    for (int o = t; o < no; o += kQB) {
      float a = 0.f;
      for (int k = 0; k < kQB; k++) a += del[o][k];
      out[qoff_b(l) + o] = a;
    }
  };
  wgrad(3, kQHid, kQOut, S.h[2], S.z);
  wgrad(2, kQHid, kQHid, S.h[1], S.d[0]);
  wgrad(1, kQHid, kQHid, S.h[0], S.d[1]);
  __syncthreads();
  // delta of layer 0 (w.r.t. its pre-activation) from the layer-1 delta d[1]
  {
    const float* Wt = S.P + qoff_w(1);
    for (int i = 0; i < kQHid; i++) {
      float a = 0.f;
      for (int o = 0; o < kQHid; o++) a = fmaf(Wt[i * kQHid + o], S.d[1][o][t], a);
      const float hv = S.h[0][i][t];
      S.d[0][i][t] = a * (1.f - hv * hv);
    }
  }
  __syncthreads();
  wgrad(0, kQIn, kQHid, S.x, S.d[0]);
  // loss partial
  __shared__ float red[kQB];
  red[t] = ce;
  __syncthreads();
  if (t == 0) {
    float a = 0.f;
    for (int k = 0; k < kQB; k++) a += red[k];
    lossp[blockIdx.x] = a;
  }
}

// grad[p] = sum over workgroups (fixed order) of part[w][p]; loss = sum of lossp / nb
__global__ void k_qnet_reduce(const float* part, int nwg, float* grad, const float* lossp,
                              int nb, float* loss) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < kQParams) {
    float a = 0.f;
    for (int w = 0; w < nwg; w++) a += part[(long long)w * kQParams + p];
    grad[p] = a;
  }
  if (p == 0 && loss) {
    float a = 0.f;
    for (int w = 0; w < nwg; w++) a += lossp[w];
    *loss = a / (float)nb;
  }
}

// Adam (Kingma & Ba 2015): m, v moments; step t >= 1
__global__ void k_qnet_adam(float* params, const float* grad, float* m, float* v, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kQParams) return;
  const float g = grad[p];
  const float mm = b1 * m[p] + (1.f - b1) * g;
  const float vv = b2 * v[p] + (1.f - b2) * g * g;
  m[p] = mm;
  v[p] = vv;
  params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

// Forward for rows [0, n): probabilities (n x kQOut) and/or the chosen class per row: the
// cheapest class whose cumulative probability reaches conf.  FP32 matrix cores
// (v_mfma_f32_16x16x4_f32: exact FP32 products, FP32 accumulation), one wave per 16 rows at a
// time, every layer computed transposed: out^T (16 features x 16 rows) = W^T x in^T, so that
// a lane holds, for row j = lane % 16, the features 16 t + 4 q + r (q = lane / 16, r = 0..3) of
// every 16-feature tile t.  The product's depth index is free to order, so layer l + 1 takes its
// k-step (t, r) over the features {16 t + 4 q + r : q} -- exactly the values each lane already
// holds: activations never leave registers.  The weights are staged in LDS once per workgroup in
// that order (one 16-byte read per lane per k-step, conflict-free).  tanh from one exp:
// sign(x) (1 - e) / (1 + e), e = exp(-2|x|) (absolute error ~1e-7, against ~1e-7 for FP32
// rounding of the sums).
constexpr int kQInfer = 256;            // threads per workgroup (4 waves)
constexpr int kQRowsWg = 4 * 16;        // rows per workgroup pass
constexpr int kQInferGridMax = 2048;    // workgroups (grid-stride over the rows)
static_assert(kQHid == 64 && kQIn <= 8 && kQOut <= 16, "tile shapes of the MFMA forward");
typedef float qf4 __attribute__((ext_vector_type(4)));

// staged parameters (floats): layer 0 [2 s][64 lanes][4 u]; layers 1, 2 [4 t][4 r][64][4 u];
// layer 3 [4 t][64][4 r]; biases [3 layers][4 u][4 q][4 r], then layer 3 [4 q][4 r]
constexpr int kQS0 = 0, kQS1 = kQS0 + 2 * 64 * 4, kQS2 = kQS1 + 16 * 64 * 4,
              kQS3 = kQS2 + 16 * 64 * 4, kQSB = kQS3 + 4 * 64 * 4, kQSB3 = kQSB + 3 * 64,
              kQStaged = kQSB3 + 16;

__device__ inline float q_tanh(float x) {
  const float e = __expf(-2.f * fabsf(x));
  return copysignf(__fdividef(1.f - e, 1.f + e), x);
}

constexpr int kQW[4] = {qoff_w(0), qoff_w(1), qoff_w(2), qoff_w(3)};
constexpr int kQBias[4] = {qoff_b(0), qoff_b(1), qoff_b(2), qoff_b(3)};

__device__ inline void q_stage(const float* __restrict__ params, float* S) {
  for (int e = threadIdx.x; e < kQStaged; e += kQInfer) {
    float v = 0.f;
    if (e < kQS1) {                       // layer 0: [s][lane][u]
      const int u = e & 3, l = (e >> 2) & 63, sk = e >> 8;
      const int f = 4 * sk + (l >> 4), o = 16 * u + (l & 15);
      if (f < kQIn) v = params[kQW[0] + f * kQHid + o];
    } else if (e < kQS3) {                // layers 1, 2: [t][r][lane][u]
      const int layer = e < kQS2 ? 1 : 2, x = e - (layer == 1 ? kQS1 : kQS2);
      const int u = x & 3, l = (x >> 2) & 63, r = (x >> 8) & 3, t = x >> 10;
      const int f = 16 * t + 4 * (l >> 4) + r, o = 16 * u + (l & 15);
      v = params[(layer == 1 ? kQW[1] : kQW[2]) + f * kQHid + o];
    } else if (e < kQSB) {                // layer 3: [t][lane][r]
      const int x = e - kQS3, r = x & 3, l = (x >> 2) & 63, t = x >> 8;
      const int f = 16 * t + 4 * (l >> 4) + r, o = l & 15;
      if (o < kQOut) v = params[kQW[3] + f * kQOut + o];
    } else if (e < kQSB3) {               // hidden biases: [layer][u][q][r] = b[16 u + 4 q + r]
      const int x = e - kQSB, layer = x >> 6, o = x & 63;
      v = params[(layer == 0 ? kQBias[0] : layer == 1 ? kQBias[1] : kQBias[2]) + o];
    } else {                              // output bias: [q][r] = b[4 q + r]
      const int o = e - kQSB3;
      if (o < kQOut) v = params[kQBias[3] + o];
    }
    S[e] = v;
  }
}

// one hidden layer (64 -> 64, tanh) of the wave's 16 rows: h (in, then out) [t][r]
__device__ inline void q_hidden(const float* S, int sw, int sb, int lane, float (&h)[16]) {
  const int q = lane >> 4;
  qf4 acc[4];
#pragma unroll
  for (int u = 0; u < 4; u++) acc[u] = *(const qf4*)&S[sb + 16 * u + 4 * q];
#pragma unroll
  for (int t = 0; t < 4; t++) {
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const qf4 a = *(const qf4*)&S[sw + ((t * 4 + r) * 64 + lane) * 4];
      const float b = h[4 * t + r];
#pragma unroll
      for (int u = 0; u < 4; u++) acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b, acc[u], 0, 0, 0);
    }
    __builtin_amdgcn_sched_barrier(0);   // at most one tile's weights in registers at a time
  }
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int r = 0; r < 4; r++) h[4 * u + r] = q_tanh(acc[u][r]);
}

__global__ void __launch_bounds__(kQInfer) k_qnet_infer(const float* __restrict__ params,
                                                        const float* __restrict__ X, int n,
                                                        float conf, float* __restrict__ probs,
                                                        int* __restrict__ cls) {
  __shared__ __attribute__((aligned(16))) float S[kQStaged];
  q_stage(params, S);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, j = lane & 15;
  for (long long base = ((long long)blockIdx.x * 4 + wave) * 16; base < n;
       base += (long long)gridDim.x * kQRowsWg) {
    const long long row = base + j;
    const bool ok = row < n;
    // layer 0 (inputs padded to 8: k-steps s = 0, 1 over the features 4 s + q)
    float h[16];
    {
      qf4 acc[4];
#pragma unroll
      for (int u = 0; u < 4; u++) acc[u] = *(const qf4*)&S[kQSB + 16 * u + 4 * q];
#pragma unroll
      for (int sk = 0; sk < 2; sk++) {
        const int f = 4 * sk + q;
        const float b = (ok && f < kQIn) ? X[row * kQIn + f] : 0.f;
        const qf4 a = *(const qf4*)&S[kQS0 + (sk * 64 + lane) * 4];
#pragma unroll
        for (int u = 0; u < 4; u++) acc[u] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], b, acc[u], 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int r = 0; r < 4; r++) h[4 * u + r] = q_tanh(acc[u][r]);
    }
    q_hidden(S, kQS1, kQSB + 64, lane, h);
    q_hidden(S, kQS2, kQSB + 128, lane, h);
    // output layer: one tile of 16 logits (kQOut used): lane holds logits 4 q + r of row j
    qf4 z = *(const qf4*)&S[kQSB3 + 4 * q];
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const qf4 a = *(const qf4*)&S[kQS3 + (t * 64 + lane) * 4];
#pragma unroll
      for (int r = 0; r < 4; r++) z = __builtin_amdgcn_mfma_f32_16x16x4f32(a[r], h[4 * t + r], z, 0, 0, 0);
    }
    // softmax over the row's kQOut logits (spread over the lanes j, j + 16, j + 32, j + 48)
    float mx = -INFINITY;
#pragma unroll
    for (int r = 0; r < 4; r++)
      if (4 * q + r < kQOut) mx = fmaxf(mx, z[r]);
    mx = fmaxf(mx, __shfl_xor(mx, 16));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    float ez[4], sum = 0.f;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      ez[r] = 4 * q + r < kQOut ? __expf(z[r] - mx) : 0.f;
      sum += ez[r];
    }
    // the row's sum in class order (one lane group after another), as the per-row loop adds
    float tot = 0.f;
#pragma unroll
    for (int g = 0; g < 4; g++)
      if (4 * g < kQOut) tot += __shfl(sum, j + 16 * g);
    const float inv = 1.f / tot;
    if (probs && ok)
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (4 * q + r < kQOut) probs[row * kQOut + 4 * q + r] = ez[r] * inv;
    if (cls) {
      // lane group 0 gathers the row's probabilities in class order
      float pz[kQOut];
#pragma unroll
      for (int o = 0; o < kQOut; o++) pz[o] = __shfl(ez[o & 3], j + 16 * (o >> 2)) * inv;
      if (q == 0 && ok) {
        float c = 0.f;
        int pick = kQOut - 1;
        bool found = false;
#pragma unroll
        for (int o = 0; o < kQOut; o++) {
          c += pz[o];
          if (!found && c >= conf) {
            pick = o;
            found = true;
          }
        }
        cls[row] = pick;
      }
    }
  }
}

}  // namespace mfb

using namespace mfb;

struct mfb_qnet_s {
  float* params = nullptr;
  float* m = nullptr;
  float* v = nullptr;
  float* grad = nullptr;
  float* part = nullptr;
  float* lossp = nullptr;
  float* loss = nullptr;
  int part_cap = 0;  // workgroups the partial buffer holds
  long long step = 0;
};

namespace {

int qnet_reserve(mfb_qnet_t q, int nwg) {
  if (nwg <= q->part_cap) return MFB_OK;
  if (q->part) (void)hipFree(q->part);
  if (q->lossp) (void)hipFree(q->lossp);
  q->part = nullptr;
  q->lossp = nullptr;
  MFB_HIP(hipMalloc(&q->part, sizeof(float) * (size_t)kQParams * nwg));
  MFB_HIP(hipMalloc(&q->lossp, sizeof(float) * nwg));
  q->part_cap = nwg;
  return MFB_OK;
}

constexpr size_t kQShared = sizeof(QShared);

// inference grid: kQRowsWg rows per workgroup pass, at most kQInferGridMax workgroups
unsigned infer_grid(int n) {
  return (unsigned)std::min<long long>(kQInferGridMax, ((long long)n + kQRowsWg - 1) / kQRowsWg);
}

// the fused kernels use more than the default 64 KB of LDS per workgroup (gfx950: 160 KB per CU)
int qnet_lds_attr() {
  static int done_dev = -1;
  int dev = 0;
  MFB_HIP(hipGetDevice(&dev));
  if (done_dev == dev) return MFB_OK;
  MFB_HIP(hipFuncSetAttribute((const void*)k_qnet_fwdbwd,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kQShared));
  done_dev = dev;
  return MFB_OK;
}

}  // namespace

extern "C" int mfb_qnet_param_count(void) { return kQParams; }

extern "C" int mfb_qnet_create(unsigned long long seed, mfb_qnet_t* out) {
  MFB_ARG(out, "null argument");
  *out = nullptr;
  mfb_qnet_t q = new mfb_qnet_s();
  // Glorot-uniform weights, zero biases, from a seeded host generator
  std::vector<float> h(kQParams, 0.f);
  std::mt19937_64 rng(seed);
  for (int l = 0; l < 4; l++) {
    const float a = std::sqrt(6.f / (float)(kQLayerIn[l] + kQLayerOut[l]));
    std::uniform_real_distribution<float> U(-a, a);
    for (int k = 0; k < kQLayerIn[l] * kQLayerOut[l]; k++) h[qoff_w(l) + k] = U(rng);
  }
  hipError_t e = hipMalloc(&q->params, sizeof(float) * kQParams);
  if (e == hipSuccess) e = hipMalloc(&q->m, sizeof(float) * kQParams);
  if (e == hipSuccess) e = hipMalloc(&q->v, sizeof(float) * kQParams);
  if (e == hipSuccess) e = hipMalloc(&q->grad, sizeof(float) * kQParams);
  if (e == hipSuccess) e = hipMalloc(&q->loss, sizeof(float));
  if (e == hipSuccess) e = hipMemcpy(q->params, h.data(), sizeof(float) * kQParams, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(q->m, 0, sizeof(float) * kQParams);
  if (e == hipSuccess) e = hipMemset(q->v, 0, sizeof(float) * kQParams);
  if (e != hipSuccess) {
    set_error("mfb_qnet_create: %s", hipGetErrorString(e));
    mfb_qnet_destroy(q);
    return MFB_ERR_HIP;
  }
  *out = q;
  return MFB_OK;
}

extern "C" int mfb_qnet_destroy(mfb_qnet_t q) {
  if (!q) return MFB_OK;
  (void)hipFree(q->params);
  (void)hipFree(q->m);
  (void)hipFree(q->v);
  (void)hipFree(q->grad);
  (void)hipFree(q->part);
  (void)hipFree(q->lossp);
  (void)hipFree(q->loss);
  delete q;
  return MFB_OK;
}

extern "C" int mfb_qnet_get_params(mfb_qnet_t q, float* host_params) {
  MFB_ARG(q && host_params, "null argument");
  MFB_HIP(hipMemcpy(host_params, q->params, sizeof(float) * kQParams, hipMemcpyDeviceToHost));
  return MFB_OK;
}

extern "C" int mfb_qnet_set_params(mfb_qnet_t q, const float* host_params) {
  MFB_ARG(q && host_params, "null argument");
  MFB_HIP(hipMemcpy(q->params, host_params, sizeof(float) * kQParams, hipMemcpyHostToDevice));
  MFB_HIP(hipMemset(q->m, 0, sizeof(float) * kQParams));
  MFB_HIP(hipMemset(q->v, 0, sizeof(float) * kQParams));
  q->step = 0;
  return MFB_OK;
}

// Gradient of the mean cross-entropy over rows idx[0..nb) (idx may be null: rows 0..nb) of the
// device arrays X (n x MFB_QNET_INPUTS), y; written to the device array grad (param_count) and
// the loss to *loss (host).  Everything device-side runs on the given stream.
static int qnet_grad(mfb_qnet_t q, const float* X, const int* y, const int* idx, int nb,
                     hipStream_t s) {
  const int nwg = (nb + kQB - 1) / kQB;
  int rc = qnet_lds_attr();
  if (rc) return rc;
  rc = qnet_reserve(q, nwg);
  if (rc) return rc;
  hipLaunchKernelGGL(k_qnet_fwdbwd, dim3(nwg), dim3(kQB), kQShared, s, q->params, X, y, idx, nb,
                     q->part, q->lossp);
  hipLaunchKernelGGL(k_qnet_reduce, dim3((kQParams + 255) / 256), dim3(256), 0, s, q->part, nwg,
                     q->grad, q->lossp, nb, q->loss);
  MFB_HIP(hipGetLastError());
  return MFB_OK;
}

extern "C" int mfb_qnet_gradient(mfb_qnet_t q, const float* X, const int* y, int n, float* grad,
                                 float* loss) {
  MFB_ARG(q && X && y && grad && n > 0, "bad argument");
  int rc = qnet_grad(q, X, y, nullptr, n, 0);
  if (rc) return rc;
  MFB_HIP(hipMemcpy(grad, q->grad, sizeof(float) * kQParams, hipMemcpyDeviceToHost));
  if (loss) MFB_HIP(hipMemcpy(loss, q->loss, sizeof(float), hipMemcpyDeviceToHost));
  return MFB_OK;
}

// Allocation-free gradient: the per-workgroup partials (and loss partials) of k_qnet_fwdbwd live
// in this fixed file-scope device array instead of the network's grow-only buffer.  Every GPU the
// library runs on holds its own copy until the process ends; calls on one GPU share it, so calls
// that overlap on one GPU race.  Each call rewrites every partial it reads (no clearing needed).
constexpr int kQNoAllocRows = 16384;                // largest n of one call
constexpr int kQNoAllocWg = kQNoAllocRows / kQB;   // 256 workgroups
static __device__ __align__(256) float scratch[(size_t)kQNoAllocWg * kQParams + kQNoAllocWg];

namespace {
constexpr int kMaxDevices = 64;
float* g_scratch[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up once per device, with it current):
// the host-side name of the array is not a device address
int scratch_address(float** out) {
  int dev = 0;
  MFB_HIP(hipGetDevice(&dev));
  MFB_ARG(dev >= 0 && dev < kMaxDevices, "device index beyond 64");
  if (!g_scratch[dev]) {
    void* p = nullptr;
    MFB_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[dev] = static_cast<float*>(p);
  }
  *out = g_scratch[dev];
  return MFB_OK;
}
}  // namespace

extern "C" int mfb_qnet_gradient_noalloc(mfb_qnet_t q, const float* X, const int* y, int n,
                                         float* grad, float* loss) {
  MFB_ARG(q && X && y && grad && n > 0, "bad argument");
  if (n > kQNoAllocRows) {
    set_error("mfb_qnet_gradient_noalloc: %d rows exceed the %d rows its fixed scratch holds "
              "(use mfb_qnet_gradient)", n, kQNoAllocRows);
    return MFB_ERR_LIMIT;
  }
  float* part = nullptr;
  int rc = scratch_address(&part);
  if (!rc) rc = qnet_lds_attr();
  if (rc) return rc;
  float* lossp = part + (size_t)kQNoAllocWg * kQParams;
  const int nwg = (n + kQB - 1) / kQB;
  hipLaunchKernelGGL(k_qnet_fwdbwd, dim3(nwg), dim3(kQB), kQShared, 0, q->params, X, y,
                     (const int*)nullptr, n, part, lossp);
  hipLaunchKernelGGL(k_qnet_reduce, dim3((kQParams + 255) / 256), dim3(256), 0, 0, part, nwg,
                     q->grad, lossp, n, q->loss);
  MFB_HIP(hipGetLastError());
  MFB_HIP(hipMemcpy(grad, q->grad, sizeof(float) * kQParams, hipMemcpyDeviceToHost));
  if (loss) MFB_HIP(hipMemcpy(loss, q->loss, sizeof(float), hipMemcpyDeviceToHost));
  return MFB_OK;
}

extern "C" int mfb_qnet_train(mfb_qnet_t q, const float* X, const int* y, int n, int epochs,
                              int batch, float lr, unsigned long long seed, float* loss_per_epoch) {
  MFB_ARG(q && X && y && n > 0 && epochs >= 0 && batch > 0, "bad argument");
  hipStream_t s = 0;
  int* didx = nullptr;
  MFB_HIP(hipMalloc(&didx, sizeof(int) * n));
  std::vector<int> perm(n);
  for (int i = 0; i < n; i++) perm[i] = i;
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int rc = MFB_OK;
  for (int ep = 0; ep < epochs && rc == MFB_OK; ep++) {
    std::shuffle(perm.begin(), perm.end(), rng);
    hipError_t e = hipMemcpyAsync(didx, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) { set_error("hipMemcpyAsync: %s", hipGetErrorString(e)); rc = MFB_ERR_HIP; break; }
    double lsum = 0;
    int nbat = 0;
    for (int b0 = 0; b0 < n; b0 += batch) {
      const int nb = std::min(batch, n - b0);
      rc = qnet_grad(q, X, y, didx + b0, nb, s);
      if (rc) break;
      q->step++;
      const float c1 = 1.f - std::pow(b1, (float)q->step), c2 = 1.f - std::pow(b2, (float)q->step);
      hipLaunchKernelGGL(k_qnet_adam, dim3((kQParams + 255) / 256), dim3(256), 0, s, q->params,
                         q->grad, q->m, q->v, lr, b1, b2, eps, c1, c2);
      if (loss_per_epoch) {
        float l = 0;
        e = hipMemcpyAsync(&l, q->loss, sizeof(float), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) { set_error("hip: %s", hipGetErrorString(e)); rc = MFB_ERR_HIP; break; }
        lsum += l;
      }
      nbat++;
    }
    if (loss_per_epoch) loss_per_epoch[ep] = (float)(lsum / std::max(nbat, 1));
  }
  hipError_t e = hipStreamSynchronize(s);
  if (rc == MFB_OK && e != hipSuccess) { set_error("hip: %s", hipGetErrorString(e)); rc = MFB_ERR_HIP; }
  (void)hipFree(didx);
  return rc;
}

extern "C" int mfb_qnet_infer(mfb_qnet_t q, const float* X, int n, float confidence, float* probs,
                              int* classes, void* stream) {
  MFB_ARG(q && X && n > 0 && (probs || classes), "bad argument");
  MFB_ARG(confidence > 0.f && confidence <= 1.f, "confidence must be in (0, 1]");
  hipStream_t s = (hipStream_t)stream;
  int rc = qnet_lds_attr();
  if (rc) return rc;
  hipLaunchKernelGGL(k_qnet_infer, dim3(infer_grid(n)), dim3(kQInfer), 0, s, q->params, X, n,
                     confidence, probs, classes);
  MFB_HIP(hipGetLastError());
  MFB_HIP(hipStreamSynchronize(s));
  return MFB_OK;
}

// device pointer of the parameters (for the assembly path in the same library)
float* mfb_qnet_device_params(struct mfb_qnet_s* q) { return q ? q->params : nullptr; }

// class per row (device arrays, on the caller's stream, no synchronisation): used by the assembly
int mfb_qnet_classes_device(struct mfb_qnet_s* q, const float* X, int n, float confidence,
                            int* classes, hipStream_t s) {
  int rc = qnet_lds_attr();
  if (rc) return rc;
  hipLaunchKernelGGL(k_qnet_infer, dim3(infer_grid(n)), dim3(kQInfer), 0, s, q->params, X, n,
                     confidence, (float*)nullptr, classes);
  MFB_HIP(hipGetLastError());
  return MFB_OK;
}
