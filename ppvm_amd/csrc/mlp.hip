// mlp.hip — ensemble MLP surrogate trained and evaluated with hand-written
// BF16 MFMA kernels (FP32 accumulation) on gfx950.
//
// Network per member: in (padded to kInPad) -> H -> H -> H -> 1, GELU after
// the hidden layers.  Weights and activations are BF16; accumulation, biases,
// pre-activations, gradients of the weights and the Adam state are FP32
// (master weights are FP32 and re-rounded to BF16 after every update).
// All members of the ensemble run in the same launches (grid z = member).
//
// GEMM: C[m][n] = sum_k A(m,k) B(n,k) with arbitrary element strides for A and
// B (so X W^T, dY W and dY^T X are the same kernel).  A block computes a
// 64x64 tile with 4 wave64s (one 32x32 sub-tile each); 64x32 slices of A and B
// are staged through LDS and fed to v_mfma_f32_32x32x16_bf16 (8 BF16 per lane
// per operand).  Lane layout of the 32x32 result: lane l, accumulator i ->
// row 8*(i/4) + 4*(l/32) + i%4, column l%32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ppvm_hip.h"

typedef unsigned short bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint64_t u64;
typedef uint32_t u32;

namespace {

thread_local std::string m_err;
int mfail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  m_err = buf;
  return code;
}
#define MCHK(expr)                                                                        \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) return mfail(PPVM_EHIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

constexpr int kInPad = 32;   // input features padded to a multiple of 16
constexpr int kOutPad = 32;  // output layer padded to one 32-column tile
constexpr int kLayers = 4;

// FP32 -> BF16 with round-to-nearest-even on the bit pattern (NaN kept quiet).
__host__ __device__ inline bf16 f2bf(float f) {
  u32 b;
  std::memcpy(&b, &f, 4);
  if ((b & 0x7fffffffu) > 0x7f800000u) return (bf16)((b >> 16) | 0x40u);
  b += 0x7fffu + ((b >> 16) & 1u);
  return (bf16)(b >> 16);
}
__host__ __device__ inline float bf2f(bf16 h) {
  u32 b = (u32)h << 16;
  float f;
  std::memcpy(&f, &b, 4);
  return f;
}

// GELU(x) = x * Phi(x) with Phi the standard normal CDF.
__device__ inline float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ inline float gelu_grad(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.3989422804014327f * expf(-0.5f * x * x);
}

enum Epi : int { E_FWD_GELU = 0, E_FWD_LINEAR = 1, E_BWD_GELU = 2, E_STORE = 3 };

struct GemmArgs {
  int M, N, K;
  const bf16* A;
  long long sam, sak, zA;   // element strides of A(m,k); member stride
  const bf16* B;
  long long sbn, sbk, zB;
  float* C;                 // fp32 output (pre-activation / linear output / dW)
  long long ldc, zC;
  const float* bias;        // per n (forward)
  long long zbias;
  bf16* Cb;                 // bf16 output (activation / gradient)
  long long ldcb, zCb;
  const float* pre;         // pre-activation for the GELU derivative (backward)
  long long ldpre, zpre;
};

constexpr int kTileM = 64, kTileN = 64, kTileK = 32, kLdsPad = 8;

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  __shared__ __attribute__((aligned(16))) bf16 As[kTileM][kTileK + kLdsPad];
  __shared__ __attribute__((aligned(16))) bf16 Bs[kTileN][kTileK + kLdsPad];
  const int z = blockIdx.z;
  const bf16* A = g.A + z * g.zA;
  const bf16* B = g.B + z * g.zB;
  const int m0 = blockIdx.y * kTileM, n0 = blockIdx.x * kTileN;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f32x16 acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
  for (int k0 = 0; k0 < g.K; k0 += kTileK) {
    // stage A[m0..m0+64) x [k0..k0+32) and B likewise: 2048 elements each, 8 per thread
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + e * 256;
      const int r = idx / kTileK, c = idx % kTileK;
      const int m = m0 + r, n = n0 + r, k = k0 + c;
      As[r][c] = (m < g.M && k < g.K) ? A[m * g.sam + k * g.sak] : (bf16)0;
      Bs[r][c] = (n < g.N && k < g.K) ? B[n * g.sbn + k * g.sbk] : (bf16)0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTileK; kk += 16) {
      const int kb = kk + 8 * (lane >> 5);
      const u16x8 av = *reinterpret_cast<const u16x8*>(&As[wm + (lane & 31)][kb]);
      const u16x8 bv = *reinterpret_cast<const u16x8*>(&Bs[wn + (lane & 31)][kb]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, av),
                                                    __builtin_bit_cast(bf16x8, bv), acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int n = n0 + wn + (lane & 31);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int m = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
    if (m >= g.M || n >= g.N) continue;
    float c = acc[i];
    if (EPI == E_FWD_GELU) {
      c += g.bias[z * g.zbias + n];
      g.C[z * g.zC + (long long)m * g.ldc + n] = c;
      g.Cb[z * g.zCb + (long long)m * g.ldcb + n] = f2bf(gelu(c));
    } else if (EPI == E_FWD_LINEAR) {
      c += g.bias[z * g.zbias + n];
      g.C[z * g.zC + (long long)m * g.ldc + n] = c;
    } else if (EPI == E_BWD_GELU) {
      c *= gelu_grad(g.pre[z * g.zpre + (long long)m * g.ldpre + n]);
      g.Cb[z * g.zCb + (long long)m * g.ldcb + n] = f2bf(c);
    } else {
      g.C[z * g.zC + (long long)m * g.ldc + n] = c;
    }
  }
}

// MSE loss on output column 0: d = 2 (out - y) / n as BF16 (other columns 0);
// per-member loss sums into loss[member].
__global__ void k_loss(const float* out, long long ldo, long long zo, const float* y, int n, bf16* d,
                       long long ldd, long long zd, float* loss) {
  __shared__ float red[256];
  const int z = blockIdx.y;
  float s = 0.f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float e = out[z * zo + (long long)i * ldo] - y[i];
    s += e * e;
    d[z * zd + (long long)i * ldd] = f2bf(2.0f * e / (float)n);
    for (int c = 1; c < kOutPad; ++c) d[z * zd + (long long)i * ldd + c] = 0;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(&loss[z], red[0] / (float)n);
}

// db[n] = sum_m d[m][n]
__global__ void k_colsum(const bf16* d, long long ldd, long long zd, int M, int N, float* db, long long zdb) {
  __shared__ float red[256];
  const int n = blockIdx.x, z = blockIdx.y;
  float s = 0.f;
  for (int m = threadIdx.x; m < M; m += blockDim.x) s += bf2f(d[z * zd + (long long)m * ldd + n]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0 && n < N) db[z * zdb + n] = red[0];
}

struct AdamArgs {
  float* p;        // fp32 master parameters (all members, contiguous)
  const float* g;  // gradients, same layout
  float* m;
  float* v;
  long long n;
  float lr, b1, b2, eps;
  int* step;       // device step counter (incremented by k_step before k_adam)
  int total_steps; // cosine decay horizon
};

__global__ void k_step(int* step) { *step += 1; }

// Adam with cosine learning-rate decay (step counter on the device, so the
// update can be replayed inside a graph).
__global__ void k_adam(AdamArgs a) {
  const int t = *a.step;
  const float frac = fminf(1.0f, (float)t / (float)a.total_steps);
  const float lr = a.lr * 0.5f * (1.0f + cosf(3.14159265358979f * frac)) + 1e-6f;
  const float c1 = 1.0f - powf(a.b1, (float)t), c2 = 1.0f - powf(a.b2, (float)t);
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < a.n; i += (long long)gridDim.x * blockDim.x) {
    const float g = a.g[i];
    const float m = a.b1 * a.m[i] + (1.0f - a.b1) * g;
    const float v = a.b2 * a.v[i] + (1.0f - a.b2) * g * g;
    a.m[i] = m;
    a.v[i] = v;
    a.p[i] -= lr * (m / c1) / (sqrtf(v / c2) + a.eps);
  }
}

__global__ void k_to_bf16(const float* src, bf16* dst, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    dst[i] = f2bf(src[i]);
}

__global__ void k_zero_f32(float* p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.f;
}

// ensemble mean / std of output column 0
__global__ void k_ensemble(const float* out, long long ldo, long long zo, int n, int members, float* mean,
                           float* sd) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // two passes (mean, then squared deviations): the spread is small next to
  // the mean, and E[v^2] - mu^2 would cancel most of its digits
  float s = 0.f;
  for (int z = 0; z < members; ++z) s += out[z * zo + (long long)i * ldo];
  const float mu = s / members;
  float s2 = 0.f;
  for (int z = 0; z < members; ++z) {
    const float dv = out[z * zo + (long long)i * ldo] - mu;
    s2 += dv * dv;
  }
  mean[i] = mu;
  sd[i] = sqrtf(s2 / members);
}

u64 sm64(u64& s) {
  s += 0x9e3779b97f4a7c15ull;
  u64 z = s;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
double unif(u64& s) { return ((sm64(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
double normal(u64& s) { return std::sqrt(-2.0 * std::log(unif(s))) * std::cos(6.283185307179586 * unif(s)); }

}  // namespace

struct ppvm_mlp {
  u32 in_dim = 0, hid = 0, members = 0, max_batch = 0;
  hipStream_t stream = nullptr;
  // layer l: N_l x K_l weights, N_l bias
  int K[kLayers], N[kLayers];
  long long woff[kLayers], boff[kLayers];  // offsets in the per-member parameter vector
  long long per_member = 0;                // parameters per member
  float* p = nullptr;   // fp32 master [members][per_member]
  float* g = nullptr;   // gradients
  float* m = nullptr;
  float* v = nullptr;
  bf16* pb = nullptr;   // bf16 copy of p
  int* step = nullptr;
  // activations: act[l] = input of layer l (bf16, [members][B][K_l]); pre[l] fp32 [members][B][N_l];
  // dpre[l] bf16 [members][B][N_l]
  bf16* act[kLayers + 1] = {};
  float* pre[kLayers] = {};
  bf16* dpre[kLayers] = {};
  float* x32 = nullptr;  // staged inputs fp32 [B][in]
  float* y = nullptr;
  float* loss = nullptr;
  float* mean = nullptr;
  float* sd = nullptr;
  hipGraphExec_t graph = nullptr;
  u32 graph_n = 0;
  int total_steps = 1;
};

namespace {

long long act_members_stride(const ppvm_mlp* h, int l) { return (long long)h->max_batch * h->K[l]; }

template <int EPI>
int gemm(ppvm_mlp* h, const GemmArgs& a) {
  dim3 grid((a.N + kTileN - 1) / kTileN, (a.M + kTileM - 1) / kTileM, h->members);
  hipLaunchKernelGGL(k_gemm<EPI>, grid, dim3(256), 0, h->stream, a);
  MCHK(hipGetLastError());
  return PPVM_OK;
}

// Forward pass of all members over n staged rows (act[0] already holds the
// bf16 inputs).  Output (fp32, column 0 of kOutPad) in pre[kLayers-1].
int forward(ppvm_mlp* h, int n) {
  int rc;
  for (int l = 0; l < kLayers; ++l) {
    GemmArgs a{};
    a.M = n;
    a.N = h->N[l];
    a.K = h->K[l];
    a.A = h->act[l];
    a.sam = h->K[l];
    a.sak = 1;
    a.zA = act_members_stride(h, l);
    a.B = h->pb + h->woff[l];
    a.sbn = h->K[l];
    a.sbk = 1;
    a.zB = h->per_member;
    a.C = h->pre[l];
    a.ldc = h->N[l];
    a.zC = (long long)h->max_batch * h->N[l];
    a.bias = h->p + h->boff[l];
    a.zbias = h->per_member;
    if (l + 1 < kLayers) {
      a.Cb = h->act[l + 1];
      a.ldcb = h->N[l];
      a.zCb = act_members_stride(h, l + 1);
      if ((rc = gemm<E_FWD_GELU>(h, a))) return rc;
    } else {
      if ((rc = gemm<E_FWD_LINEAR>(h, a))) return rc;
    }
  }
  return PPVM_OK;
}

// Backward pass + gradients into h->g (loss gradient in dpre[kLayers-1]).
int backward(ppvm_mlp* h, int n) {
  int rc;
  for (int l = kLayers - 1; l >= 0; --l) {
    // dW_l[nn][k] = sum_m dpre_l[m][nn] act_l[m][k]
    GemmArgs a{};
    a.M = h->N[l];
    a.N = h->K[l];
    a.K = n;
    a.A = h->dpre[l];
    a.sam = 1;
    a.sak = h->N[l];
    a.zA = (long long)h->max_batch * h->N[l];
    a.B = h->act[l];
    a.sbn = 1;
    a.sbk = h->K[l];
    a.zB = act_members_stride(h, l);
    a.C = h->g + h->woff[l];
    a.ldc = h->K[l];
    a.zC = h->per_member;
    if ((rc = gemm<E_STORE>(h, a))) return rc;
    hipLaunchKernelGGL(k_colsum, dim3(h->N[l], h->members), dim3(256), 0, h->stream, h->dpre[l],
                       (long long)h->N[l], (long long)h->max_batch * h->N[l], n, h->N[l], h->g + h->boff[l],
                       h->per_member);
    MCHK(hipGetLastError());
    if (l == 0) break;
    // dpre_{l-1}[m][k] = (sum_nn dpre_l[m][nn] W_l[nn][k]) * gelu'(pre_{l-1}[m][k])
    GemmArgs b{};
    b.M = n;
    b.N = h->K[l];
    b.K = h->N[l];
    b.A = h->dpre[l];
    b.sam = h->N[l];
    b.sak = 1;
    b.zA = (long long)h->max_batch * h->N[l];
    b.B = h->pb + h->woff[l];
    b.sbn = 1;
    b.sbk = h->K[l];
    b.zB = h->per_member;
    b.Cb = h->dpre[l - 1];
    b.ldcb = h->N[l - 1];
    b.zCb = (long long)h->max_batch * h->N[l - 1];
    b.pre = h->pre[l - 1];
    b.ldpre = h->N[l - 1];
    b.zpre = (long long)h->max_batch * h->N[l - 1];
    if ((rc = gemm<E_BWD_GELU>(h, b))) return rc;
  }
  return PPVM_OK;
}

int stage_inputs(ppvm_mlp* h, const float* X, const float* Y, u32 n) {
  // inputs padded to kInPad columns, rounded to bf16, replicated per member
  std::vector<bf16> xb((size_t)n * kInPad, 0);
  for (u32 i = 0; i < n; ++i)
    for (u32 j = 0; j < h->in_dim; ++j) xb[(size_t)i * kInPad + j] = f2bf(X[(size_t)i * h->in_dim + j]);
  for (u32 z = 0; z < h->members; ++z)
    MCHK(hipMemcpyAsync(h->act[0] + z * act_members_stride(h, 0), xb.data(), xb.size() * 2, hipMemcpyHostToDevice,
                        h->stream));
  if (Y) MCHK(hipMemcpyAsync(h->y, Y, n * 4, hipMemcpyHostToDevice, h->stream));
  return PPVM_OK;
}

int loss_and_grad(ppvm_mlp* h, int n) {
  hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, h->stream, h->loss, (long long)h->members);
  hipLaunchKernelGGL(k_loss, dim3(std::min(64, (n + 255) / 256), h->members), dim3(256), 0, h->stream,
                     h->pre[kLayers - 1], (long long)kOutPad, (long long)h->max_batch * kOutPad, h->y, n,
                     h->dpre[kLayers - 1], (long long)kOutPad, (long long)h->max_batch * kOutPad, h->loss);
  MCHK(hipGetLastError());
  return PPVM_OK;
}

}  // namespace

extern "C" {

const char* ppvm_mlp_last_error(void) { return m_err.c_str(); }

int ppvm_mlp_create(uint32_t in_dim, uint32_t hidden, uint32_t members, uint32_t max_batch, uint64_t seed,
                    ppvm_mlp_handle* out) {
  if (!out) return mfail(PPVM_EINVAL, "null output");
  *out = nullptr;
  if (in_dim == 0 || in_dim > (u32)kInPad) return mfail(PPVM_EINVAL, "in_dim must be in 1..%d", kInPad);
  if (hidden == 0 || hidden % 64) return mfail(PPVM_EINVAL, "hidden must be a positive multiple of 64");
  if (members == 0 || members > 16) return mfail(PPVM_EINVAL, "members must be in 1..16");
  if (max_batch == 0 || max_batch > (1u << 20)) return mfail(PPVM_EINVAL, "max_batch out of range");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return mfail(PPVM_ENODEV, "no HIP device available");
  }
  ppvm_mlp* h = new ppvm_mlp();
  h->in_dim = in_dim;
  h->hid = hidden;
  h->members = members;
  h->max_batch = max_batch;
  const int Ks[kLayers] = {kInPad, (int)hidden, (int)hidden, (int)hidden};
  const int Ns[kLayers] = {(int)hidden, (int)hidden, (int)hidden, kOutPad};
  long long off = 0;
  for (int l = 0; l < kLayers; ++l) {
    h->K[l] = Ks[l];
    h->N[l] = Ns[l];
    h->woff[l] = off;
    off += (long long)Ns[l] * Ks[l];
    h->boff[l] = off;
    off += Ns[l];
  }
  h->per_member = off;
  const long long P = off * members;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return mfail(PPVM_EHIP, "stream creation failed");
  }
  bool ok = hipMalloc(&h->p, P * 4) == hipSuccess && hipMalloc(&h->g, P * 4) == hipSuccess &&
            hipMalloc(&h->m, P * 4) == hipSuccess && hipMalloc(&h->v, P * 4) == hipSuccess &&
            hipMalloc(&h->pb, P * 2) == hipSuccess && hipMalloc(&h->step, 4) == hipSuccess &&
            hipMalloc(&h->y, (size_t)max_batch * 4) == hipSuccess &&
            hipMalloc(&h->loss, members * 4) == hipSuccess && hipMalloc(&h->mean, (size_t)max_batch * 4) == hipSuccess &&
            hipMalloc(&h->sd, (size_t)max_batch * 4) == hipSuccess;
  for (int l = 0; l < kLayers && ok; ++l) {
    ok = hipMalloc(&h->act[l], (size_t)members * max_batch * h->K[l] * 2) == hipSuccess &&
         hipMalloc(&h->pre[l], (size_t)members * max_batch * h->N[l] * 4) == hipSuccess &&
         hipMalloc(&h->dpre[l], (size_t)members * max_batch * h->N[l] * 2) == hipSuccess;
  }
  if (!ok) {
    (void)hipGetLastError();
    ppvm_mlp_destroy(h);
    return mfail(PPVM_ENOMEM, "MLP buffer allocation failed");
  }
  // He-style init (normal, std sqrt(2 / fan_in)), zero biases; padded input
  // columns and padded output rows start at zero.
  std::vector<float> hp(P, 0.f);
  u64 s = seed;
  for (u32 z = 0; z < members; ++z)
    for (int l = 0; l < kLayers; ++l) {
      const int fan = l == 0 ? (int)in_dim : h->K[l];
      const double sd = std::sqrt((l + 1 < kLayers ? 2.0 : 1.0) / fan);
      for (int nn = 0; nn < h->N[l]; ++nn)
        for (int k = 0; k < h->K[l]; ++k) {
          const bool live = (l != 0 || k < (int)in_dim) && (l + 1 < kLayers || nn == 0);
          hp[z * off + h->woff[l] + (long long)nn * h->K[l] + k] = live ? (float)(sd * normal(s)) : 0.f;
        }
    }
  hipError_t e = hipMemcpy(h->p, hp.data(), P * 4, hipMemcpyHostToDevice);
  std::vector<float> zeros(P, 0.f);
  if (e == hipSuccess) e = hipMemcpy(h->m, zeros.data(), P * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->v, zeros.data(), P * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(h->g, zeros.data(), P * 4, hipMemcpyHostToDevice);
  int zero = 0;
  if (e == hipSuccess) e = hipMemcpy(h->step, &zero, 4, hipMemcpyHostToDevice);
  for (int l = 0; l < kLayers && e == hipSuccess; ++l)
    e = hipMemset(h->act[l], 0, (size_t)members * max_batch * h->K[l] * 2);
  if (e != hipSuccess) {
    ppvm_mlp_destroy(h);
    return mfail(PPVM_EHIP, "MLP initialisation failed: %s", hipGetErrorString(e));
  }
  hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, h->stream, h->p, h->pb, P);
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    ppvm_mlp_destroy(h);
    return mfail(PPVM_EHIP, "MLP initialisation failed");
  }
  *out = h;
  return PPVM_OK;
}

int ppvm_mlp_destroy(ppvm_mlp_handle h) {
  if (!h) return PPVM_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->graph) (void)hipGraphExecDestroy(h->graph);
  void* ptrs[] = {h->p, h->g, h->m, h->v, h->pb, h->step, h->x32, h->y, h->loss, h->mean, h->sd};
  for (void* q : ptrs)
    if (q) (void)hipFree(q);
  for (int l = 0; l < kLayers; ++l) {
    if (h->act[l]) (void)hipFree(h->act[l]);
    if (h->pre[l]) (void)hipFree(h->pre[l]);
    if (h->dpre[l]) (void)hipFree(h->dpre[l]);
  }
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return PPVM_OK;
}

uint64_t ppvm_mlp_param_count(ppvm_mlp_handle h) { return h ? (uint64_t)h->per_member : 0; }

int ppvm_mlp_get_params(ppvm_mlp_handle h, uint32_t member, float* out, size_t cap) {
  if (!h || !out || member >= h->members) return mfail(PPVM_EINVAL, "bad argument");
  if (cap < (size_t)h->per_member) return mfail(PPVM_EINVAL, "buffer too small");
  MCHK(hipStreamSynchronize(h->stream));
  MCHK(hipMemcpy(out, h->p + member * h->per_member, h->per_member * 4, hipMemcpyDeviceToHost));
  return PPVM_OK;
}

int ppvm_mlp_set_params(ppvm_mlp_handle h, uint32_t member, const float* in, size_t n) {
  if (!h || !in || member >= h->members || n != (size_t)h->per_member) return mfail(PPVM_EINVAL, "bad argument");
  MCHK(hipStreamSynchronize(h->stream));
  MCHK(hipMemcpy(h->p + member * h->per_member, in, n * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_to_bf16, dim3(256), dim3(256), 0, h->stream, h->p, h->pb, h->per_member * h->members);
  MCHK(hipStreamSynchronize(h->stream));
  return PPVM_OK;
}

int ppvm_mlp_grads(ppvm_mlp_handle h, const float* X, const float* Y, uint32_t n, uint32_t member, float* grads,
                   size_t cap, float* loss) {
  if (!h || !X || !Y || !grads || member >= h->members) return mfail(PPVM_EINVAL, "bad argument");
  if (n == 0 || n > h->max_batch) return mfail(PPVM_EINVAL, "n out of range");
  if (cap < (size_t)h->per_member) return mfail(PPVM_EINVAL, "buffer too small");
  int rc;
  if ((rc = stage_inputs(h, X, Y, n))) return rc;
  if ((rc = forward(h, (int)n))) return rc;
  if ((rc = loss_and_grad(h, (int)n))) return rc;
  if ((rc = backward(h, (int)n))) return rc;
  MCHK(hipStreamSynchronize(h->stream));
  MCHK(hipMemcpy(grads, h->g + member * h->per_member, h->per_member * 4, hipMemcpyDeviceToHost));
  if (loss) MCHK(hipMemcpy(loss, h->loss + member, 4, hipMemcpyDeviceToHost));
  return PPVM_OK;
}

int ppvm_mlp_train(ppvm_mlp_handle h, const float* X, const float* Y, uint32_t n, uint32_t steps, float lr,
                   uint32_t use_graph, float* final_loss) {
  if (!h || !X || !Y) return mfail(PPVM_EINVAL, "null argument");
  if (n == 0 || n > h->max_batch) return mfail(PPVM_EINVAL, "n out of range");
  if (!(lr > 0.f)) return mfail(PPVM_EINVAL, "lr must be positive");
  int rc;
  if ((rc = stage_inputs(h, X, Y, n))) return rc;
  h->total_steps = (int)std::max<u32>(1, steps);
  int zero = 0;
  const long long P = h->per_member * h->members;
  // each call is a fresh Adam run: step counter and both moments start at zero
  MCHK(hipMemcpyAsync(h->step, &zero, 4, hipMemcpyHostToDevice, h->stream));
  hipLaunchKernelGGL(k_zero_f32, dim3(512), dim3(256), 0, h->stream, h->m, P);
  hipLaunchKernelGGL(k_zero_f32, dim3(512), dim3(256), 0, h->stream, h->v, P);
  MCHK(hipGetLastError());
  MCHK(hipStreamSynchronize(h->stream));
  AdamArgs a{h->p, h->g, h->m, h->v, P, lr, 0.9f, 0.999f, 1e-8f, h->step, h->total_steps};
  auto one = [&](hipStream_t st) -> int {
    int r;
    if ((r = forward(h, (int)n))) return r;
    if ((r = loss_and_grad(h, (int)n))) return r;
    if ((r = backward(h, (int)n))) return r;
    hipLaunchKernelGGL(k_step, dim3(1), dim3(1), 0, st, h->step);
    hipLaunchKernelGGL(k_adam, dim3(512), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_to_bf16, dim3(512), dim3(256), 0, st, h->p, h->pb, P);
    MCHK(hipGetLastError());
    return PPVM_OK;
  };
  if (use_graph) {
    if (h->graph) {
      (void)hipGraphExecDestroy(h->graph);
      h->graph = nullptr;
    }
    hipGraph_t g;
    MCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    rc = one(h->stream);
    hipError_t ce = hipStreamEndCapture(h->stream, &g);
    if (rc) return rc;
    if (ce != hipSuccess) return mfail(PPVM_EHIP, "graph capture failed: %s", hipGetErrorString(ce));
    MCHK(hipGraphInstantiate(&h->graph, g, nullptr, nullptr, 0));
    (void)hipGraphDestroy(g);
    for (u32 s = 0; s < steps; ++s) MCHK(hipGraphLaunch(h->graph, h->stream));
  } else {
    for (u32 s = 0; s < steps; ++s)
      if ((rc = one(h->stream))) return rc;
  }
  MCHK(hipStreamSynchronize(h->stream));
  if (final_loss) MCHK(hipMemcpy(final_loss, h->loss, h->members * 4, hipMemcpyDeviceToHost));
  return PPVM_OK;
}

int ppvm_mlp_predict(ppvm_mlp_handle h, const float* X, uint32_t n, float* mean, float* sd, float* per_member) {
  if (!h || !X || !mean || !sd) return mfail(PPVM_EINVAL, "null argument");
  if (n == 0 || n > h->max_batch) return mfail(PPVM_EINVAL, "n out of range");
  int rc;
  if ((rc = stage_inputs(h, X, nullptr, n))) return rc;
  if ((rc = forward(h, (int)n))) return rc;
  hipLaunchKernelGGL(k_ensemble, dim3((n + 255) / 256), dim3(256), 0, h->stream, h->pre[kLayers - 1],
                     (long long)kOutPad, (long long)h->max_batch * kOutPad, (int)n, (int)h->members, h->mean, h->sd);
  MCHK(hipGetLastError());
  MCHK(hipMemcpyAsync(mean, h->mean, n * 4, hipMemcpyDeviceToHost, h->stream));
  MCHK(hipMemcpyAsync(sd, h->sd, n * 4, hipMemcpyDeviceToHost, h->stream));
  if (per_member) {
    std::vector<float> tmp((size_t)n * kOutPad);
    for (u32 z = 0; z < h->members; ++z) {
      MCHK(hipMemcpyAsync(tmp.data(), h->pre[kLayers - 1] + (long long)z * h->max_batch * kOutPad, (size_t)n * kOutPad * 4,
                          hipMemcpyDeviceToHost, h->stream));
      MCHK(hipStreamSynchronize(h->stream));
      for (u32 i = 0; i < n; ++i) per_member[(size_t)z * n + i] = tmp[(size_t)i * kOutPad];
    }
  }
  MCHK(hipStreamSynchronize(h->stream));
  return PPVM_OK;
}

// Plain GEMM through the MFMA kernel, for layout tests: C[M][N] = A[M][K] B[N][K]^T
// (all row-major, bf16 inputs given as fp32 and rounded, fp32 output).
int ppvm_mlp_gemm_check(const float* A, const float* B, int M, int N, int K, float* C) {
  if (!A || !B || !C || M <= 0 || N <= 0 || K <= 0) return mfail(PPVM_EINVAL, "bad argument");
  std::vector<bf16> ab((size_t)M * K), bb((size_t)N * K);
  for (size_t i = 0; i < ab.size(); ++i) ab[i] = f2bf(A[i]);
  for (size_t i = 0; i < bb.size(); ++i) bb[i] = f2bf(B[i]);
  bf16 *dA = nullptr, *dB = nullptr;
  float* dC = nullptr;
  MCHK(hipMalloc(&dA, ab.size() * 2));
  MCHK(hipMalloc(&dB, bb.size() * 2));
  MCHK(hipMalloc(&dC, (size_t)M * N * 4));
  MCHK(hipMemcpy(dA, ab.data(), ab.size() * 2, hipMemcpyHostToDevice));
  MCHK(hipMemcpy(dB, bb.data(), bb.size() * 2, hipMemcpyHostToDevice));
  GemmArgs a{};
  a.M = M;
  a.N = N;
  a.K = K;
  a.A = dA;
  a.sam = K;
  a.sak = 1;
  a.B = dB;
  a.sbn = K;
  a.sbk = 1;
  a.C = dC;
  a.ldc = N;
  dim3 grid((N + kTileN - 1) / kTileN, (M + kTileM - 1) / kTileM, 1);
  hipLaunchKernelGGL(k_gemm<E_STORE>, grid, dim3(256), 0, 0, a);
  MCHK(hipGetLastError());
  MCHK(hipDeviceSynchronize());
  MCHK(hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost));
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  return PPVM_OK;
}

}  // extern "C"
