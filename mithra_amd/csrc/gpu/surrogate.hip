// Learned surrogate of a scan point: an ensemble of multilayer perceptrons (SiLU hidden layers,
// linear output) trained and evaluated on the GPU.  Weights are kept in FP32 (the Adam master
// copy) and in BF16 (what the products read); every product is a BF16 x BF16 -> FP32 MFMA GEMM
// (v_mfma_f32_32x32x16_bf16) with FP32 accumulation.  Activations enter the products as BF16;
// pre-activations, losses and gradients stay FP32 until they feed the next product.
//
// GEMM: C[M x N] (FP32, row-major, ldc) = op(A)[M x K] * op(B)[K x N], where op(A)(m, k) is read
// at A[m * sam + k * sak] and op(B)(k, n) at B[k * sbk + n * sbn] (BF16): one kernel serves the
// forward product and both backward products (transposed operands) without copies.  Workgroup
// tile 64 x 64, four wavefronts of 32 x 32, K staged through LDS 32 at a time.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <vector>

#include "common.h"

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

namespace {

inline int status(hipError_t e) { return e == hipSuccess ? MITHRA_OK : MITHRA_E_HIP - (int)e; }

#define SG_TRY(expr)                          \
  do {                                        \
    hipError_t _e = (expr);                   \
    if (_e != hipSuccess) return status(_e);  \
  } while (0)

constexpr int kTM = 64, kTN = 64, kTK = 32;

// Lane layout of v_mfma_f32_32x32x16_bf16 (64 lanes, l = lane): A supplies row l % 32 and
// k = 8 (l / 32) .. +7, B column l % 32 and the same k; accumulator i of lane l is row
// 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32.
__global__ __launch_bounds__(256) void k_gemm(const __bf16* __restrict__ A, long sam, long sak,
                                              const __bf16* __restrict__ B, long sbk, long sbn,
                                              float* __restrict__ C, long ldc, int M, int N, int K) {
  __shared__ __attribute__((aligned(16))) __bf16 As[kTM][kTK + 8];  // rows m, k contiguous
  __shared__ __attribute__((aligned(16))) __bf16 Bs[kTN][kTK + 8];  // rows n, k contiguous (B transposed)
  const int m0 = blockIdx.y * kTM, n0 = blockIdx.x * kTN;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  floatx16 acc = {};
  for (int k0 = 0; k0 < K; k0 += kTK) {
    for (int e = tid; e < kTM * kTK; e += 256) {  // stage A(m0.., k0..) and B(k0.., n0..)
      const int r = e / kTK, c = e % kTK;
      const int m = m0 + r, k = k0 + c;
      As[r][c] = (m < M && k < K) ? A[m * sam + k * sak] : (__bf16)0.0f;
      const int n = n0 + r;
      Bs[r][c] = (n < N && k < K) ? B[k * sbk + n * sbn] : (__bf16)0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTK; kk += 16) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm + (lane & 31)][kk + 8 * (lane >> 5)]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn + (lane & 31)][kk + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int r = m0 + wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3), c = n0 + wn + (lane & 31);
    if (r < M && c < N) C[(long)r * ldc + c] = acc[i];
  }
}

__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float silu_grad(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Z += bias (row-broadcast); hidden layers also write A = bf16(silu(Z)), the output layer Y = Z.
__global__ void k_bias_act(float* __restrict__ Z, const float* __restrict__ b, __bf16* __restrict__ A, long n,
                           int d, int act) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * d) return;
  const float z = Z[t] + b[t % d];
  Z[t] = z;
  if (act) A[t] = (__bf16)silu(z);
}

__global__ void k_to_bf16(const float* __restrict__ x, __bf16* __restrict__ y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = (__bf16)x[t];
}

// dY = 2 (Y - T) / (n d); per-workgroup partial sums of (Y - T)^2 for the loss
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, const float* __restrict__ T,
                                             float* __restrict__ dY, float* __restrict__ partial, long count,
                                             float scale) {
  __shared__ float s[256];
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (t < count) {
    const float e = Y[t] - T[t];
    dY[t] = 2.0f * e * scale;
    e2 = e * e;
  }
  s[threadIdx.x] = e2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// dZ = dA * silu'(Z) (in place on dA), and its BF16 copy for the next products
__global__ void k_silu_back(float* __restrict__ dA, const float* __restrict__ Z, __bf16* __restrict__ dZb, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float g = dA[t] * silu_grad(Z[t]);
  dA[t] = g;
  dZb[t] = (__bf16)g;
}

// db[j] = sum over rows of dZ[:, j] (one thread per column, fixed order)
__global__ void k_colsum(const float* __restrict__ dZ, float* __restrict__ db, long n, int d) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= d) return;
  float s = 0.0f;
  for (long r = 0; r < n; ++r) s += dZ[r * d + j];
  db[j] = s;
}

// Adam on the FP32 master weights; the BF16 copy follows
__global__ void k_adam(float* __restrict__ w, __bf16* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float gi = g[t];
  const float mi = b1 * m[t] + (1.0f - b1) * gi;
  const float vi = b2 * v[t] + (1.0f - b2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  const float wi = w[t] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[t] = wi;
  if (wb) wb[t] = (__bf16)wi;
}

unsigned blocks(long n, int per = 256) { return (unsigned)((n + per - 1) / per); }

}  // namespace

struct mithra_mlp {
  int members = 0, layers = 0;
  std::vector<int> dims;            // layers + 1 widths
  std::vector<long> woff, boff;     // offsets of W_l (d_l x d_{l+1}, row-major) and b_l in one member
  long per_member = 0;              // parameters per member
  hipStream_t stream = nullptr;
  float* w = nullptr;               // FP32 master parameters, members x per_member
  __bf16* wb = nullptr;             // BF16 copy (the biases' entries unused)
  float* g = nullptr;               // gradients of one member
  float* m = nullptr;               // Adam moments
  float* v = nullptr;
  long step = 0;
  // activations for a batch of up to cap rows
  long cap = 0;
  std::vector<__bf16*> X;           // layer inputs (BF16), n x d_l
  std::vector<float*> Z;            // pre-activations (FP32), n x d_{l+1}
  float* dA = nullptr;              // gradient buffer, n x max width
  float* dB = nullptr;
  __bf16* dZb = nullptr;
  float* xin = nullptr;             // FP32 staging of inputs / targets
  float* tgt = nullptr;
  float* partial = nullptr;
  int maxw = 0;
};

namespace {

void free_acts(mithra_mlp* p) {
  for (auto* x : p->X) (void)hipFree(x);
  for (auto* z : p->Z) (void)hipFree(z);
  p->X.clear();
  p->Z.clear();
  (void)hipFree(p->dA); (void)hipFree(p->dB); (void)hipFree(p->dZb);
  (void)hipFree(p->xin); (void)hipFree(p->tgt); (void)hipFree(p->partial);
  p->dA = p->dB = nullptr;
  p->dZb = nullptr;
  p->xin = p->tgt = p->partial = nullptr;
  p->cap = 0;
}

int ensure_acts(mithra_mlp* p, long n) {
  if (n <= p->cap) return MITHRA_OK;
  SG_TRY(hipStreamSynchronize(p->stream));
  free_acts(p);
  const int L = p->layers;
  p->X.assign(L, nullptr);
  p->Z.assign(L, nullptr);
  bool ok = true;
  for (int l = 0; l < L && ok; ++l) {
    ok = hipMalloc(&p->X[l], sizeof(__bf16) * n * p->dims[l]) == hipSuccess &&
         hipMalloc(&p->Z[l], sizeof(float) * n * p->dims[l + 1]) == hipSuccess;
  }
  ok = ok && hipMalloc(&p->dA, sizeof(float) * n * p->maxw) == hipSuccess &&
       hipMalloc(&p->dB, sizeof(float) * n * p->maxw) == hipSuccess &&
       hipMalloc(&p->dZb, sizeof(__bf16) * n * p->maxw) == hipSuccess &&
       hipMalloc(&p->xin, sizeof(float) * n * p->maxw) == hipSuccess &&
       hipMalloc(&p->tgt, sizeof(float) * n * p->dims[L]) == hipSuccess &&
       hipMalloc(&p->partial, sizeof(float) * (blocks(n * p->dims[L]) + 1)) == hipSuccess;
  if (!ok) {
    free_acts(p);
    return MITHRA_E_NOMEM;
  }
  p->cap = n;
  return MITHRA_OK;
}

hipError_t gemm(hipStream_t s, const __bf16* A, long sam, long sak, const __bf16* B, long sbk, long sbn, float* C,
                long ldc, int M, int N, int K) {
  hipLaunchKernelGGL(k_gemm, dim3((N + kTN - 1) / kTN, (M + kTM - 1) / kTM), dim3(256), 0, s, A, sam, sak, B, sbk,
                     sbn, C, ldc, M, N, K);
  return hipGetLastError();
}

// Forward pass of one member on X[0] (already BF16): Z[l] and X[l+1]; output in Z[L-1].
int forward(mithra_mlp* p, int member, long n) {
  hipStream_t s = p->stream;
  const float* w = p->w + (size_t)member * p->per_member;
  const __bf16* wb = p->wb + (size_t)member * p->per_member;
  for (int l = 0; l < p->layers; ++l) {
    const int di = p->dims[l], dout = p->dims[l + 1];
    SG_TRY(gemm(s, p->X[l], di, 1, wb + p->woff[l], dout, 1, p->Z[l], dout, (int)n, dout, di));
    const bool hidden = l + 1 < p->layers;
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(n * dout)), dim3(256), 0, s, p->Z[l], w + p->boff[l],
                       hidden ? p->X[l + 1] : nullptr, n, dout, hidden ? 1 : 0);
    SG_TRY(hipGetLastError());
  }
  return MITHRA_OK;
}

// Loss and gradients of one member into p->g (the forward pass must have run on the batch).
int backward(mithra_mlp* p, int member, long n, float* loss) {
  hipStream_t s = p->stream;
  const int L = p->layers, dL = p->dims[L];
  const __bf16* wb = p->wb + (size_t)member * p->per_member;
  const long count = n * dL;
  const unsigned nb = blocks(count);
  hipLaunchKernelGGL(k_mse, dim3(nb), dim3(256), 0, s, p->Z[L - 1], p->tgt, p->dA, p->partial, count,
                     1.0f / (float)count);
  SG_TRY(hipGetLastError());
  if (loss) {
    std::vector<float> part(nb);
    SG_TRY(hipMemcpyAsync(part.data(), p->partial, sizeof(float) * nb, hipMemcpyDeviceToHost, s));
    SG_TRY(hipStreamSynchronize(s));
    double sum = 0.0;
    for (float x : part) sum += x;
    *loss = (float)(sum / (double)count);
  }
  hipLaunchKernelGGL(k_to_bf16, dim3(nb), dim3(256), 0, s, p->dA, p->dZb, count);
  SG_TRY(hipGetLastError());
  float* dcur = p->dA;  // dZ of layer l (FP32), BF16 copy in dZb
  float* dnext = p->dB;
  for (int l = L - 1; l >= 0; --l) {
    const int di = p->dims[l], dout = p->dims[l + 1];
    // dW_l = X_l^T dZ_l  (di x dout)
    SG_TRY(gemm(s, p->X[l], 1, di, p->dZb, dout, 1, p->g + p->woff[l], dout, di, dout, (int)n));
    hipLaunchKernelGGL(k_colsum, dim3(blocks(dout)), dim3(256), 0, s, dcur, p->g + p->boff[l], n, dout);
    SG_TRY(hipGetLastError());
    if (l == 0) break;
    // dA_{l-1} = dZ_l W_l^T  (n x di), then dZ_{l-1} = dA_{l-1} * silu'(Z_{l-1})
    SG_TRY(gemm(s, p->dZb, dout, 1, wb + p->woff[l], 1, dout, dnext, di, (int)n, di, dout));
    hipLaunchKernelGGL(k_silu_back, dim3(blocks(n * di)), dim3(256), 0, s, dnext, p->Z[l - 1], p->dZb, n * di);
    SG_TRY(hipGetLastError());
    float* t = dcur;
    dcur = dnext;
    dnext = t;
  }
  return MITHRA_OK;
}

int load_inputs(mithra_mlp* p, const float* x, const float* t, long n) {
  if (int rc = ensure_acts(p, n)) return rc;
  hipStream_t s = p->stream;
  const long nx = n * p->dims[0];
  SG_TRY(hipMemcpyAsync(p->xin, x, sizeof(float) * nx, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(nx)), dim3(256), 0, s, p->xin, p->X[0], nx);
  SG_TRY(hipGetLastError());
  if (t) SG_TRY(hipMemcpyAsync(p->tgt, t, sizeof(float) * n * p->dims[p->layers], hipMemcpyHostToDevice, s));
  SG_TRY(hipStreamSynchronize(s));  // the host arrays may go away on return
  return MITHRA_OK;
}

}  // namespace

extern "C" {

int mithra_mlp_create(int members, int layers, const int* dims, unsigned long long seed, mithra_mlp** out) {
  if (!out || members < 1 || layers < 1 || !dims) return MITHRA_E_ARG;
  for (int l = 0; l <= layers; ++l)
    if (dims[l] < 1) return MITHRA_E_ARG;
  *out = nullptr;
  mithra_mlp* p = new (std::nothrow) mithra_mlp();
  if (!p) return MITHRA_E_NOMEM;
  p->members = members;
  p->layers = layers;
  p->dims.assign(dims, dims + layers + 1);
  long off = 0;
  for (int l = 0; l < layers; ++l) {
    p->woff.push_back(off);
    off += (long)dims[l] * dims[l + 1];
    p->boff.push_back(off);
    off += dims[l + 1];
    p->maxw = std::max(p->maxw, std::max(dims[l], dims[l + 1]));
  }
  p->per_member = off;
  const size_t total = (size_t)members * off;
  int rc = status(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  if (!rc && (hipMalloc(&p->w, sizeof(float) * total) != hipSuccess ||
              hipMalloc(&p->wb, sizeof(__bf16) * total) != hipSuccess ||
              hipMalloc(&p->g, sizeof(float) * off) != hipSuccess ||
              hipMalloc(&p->m, sizeof(float) * total) != hipSuccess ||
              hipMalloc(&p->v, sizeof(float) * total) != hipSuccess))
    rc = MITHRA_E_NOMEM;
  if (rc) {
    mithra_mlp_destroy(p);
    return rc;
  }
  // He-normal weights, zero biases; each member from its own stream of the seed
  std::vector<float> h(total, 0.0f);
  for (int e = 0; e < members; ++e) {
    std::mt19937_64 gen(seed * 1000003ull + (unsigned long long)e);
    for (int l = 0; l < layers; ++l) {
      std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (float)dims[l]));
      float* W = h.data() + (size_t)e * off + p->woff[l];
      for (long i = 0; i < (long)dims[l] * dims[l + 1]; ++i) W[i] = nd(gen);
    }
  }
  rc = mithra_mlp_set_params(p, -1, h.data());
  if (!rc) {
    std::fill(h.begin(), h.end(), 0.0f);
    rc = status(hipMemcpy(p->m, h.data(), sizeof(float) * total, hipMemcpyHostToDevice));
    if (!rc) rc = status(hipMemcpy(p->v, h.data(), sizeof(float) * total, hipMemcpyHostToDevice));
  }
  if (rc) {
    mithra_mlp_destroy(p);
    return rc;
  }
  *out = p;
  return MITHRA_OK;
}

int mithra_mlp_destroy(mithra_mlp* p) {
  if (!p) return MITHRA_E_ARG;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  free_acts(p);
  (void)hipFree(p->w); (void)hipFree(p->wb); (void)hipFree(p->g); (void)hipFree(p->m); (void)hipFree(p->v);
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
  return MITHRA_OK;
}

long mithra_mlp_param_count(const mithra_mlp* p) { return p ? p->per_member : -1; }

// member = -1: all members from params (members x per-member); else one member.
int mithra_mlp_set_params(mithra_mlp* p, int member, const float* params) {
  if (!p || !params || member < -1 || member >= p->members) return MITHRA_E_ARG;
  const size_t n = (size_t)(member < 0 ? p->members : 1) * p->per_member;
  float* w = p->w + (member < 0 ? 0 : (size_t)member * p->per_member);
  __bf16* wb = p->wb + (member < 0 ? 0 : (size_t)member * p->per_member);
  SG_TRY(hipMemcpyAsync(w, params, sizeof(float) * n, hipMemcpyHostToDevice, p->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks((long)n)), dim3(256), 0, p->stream, w, wb, (long)n);
  SG_TRY(hipGetLastError());
  SG_TRY(hipStreamSynchronize(p->stream));
  return MITHRA_OK;
}

int mithra_mlp_get_params(mithra_mlp* p, int member, float* params) {
  if (!p || !params || member < -1 || member >= p->members) return MITHRA_E_ARG;
  const size_t n = (size_t)(member < 0 ? p->members : 1) * p->per_member;
  SG_TRY(hipStreamSynchronize(p->stream));
  SG_TRY(hipMemcpy(params, p->w + (member < 0 ? 0 : (size_t)member * p->per_member), sizeof(float) * n,
                   hipMemcpyDeviceToHost));
  return MITHRA_OK;
}

// y: members x n x d_out, each member's prediction for the n rows of x (n x d_in).
int mithra_mlp_forward(mithra_mlp* p, const float* x, long n, float* y) {
  if (!p || !x || !y || n < 1) return MITHRA_E_ARG;
  if (int rc = load_inputs(p, x, nullptr, n)) return rc;
  const int dL = p->dims[p->layers];
  for (int e = 0; e < p->members; ++e) {
    if (int rc = forward(p, e, n)) return rc;
    SG_TRY(hipMemcpyAsync(y + (size_t)e * n * dL, p->Z[p->layers - 1], sizeof(float) * n * dL,
                          hipMemcpyDeviceToHost, p->stream));
  }
  SG_TRY(hipStreamSynchronize(p->stream));
  return MITHRA_OK;
}

// Loss (mean squared error over the n x d_out targets) and its gradient for one member, in the
// parameter layout of mithra_mlp_get_params.
int mithra_mlp_gradients(mithra_mlp* p, int member, const float* x, const float* t, long n, float* grads,
                         float* loss) {
  if (!p || !x || !t || n < 1 || member < 0 || member >= p->members) return MITHRA_E_ARG;
  if (int rc = load_inputs(p, x, t, n)) return rc;
  if (int rc = forward(p, member, n)) return rc;
  if (int rc = backward(p, member, n, loss)) return rc;
  if (grads) {
    SG_TRY(hipMemcpyAsync(grads, p->g, sizeof(float) * p->per_member, hipMemcpyDeviceToHost, p->stream));
    SG_TRY(hipStreamSynchronize(p->stream));
  }
  return MITHRA_OK;
}

// One Adam step of every member on the same mini-batch; loss[e] (optional) before the step.
int mithra_mlp_train_step(mithra_mlp* p, const float* x, const float* t, long n, float lr, float* loss) {
  if (!p || !x || !t || n < 1 || !(lr > 0.0f)) return MITHRA_E_ARG;
  if (int rc = load_inputs(p, x, t, n)) return rc;
  ++p->step;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - std::pow(b1, (float)p->step), c2 = 1.0f - std::pow(b2, (float)p->step);
  for (int e = 0; e < p->members; ++e) {
    if (int rc = forward(p, e, n)) return rc;
    if (int rc = backward(p, e, n, loss ? loss + e : nullptr)) return rc;
    const size_t o = (size_t)e * p->per_member;
    hipLaunchKernelGGL(k_adam, dim3(blocks(p->per_member)), dim3(256), 0, p->stream, p->w + o, p->wb + o, p->g,
                       p->m + o, p->v + o, p->per_member, lr, b1, b2, eps, c1, c2);
    SG_TRY(hipGetLastError());
  }
  SG_TRY(hipStreamSynchronize(p->stream));
  return MITHRA_OK;
}

// One product of the surrogate's GEMM on host arrays (BF16 given as raw 16-bit patterns), for
// tests and timing: C = op(A) op(B) with the strides of k_gemm.
int mithra_mlp_gemm(const unsigned short* A, long sam, long sak, long a_len, const unsigned short* B, long sbk,
                    long sbn, long b_len, float* C, int M, int N, int K, int reps, float* ms) {
  if (!A || !B || !C || M < 1 || N < 1 || K < 1 || reps < 1) return MITHRA_E_ARG;
  __bf16 *dA = nullptr, *dB = nullptr;
  float* dC = nullptr;
  int rc = MITHRA_OK;
  if (hipMalloc(&dA, 2 * a_len) != hipSuccess || hipMalloc(&dB, 2 * b_len) != hipSuccess ||
      hipMalloc(&dC, sizeof(float) * (size_t)M * N) != hipSuccess)
    rc = MITHRA_E_NOMEM;
  if (!rc) rc = status(hipMemcpy(dA, A, 2 * a_len, hipMemcpyHostToDevice));
  if (!rc) rc = status(hipMemcpy(dB, B, 2 * b_len, hipMemcpyHostToDevice));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) rc = status(hipEventCreate(&e0));
  if (!rc) rc = status(hipEventCreate(&e1));
  if (!rc) rc = status(gemm(0, dA, sam, sak, dB, sbk, sbn, dC, N, M, N, K));  // warm-up
  if (!rc) rc = status(hipEventRecord(e0, 0));
  for (int r = 0; r < reps && !rc; ++r) rc = status(gemm(0, dA, sam, sak, dB, sbk, sbn, dC, N, M, N, K));
  if (!rc) rc = status(hipEventRecord(e1, 0));
  if (!rc) rc = status(hipEventSynchronize(e1));
  float t = 0.0f;
  if (!rc) rc = status(hipEventElapsedTime(&t, e0, e1));
  if (!rc && ms) *ms = t / reps;
  if (!rc) rc = status(hipMemcpy(C, dC, sizeof(float) * (size_t)M * N, hipMemcpyDeviceToHost));
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  return rc;
}

}  // extern "C"
