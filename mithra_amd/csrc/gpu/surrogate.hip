// Learned surrogate of a scan point: an ensemble of multilayer perceptrons (SiLU hidden layers,
// linear output) trained and evaluated on the GPU.  Weights are kept in FP32 (the Adam master
// copy) and in BF16 (what the products read); every product is a BF16 x BF16 -> FP32 MFMA GEMM
// (v_mfma_f32_32x32x16_bf16) with FP32 accumulation.  Activations enter the products as BF16;
// pre-activations, losses and gradients stay FP32 until they feed the next product.
//
// The members of the ensemble are marched together: every kernel takes the member from
// blockIdx.z (GEMMs) or blockIdx.y (elementwise), so one launch per layer and pass serves the
// whole ensemble.  Training on a resident data set (mithra_mlp_set_data / _train_epochs) runs
// whole epochs as one captured HIP graph: the mini-batch gather, the forward and backward
// passes, the loss and the Adam update of every member, with the epoch number, learning rate
// and Adam bias corrections read on the device.
//
// GEMM: C[M x N] (FP32, row-major, ldc) = op(A)[M x K] * op(B)[K x N], where op(A)(m, k) is read
// at A[m * sam + k * sak] and op(B)(k, n) at B[k * sbk + n * sbn] (BF16): one kernel serves the
// forward product and both backward products (transposed operands) without copies.  Four
// wavefronts of (32 WM) x (32 WN) make a (64 WM) x (64 WN) workgroup tile; K is staged through
// LDS 32 or 64 at a time, double-buffered: the next stage's 16-byte global loads are in flight
// while the current stage feeds the matrix cores.  Operands contiguous along k are written to
// LDS as they come; operands contiguous along m (or n) are transposed on the way in.
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

constexpr int kPadK = 8;  // LDS row padding (bf16): rows stay 16-byte aligned

struct GemmArgs {
  const __bf16* A;
  long sam, sak, sae;  // op(A)(m, k) of member e at A[e sae + m sam + k sak]
  const __bf16* B;
  long sbk, sbn, sbe;  // op(B)(k, n) of member e at B[e sbe + k sbk + n sbn]
  float* C;
  long ldc, sce;       // C(m, n) of member e at C[e sce + m ldc + n]
  int M, N, K;
  int veca, vecb;      // 16-byte loads allowed (aligned bases and strides along the vector)
};

// One 8-element vector of an operand tile: elements (r0 + i dr, c0 + i dc), i < 8, of a matrix
// whose element (r, c) is at P[r sr + c sc]; rows < R and columns < Cn are in range.  `vec`:
// the 8 elements are contiguous and 16-byte aligned whenever they are all in range.
__device__ inline bf16x8 load8(const __bf16* __restrict__ P, long sr, long sc, int r0, int c0, int dr, int dc, int R,
                               int Cn, bool vec) {
  const bool full = r0 + 7 * dr < R && c0 + 7 * dc < Cn && r0 < R && c0 < Cn;
  if (vec && full) return *reinterpret_cast<const bf16x8*>(P + r0 * sr + c0 * sc);
  bf16x8 v;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const int r = r0 + i * dr, c = c0 + i * dc;
    v[i] = (r < R && c < Cn) ? P[r * sr + c * sc] : (__bf16)0.0f;
  }
  return v;
}

// Staging of one operand tile of ROWS x TK (rows m of op(A), or n of op(B); k contiguous in LDS)
// by 256 threads.  KC: the operand is contiguous along k, each thread moves 8-element vectors
// straight into LDS rows.  Otherwise it is contiguous along the rows: a thread loads the same 8
// rows at two neighbouring k and writes each row's pair as one 32-bit LDS word (the transpose).
template <int ROWS, int TK, bool KC>
struct Stage {
  static constexpr int kChunks = KC ? ROWS * TK / 8 : (TK / 2) * (ROWS / 8);
  static constexpr int kPasses = kChunks / 256;
  static_assert(kPasses >= 1 && kChunks % 256 == 0, "tile not covered by whole passes");
  bf16x8 r[kPasses][KC ? 1 : 2];
  // P[row sr + k sc] with sr/sc the strides of (row, k); R rows, K columns in range
  __device__ inline void load(const __bf16* __restrict__ P, long sr, long sk, int row0, int k0, int R, int K, bool vec) {
#pragma unroll
    for (int p = 0; p < kPasses; ++p) {
      const int c = threadIdx.x + 256 * p;
      if (KC) {
        r[p][0] = load8(P, sr, sk, row0 + c / (TK / 8), k0 + (c % (TK / 8)) * 8, 0, 1, R, K, vec);
      } else {
        const int row = row0 + (c % (ROWS / 8)) * 8, k = k0 + 2 * (c / (ROWS / 8));
        r[p][0] = load8(P, sr, sk, row, k, 1, 0, R, K, vec);
        r[p][1] = load8(P, sr, sk, row, k + 1, 1, 0, R, K, vec);
      }
    }
  }
  __device__ inline void store(__bf16 (*T)[TK + kPadK]) const {
#pragma unroll
    for (int p = 0; p < kPasses; ++p) {
      const int c = threadIdx.x + 256 * p;
      if (KC) {
        *reinterpret_cast<bf16x8*>(&T[c / (TK / 8)][(c % (TK / 8)) * 8]) = r[p][0];
      } else {
        const int row = (c % (ROWS / 8)) * 8, k = 2 * (c / (ROWS / 8));
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          typedef __attribute__((__vector_size__(2 * sizeof(__bf16)))) __bf16 bf16x2;
          const bf16x2 w = {r[p][0][i], r[p][1][i]};
          *reinterpret_cast<bf16x2*>(&T[row + i][k]) = w;
        }
      }
    }
  }
};

// Lane layout of v_mfma_f32_32x32x16_bf16 (64 lanes, l = lane): A supplies row l % 32 and
// k = 8 (l / 32) .. +7, B column l % 32 and the same k; accumulator i of lane l is row
// 8 (i / 4) + 4 (l / 32) + i % 4, column l % 32.
// AK: op(A) contiguous along k (sak = 1), else along m (sam = 1); BK: op(B) along k (sbk = 1),
// else along n (sbn = 1).  With general strides the vector flags are off and any form is exact.
template <int WM, int WN, int TK, bool AK, bool BK>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
  constexpr int BM = 64 * WM, BN = 64 * WN, LD = TK + kPadK;
  __shared__ __attribute__((aligned(16))) __bf16 As[2][BM][LD];  // rows m, k contiguous
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][BN][LD];  // rows n, k contiguous (B transposed)
  const int e = blockIdx.z;
  const __bf16* __restrict__ A = g.A + e * g.sae;
  const __bf16* __restrict__ B = g.B + e * g.sbe;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = (w >> 1) * 32 * WM, wn = (w & 1) * 32 * WN;
  const bool va = g.veca, vb = g.vecb;
  Stage<BM, TK, AK> sa;
  Stage<BN, TK, BK> sb;
  // op(B)(k, n) = B[k sbk + n sbn] is staged as rows n: the matrix (n, k) with strides (sbn, sbk)
  auto load = [&](int k0) {
    sa.load(A, g.sam, g.sak, m0, k0, g.M, g.K, va);
    sb.load(B, g.sbn, g.sbk, n0, k0, g.N, g.K, vb);
  };
  floatx16 acc[WM][WN];
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = floatx16{};
  load(0);
  sa.store(As[0]);
  sb.store(Bs[0]);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < g.K; k0 += TK) {
    const bool next = k0 + TK < g.K;
    if (next) load(k0 + TK);  // in flight while this stage computes
#pragma unroll
    for (int kk = 0; kk < TK; kk += 16) {
      bf16x8 a[WM], b[WN];
#pragma unroll
      for (int i = 0; i < WM; ++i)
        a[i] = *reinterpret_cast<const bf16x8*>(&As[buf][wm + 32 * i + (lane & 31)][kk + 8 * (lane >> 5)]);
#pragma unroll
      for (int j = 0; j < WN; ++j)
        b[j] = *reinterpret_cast<const bf16x8*>(&Bs[buf][wn + 32 * j + (lane & 31)][kk + 8 * (lane >> 5)]);
#pragma unroll
      for (int i = 0; i < WM; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    // the other buffer was last read in the previous stage, which the barrier below it closed
    if (next) {
      sa.store(As[buf ^ 1]);
      sb.store(Bs[buf ^ 1]);
    }
    __syncthreads();
    buf ^= 1;
  }
  float* __restrict__ C = g.C + e * g.sce;
#pragma unroll
  for (int i = 0; i < WM; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int r = m0 + wm + 32 * i + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        const int c = n0 + wn + 32 * j + (lane & 31);
        if (r < g.M && c < g.N) C[(long)r * g.ldc + c] = acc[i][j][v];
      }
}

__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float silu_grad(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

// Member blockIdx.y: Z += bias (row-broadcast); hidden layers also write A = bf16(silu(Z)).
__global__ void k_bias_act(float* __restrict__ Z, long sz, const float* __restrict__ b, long sb,
                           __bf16* __restrict__ A, long sa, long n, int d, int act) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * d) return;
  const int e = blockIdx.y;
  const float z = Z[e * sz + t] + b[e * sb + t % d];
  Z[e * sz + t] = z;
  if (act) A[e * sa + t] = (__bf16)silu(z);
}

__global__ void k_to_bf16(const float* __restrict__ x, __bf16* __restrict__ y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = (__bf16)x[t];
}

// Member blockIdx.y: dY = 2 (Y - T) / (n d) and its BF16 copy; per-workgroup partial sums of
// (Y - T)^2 for the loss (partial[e nb + blockIdx.x]).
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ Y, long sy, const float* __restrict__ T,
                                             float* __restrict__ dY, __bf16* __restrict__ dYb, long sd,
                                             float* __restrict__ partial, long count, float scale) {
  __shared__ float s[256];
  const int e = blockIdx.y;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0.0f;
  if (t < count) {
    const float d = Y[e * sy + t] - T[t];
    const float gr = 2.0f * d * scale;
    dY[e * sd + t] = gr;
    dYb[e * sd + t] = (__bf16)gr;
    e2 = d * d;
  }
  s[threadIdx.x] = e2;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] += s[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[(long)e * gridDim.x + blockIdx.x] = s[0];
}

// Member blockIdx.y: dZ = dA * silu'(Z) (in place on dA), and its BF16 copy
__global__ void k_silu_back(float* __restrict__ dA, long sda, const float* __restrict__ Z, long sz,
                            __bf16* __restrict__ dZb, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int e = blockIdx.y;
  const float g = dA[e * sda + t] * silu_grad(Z[e * sz + t]);
  dA[e * sda + t] = g;
  dZb[e * sda + t] = (__bf16)g;
}

// Member blockIdx.y: db[j] = sum over rows of dZ[:, j].  A workgroup takes 64 columns; its 4
// wavefronts each sum every fourth row (coalesced 256-byte rows), and the four partial sums are
// added in wavefront order: a fixed order, so the result does not vary from run to run.
__global__ __launch_bounds__(256) void k_colsum(const float* __restrict__ dZ, long sdz, float* __restrict__ db,
                                                long sdb, long n, int d) {
  __shared__ float part[4][64];
  const int e = blockIdx.y, q = threadIdx.x >> 6, j = blockIdx.x * 64 + (threadIdx.x & 63);
  float s = 0.0f;
  if (j < d)
    for (long r = q; r < n; r += 4) s += dZ[e * sdz + r * d + j];
  part[q][threadIdx.x & 63] = s;
  __syncthreads();
  if (q == 0 && j < d) db[e * sdb + j] = ((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) +
                                         part[3][threadIdx.x];
}

// The schedule of a captured epoch: the epoch number lives on the device (*epoch), so one graph
// serves every epoch.  Step t = step0 + epoch steps + b + 1 (Adam's count, from 1); learning rate
// max(lr_max (1 + cos(pi epoch / epochs)) / 2, lr_min).
struct EpochSched {
  const int* epoch;
  long step0;
  int steps, b, epochs;
  float lr_max, lr_min;
};

constexpr float kB1 = 0.9f, kB2 = 0.999f, kEps = 1e-8f;

// Adam on the FP32 master weights (all members); the BF16 copy follows.  With s.epoch set, the
// rate and bias corrections come from the schedule, else from lr, c1, c2.
__global__ void k_adam(float* __restrict__ w, __bf16* __restrict__ wb, const float* __restrict__ g,
                       float* __restrict__ m, float* __restrict__ v, long n, float lr, float c1, float c2,
                       EpochSched s) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (s.epoch) {
    const int ep = *s.epoch;
    const float step = (float)(s.step0 + (long)ep * s.steps + s.b + 1);
    c1 = 1.0f - powf(kB1, step);
    c2 = 1.0f - powf(kB2, step);
    lr = fmaxf(s.lr_max * 0.5f * (1.0f + cosf(3.14159265358979f * (float)ep / (float)s.epochs)), s.lr_min);
  }
  const float gi = g[t];
  const float mi = kB1 * m[t] + (1.0f - kB1) * gi;
  const float vi = kB2 * v[t] + (1.0f - kB2) * gi * gi;
  m[t] = mi;
  v[t] = vi;
  const float wi = w[t] - lr * (mi / c1) / (sqrtf(vi / c2) + kEps);
  w[t] = wi;
  wb[t] = (__bf16)wi;
}

// Mini-batch b of the epoch *epoch: rows perm[epoch n_data + b batch + r] of the resident data
// set, inputs (BF16) and targets (FP32).
__global__ void k_gather(const __bf16* __restrict__ dx, const float* __restrict__ dt, const int* __restrict__ perm,
                         const int* __restrict__ epoch, long n_data, int b, int batch, int d0, int dL,
                         __bf16* __restrict__ xb, float* __restrict__ tb) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = blockIdx.y;
  if (r >= batch) return;
  const long src = perm[(long)*epoch * n_data + (long)b * batch + r];
  if (t < d0) xb[(long)r * d0 + t] = dx[src * d0 + t];
  if (t < dL) tb[(long)r * dL + t] = dt[src * dL + t];
}

// The loss of a mini-batch, per member, added to the epoch's entry (fixed order).
__global__ void k_loss_add(const float* __restrict__ partial, int nb, const int* __restrict__ epoch,
                           float* __restrict__ out, int members, float scale) {
  const int e = threadIdx.x;
  if (e >= members) return;
  float s = 0.0f;
  for (int i = 0; i < nb; ++i) s += partial[(long)e * nb + i];
  out[(long)*epoch * members + e] += s * scale;
}

__global__ void k_next_epoch(int* epoch) {
  if (threadIdx.x == 0) *epoch += 1;
}

__global__ void k_fill(float* __restrict__ p, long n, float v) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = v;
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
  float* g = nullptr;               // gradients, members x per_member
  float* m = nullptr;               // Adam moments
  float* v = nullptr;
  long step = 0;                    // Adam steps taken
  // activations for a batch of up to cap rows; every buffer but X[0] holds one slab per member
  long cap = 0;
  std::vector<__bf16*> X;           // layer inputs (BF16): X[0] n x d_0 (shared), X[l] members x n x d_l
  std::vector<float*> Z;            // pre-activations (FP32), members x n x d_{l+1}
  float* dA = nullptr;              // gradient buffers, members x n x max width
  float* dB = nullptr;
  __bf16* dZb = nullptr;
  float* xin = nullptr;             // FP32 staging of the inputs, n x d_0
  float* tgt = nullptr;             // targets, n x d_L (shared)
  float* partial = nullptr;         // loss partial sums, members x blocks
  int maxw = 0;
  // resident training data
  long n_data = 0;
  __bf16* data_x = nullptr;         // n_data x d_0
  float* data_t = nullptr;          // n_data x d_L
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
  const long E = p->members;
  p->X.assign(L, nullptr);
  p->Z.assign(L, nullptr);
  bool ok = true;
  for (int l = 0; l < L && ok; ++l) {
    ok = hipMalloc(&p->X[l], sizeof(__bf16) * (l ? E : 1) * n * p->dims[l]) == hipSuccess &&
         hipMalloc(&p->Z[l], sizeof(float) * E * n * p->dims[l + 1]) == hipSuccess;
  }
  ok = ok && hipMalloc(&p->dA, sizeof(float) * E * n * p->maxw) == hipSuccess &&
       hipMalloc(&p->dB, sizeof(float) * E * n * p->maxw) == hipSuccess &&
       hipMalloc(&p->dZb, sizeof(__bf16) * E * n * p->maxw) == hipSuccess &&
       hipMalloc(&p->xin, sizeof(float) * n * p->dims[0]) == hipSuccess &&
       hipMalloc(&p->tgt, sizeof(float) * n * p->dims[L]) == hipSuccess &&
       hipMalloc(&p->partial, sizeof(float) * E * (blocks(n * p->dims[L]) + 1)) == hipSuccess;
  if (!ok) {
    free_acts(p);
    return MITHRA_E_NOMEM;
  }
  p->cap = n;
  return MITHRA_OK;
}

bool aligned16(const void* q) { return ((uintptr_t)q & 15) == 0; }

// tile: 1 = 64 x 64 workgroups, 2 = 128 x 128, 0 = the larger one where it still makes a
// workgroup per compute unit.  The vector flags need each 8-element run along the contiguous
// direction to start 16-byte aligned, and the transposing stage loads runs along m (or n).
hipError_t gemm(hipStream_t s, GemmArgs g, int members, int tile = 0) {
  const bool ak = g.sak == 1 || g.sam != 1, bk = g.sbk == 1 || g.sbn != 1;
  // a vector runs along the contiguous direction: rows (or columns) must start 16-byte aligned
  g.veca = aligned16(g.A) && g.sae % 8 == 0 &&
           (ak ? g.sak == 1 && g.sam % 8 == 0 : g.sam == 1 && g.sak % 8 == 0);
  g.vecb = aligned16(g.B) && g.sbe % 8 == 0 &&
           (bk ? g.sbk == 1 && g.sbn % 8 == 0 : g.sbn == 1 && g.sbk % 8 == 0);
  if (tile == 0) {
    const long big = (long)((g.M + 127) / 128) * ((g.N + 127) / 128) * members;
    tile = big >= 256 ? 2 : 1;
  }
  const int T = tile == 2 ? 128 : 64;
  const dim3 grid((unsigned)((g.N + T - 1) / T), (unsigned)((g.M + T - 1) / T), (unsigned)members);
#define SG_GEMM(WM, TK, AK, BK) hipLaunchKernelGGL((k_gemm<WM, WM, TK, AK, BK>), grid, dim3(256), 0, s, g)
  if (tile == 2) {  // 128 x 128 workgroups, k staged 32 at a time
    if (ak && bk) SG_GEMM(2, 32, true, true);
    else if (ak) SG_GEMM(2, 32, true, false);
    else if (bk) SG_GEMM(2, 32, false, true);
    else SG_GEMM(2, 32, false, false);
  } else {  // 64 x 64 workgroups, k staged 64 at a time (half the barriers on short products)
    if (ak && bk) SG_GEMM(1, 64, true, true);
    else if (ak) SG_GEMM(1, 64, true, false);
    else if (bk) SG_GEMM(1, 64, false, true);
    else SG_GEMM(1, 64, false, false);
  }
#undef SG_GEMM
  return hipGetLastError();
}

// Forward pass of members e0 .. e0 + ne - 1 on X[0] (already BF16, n rows): Z[l] and X[l+1] in
// the members' slabs; outputs in Z[L-1].
int forward(mithra_mlp* p, int e0, int ne, long n) {
  hipStream_t s = p->stream;
  const long pm = p->per_member;
  for (int l = 0; l < p->layers; ++l) {
    const int di = p->dims[l], dout = p->dims[l + 1];
    GemmArgs g{};
    g.A = p->X[l]; g.sam = di; g.sak = 1; g.sae = l ? n * di : 0;
    g.B = p->wb + e0 * pm + p->woff[l]; g.sbk = dout; g.sbn = 1; g.sbe = pm;
    g.C = p->Z[l]; g.ldc = dout; g.sce = n * dout;
    g.M = (int)n; g.N = dout; g.K = di;
    SG_TRY(gemm(s, g, ne));
    const bool hidden = l + 1 < p->layers;
    hipLaunchKernelGGL(k_bias_act, dim3(blocks(n * dout), ne), dim3(256), 0, s, p->Z[l], n * dout,
                       p->w + e0 * pm + p->boff[l], pm, hidden ? p->X[l + 1] : nullptr, n * dout, n, dout,
                       hidden ? 1 : 0);
    SG_TRY(hipGetLastError());
  }
  return MITHRA_OK;
}

// Loss partial sums and gradients of members e0 .. e0 + ne - 1 into their slots of p->g (the
// forward pass must have run on the batch); the per-member loss to the host when `loss` is set.
int backward(mithra_mlp* p, int e0, int ne, long n, float* loss) {
  hipStream_t s = p->stream;
  const int L = p->layers, dL = p->dims[L];
  const long pm = p->per_member, slab = n * p->maxw;
  const long count = n * dL;
  const unsigned nb = blocks(count);
  hipLaunchKernelGGL(k_mse, dim3(nb, ne), dim3(256), 0, s, p->Z[L - 1], count, p->tgt, p->dA, p->dZb, slab,
                     p->partial, count, 1.0f / (float)count);
  SG_TRY(hipGetLastError());
  if (loss) {
    std::vector<float> part((size_t)ne * nb);
    SG_TRY(hipMemcpyAsync(part.data(), p->partial, sizeof(float) * part.size(), hipMemcpyDeviceToHost, s));
    SG_TRY(hipStreamSynchronize(s));
    for (int e = 0; e < ne; ++e) {
      double sum = 0.0;
      for (unsigned i = 0; i < nb; ++i) sum += part[(size_t)e * nb + i];
      loss[e] = (float)(sum / (double)count);
    }
  }
  float* dcur = p->dA;  // dZ of layer l (FP32), BF16 copy in dZb
  float* dnext = p->dB;
  for (int l = L - 1; l >= 0; --l) {
    const int di = p->dims[l], dout = p->dims[l + 1];
    // dW_l = X_l^T dZ_l  (di x dout)
    GemmArgs g{};
    g.A = p->X[l]; g.sam = 1; g.sak = di; g.sae = l ? n * di : 0;
    g.B = p->dZb; g.sbk = dout; g.sbn = 1; g.sbe = slab;
    g.C = p->g + e0 * pm + p->woff[l]; g.ldc = dout; g.sce = pm;
    g.M = di; g.N = dout; g.K = (int)n;
    SG_TRY(gemm(s, g, ne));
    hipLaunchKernelGGL(k_colsum, dim3(blocks(dout, 64), ne), dim3(256), 0, s, dcur, slab,
                       p->g + e0 * pm + p->boff[l], pm, n, dout);
    SG_TRY(hipGetLastError());
    if (l == 0) break;
    // dA_{l-1} = dZ_l W_l^T  (n x di), then dZ_{l-1} = dA_{l-1} * silu'(Z_{l-1})
    GemmArgs h{};
    h.A = p->dZb; h.sam = dout; h.sak = 1; h.sae = slab;
    h.B = p->wb + e0 * pm + p->woff[l]; h.sbk = 1; h.sbn = dout; h.sbe = pm;
    h.C = dnext; h.ldc = di; h.sce = slab;
    h.M = (int)n; h.N = di; h.K = dout;
    SG_TRY(gemm(s, h, ne));
    hipLaunchKernelGGL(k_silu_back, dim3(blocks(n * di), ne), dim3(256), 0, s, dnext, slab, p->Z[l - 1], n * di,
                       p->dZb, n * di);
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

hipError_t adam(mithra_mlp* p, float lr, float c1, float c2, EpochSched sched) {
  const long n = (long)p->members * p->per_member;
  hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, p->stream, p->w, p->wb, p->g, p->m, p->v, n, lr, c1, c2,
                     sched);
  return hipGetLastError();
}

// One epoch of `steps` mini-batches of the resident data, every member, as stream work (for
// capture): gather, forward, backward, loss, Adam; then the epoch counter advances.
int enqueue_epoch(mithra_mlp* p, const int* perm, int* epoch, float* epoch_loss, int steps, int batch, int epochs,
                  float lr_max, float lr_min) {
  hipStream_t s = p->stream;
  const int d0 = p->dims[0], dL = p->dims[p->layers];
  const unsigned nb = blocks((long)batch * dL);
  for (int b = 0; b < steps; ++b) {
    hipLaunchKernelGGL(k_gather, dim3(blocks(std::max(d0, dL)), batch), dim3(256), 0, s, p->data_x, p->data_t, perm,
                       epoch, p->n_data, b, batch, d0, dL, p->X[0], p->tgt);
    SG_TRY(hipGetLastError());
    if (int rc = forward(p, 0, p->members, batch)) return rc;
    if (int rc = backward(p, 0, p->members, batch, nullptr)) return rc;
    hipLaunchKernelGGL(k_loss_add, dim3(1), dim3(64), 0, s, p->partial, (int)nb, epoch, epoch_loss, p->members,
                       1.0f / ((float)batch * dL * steps * p->members));
    SG_TRY(hipGetLastError());
    EpochSched sched{epoch, p->step, steps, b, epochs, lr_max, lr_min};
    SG_TRY(adam(p, 0.0f, 1.0f, 1.0f, sched));
  }
  hipLaunchKernelGGL(k_next_epoch, dim3(1), dim3(64), 0, s, epoch);
  SG_TRY(hipGetLastError());
  return MITHRA_OK;
}

}  // namespace

extern "C" {

int mithra_mlp_create(int members, int layers, const int* dims, unsigned long long seed, mithra_mlp** out) {
  if (!out || members < 1 || members > 64 || layers < 1 || !dims) return MITHRA_E_ARG;
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
              hipMalloc(&p->g, sizeof(float) * total) != hipSuccess ||
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
  (void)hipFree(p->data_x); (void)hipFree(p->data_t);
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
  if (int rc = forward(p, 0, p->members, n)) return rc;
  SG_TRY(hipMemcpyAsync(y, p->Z[p->layers - 1], sizeof(float) * p->members * n * dL, hipMemcpyDeviceToHost,
                        p->stream));
  SG_TRY(hipStreamSynchronize(p->stream));
  return MITHRA_OK;
}

// Loss (mean squared error over the n x d_out targets) and its gradient for one member, in the
// parameter layout of mithra_mlp_get_params.
int mithra_mlp_gradients(mithra_mlp* p, int member, const float* x, const float* t, long n, float* grads,
                         float* loss) {
  if (!p || !x || !t || n < 1 || member < 0 || member >= p->members) return MITHRA_E_ARG;
  if (int rc = load_inputs(p, x, t, n)) return rc;
  if (int rc = forward(p, member, 1, n)) return rc;
  if (int rc = backward(p, member, 1, n, loss)) return rc;
  if (grads) {
    SG_TRY(hipMemcpyAsync(grads, p->g + (size_t)member * p->per_member, sizeof(float) * p->per_member,
                          hipMemcpyDeviceToHost, p->stream));
    SG_TRY(hipStreamSynchronize(p->stream));
  }
  return MITHRA_OK;
}

// One Adam step of every member on the same mini-batch; loss[e] (optional) before the step.
int mithra_mlp_train_step(mithra_mlp* p, const float* x, const float* t, long n, float lr, float* loss) {
  if (!p || !x || !t || n < 1 || !(lr > 0.0f)) return MITHRA_E_ARG;
  if (int rc = load_inputs(p, x, t, n)) return rc;
  ++p->step;
  const float c1 = 1.0f - std::pow(kB1, (float)p->step), c2 = 1.0f - std::pow(kB2, (float)p->step);
  if (int rc = forward(p, 0, p->members, n)) return rc;
  if (int rc = backward(p, 0, p->members, n, loss)) return rc;
  SG_TRY(adam(p, lr, c1, c2, EpochSched{}));
  SG_TRY(hipStreamSynchronize(p->stream));
  return MITHRA_OK;
}

// The training set, resident on the device: n rows of inputs x (n x d_0, rounded to BF16) and
// targets t (n x d_out), replacing any set before.
int mithra_mlp_set_data(mithra_mlp* p, const float* x, const float* t, long n) {
  if (!p || !x || !t || n < 1) return MITHRA_E_ARG;
  SG_TRY(hipStreamSynchronize(p->stream));
  (void)hipFree(p->data_x);
  (void)hipFree(p->data_t);
  p->data_x = nullptr;
  p->data_t = nullptr;
  p->n_data = 0;
  const long nx = n * p->dims[0], nt = n * p->dims[p->layers];
  float* stage = nullptr;
  if (hipMalloc(&p->data_x, sizeof(__bf16) * nx) != hipSuccess ||
      hipMalloc(&p->data_t, sizeof(float) * nt) != hipSuccess || hipMalloc(&stage, sizeof(float) * nx) != hipSuccess) {
    (void)hipFree(stage);
    (void)hipFree(p->data_x);
    (void)hipFree(p->data_t);
    p->data_x = nullptr;
    p->data_t = nullptr;
    return MITHRA_E_NOMEM;
  }
  hipError_t e = hipMemcpyAsync(stage, x, sizeof(float) * nx, hipMemcpyHostToDevice, p->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_to_bf16, dim3(blocks(nx)), dim3(256), 0, p->stream, stage, p->data_x, nx);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(p->data_t, t, sizeof(float) * nt, hipMemcpyHostToDevice, p->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(p->stream);
  (void)hipFree(stage);
  SG_TRY(e);
  p->n_data = n;
  return MITHRA_OK;
}

// `epochs` epochs of Adam on the resident training set, every member on the same mini-batches of
// `batch` rows (n / batch per epoch, each epoch a fresh permutation drawn from `seed`), learning
// rate max(lr_max (1 + cos(pi epoch / epochs)) / 2, lr_min).  One epoch is one captured HIP graph,
// replayed `epochs` times without the host.  loss (optional, epochs entries): each epoch's mean
// training loss over the members and mini-batches, before each step.
int mithra_mlp_train_epochs(mithra_mlp* p, int epochs, int batch, float lr_max, float lr_min,
                            unsigned long long seed, float* loss) {
  if (!p || epochs < 1 || batch < 1 || !(lr_max > 0.0f) || !(lr_min >= 0.0f)) return MITHRA_E_ARG;
  if (!p->data_x || batch > p->n_data) return MITHRA_E_STATE;
  const int steps = (int)(p->n_data / batch);
  const long nperm = (long)epochs * p->n_data;
  if (int rc = ensure_acts(p, batch)) return rc;
  // the permutations of every epoch (Fisher-Yates on one generator)
  std::vector<int> perm((size_t)nperm);
  std::mt19937_64 gen(seed);
  for (int ep = 0; ep < epochs; ++ep) {
    int* q = perm.data() + (size_t)ep * p->n_data;
    for (long i = 0; i < p->n_data; ++i) q[i] = (int)i;
    for (long i = p->n_data - 1; i > 0; --i) {
      std::uniform_int_distribution<long> u(0, i);
      std::swap(q[i], q[u(gen)]);
    }
  }
  int* dperm = nullptr;
  int* depoch = nullptr;
  float* dloss = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int rc = MITHRA_OK;
  if (hipMalloc(&dperm, sizeof(int) * nperm) != hipSuccess || hipMalloc(&depoch, sizeof(int)) != hipSuccess ||
      hipMalloc(&dloss, sizeof(float) * (size_t)epochs * p->members) != hipSuccess)
    rc = MITHRA_E_NOMEM;
  const int zero = 0;
  hipStream_t s = p->stream;
  if (!rc) rc = status(hipMemcpyAsync(dperm, perm.data(), sizeof(int) * nperm, hipMemcpyHostToDevice, s));
  if (!rc) rc = status(hipMemcpyAsync(depoch, &zero, sizeof(int), hipMemcpyHostToDevice, s));
  if (!rc) {
    hipLaunchKernelGGL(k_fill, dim3(blocks((long)epochs * p->members)), dim3(256), 0, s, dloss,
                       (long)epochs * p->members, 0.0f);
    rc = status(hipGetLastError());
  }
  if (!rc) rc = status(hipStreamSynchronize(s));  // `zero` and `perm` are host locals
  if (!rc) {
    rc = status(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    if (!rc) {
      const int erc = enqueue_epoch(p, dperm, depoch, dloss, steps, batch, epochs, lr_max, lr_min);
      const hipError_t ec = hipStreamEndCapture(s, &graph);
      rc = erc ? erc : status(ec);
    }
  }
  if (!rc) rc = status(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  for (int ep = 0; ep < epochs && !rc; ++ep) rc = status(hipGraphLaunch(exec, s));
  if (!rc) rc = status(hipStreamSynchronize(s));
  if (!rc) p->step += (long)epochs * steps;
  if (!rc && loss) {
    std::vector<float> l((size_t)epochs * p->members);
    rc = status(hipMemcpy(l.data(), dloss, sizeof(float) * l.size(), hipMemcpyDeviceToHost));
    for (int ep = 0; ep < epochs && !rc; ++ep) {
      double sum = 0.0;
      for (int e = 0; e < p->members; ++e) sum += l[(size_t)ep * p->members + e];
      loss[ep] = (float)sum;
    }
  }
  (void)hipStreamSynchronize(s);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipFree(dperm);
  (void)hipFree(depoch);
  (void)hipFree(dloss);
  return rc;
}

// One product of the surrogate's GEMM on host arrays (BF16 given as raw 16-bit patterns), for
// tests and timing: C = op(A) op(B) with the strides of k_gemm; tile 0 (by size), 1 (64 x 64)
// or 2 (128 x 128).
int mithra_mlp_gemm(const unsigned short* A, long sam, long sak, long a_len, const unsigned short* B, long sbk,
                    long sbn, long b_len, float* C, int M, int N, int K, int tile, int reps, float* ms) {
  if (!A || !B || !C || M < 1 || N < 1 || K < 1 || reps < 1 || tile < 0 || tile > 2) return MITHRA_E_ARG;
  // every element op(A) and op(B) reach must lie inside the arrays given
  if ((M - 1) * sam + (K - 1) * sak >= a_len || (K - 1) * sbk + (N - 1) * sbn >= b_len || sam < 0 || sak < 0 ||
      sbk < 0 || sbn < 0)
    return MITHRA_E_ARG;
  __bf16 *dA = nullptr, *dB = nullptr;
  float* dC = nullptr;
  int rc = MITHRA_OK;
  if (hipMalloc(&dA, 2 * a_len) != hipSuccess || hipMalloc(&dB, 2 * b_len) != hipSuccess ||
      hipMalloc(&dC, sizeof(float) * (size_t)M * N) != hipSuccess)
    rc = MITHRA_E_NOMEM;
  if (!rc) rc = status(hipMemcpy(dA, A, 2 * a_len, hipMemcpyHostToDevice));
  if (!rc) rc = status(hipMemcpy(dB, B, 2 * b_len, hipMemcpyHostToDevice));
  GemmArgs g{};
  g.A = dA; g.sam = sam; g.sak = sak; g.sae = 0;
  g.B = dB; g.sbk = sbk; g.sbn = sbn; g.sbe = 0;
  g.C = dC; g.ldc = N; g.sce = 0;
  g.M = M; g.N = N; g.K = K;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) rc = status(hipEventCreate(&e0));
  if (!rc) rc = status(hipEventCreate(&e1));
  if (!rc) rc = status(gemm(0, g, 1, tile));  // warm-up
  if (!rc) rc = status(hipEventRecord(e0, 0));
  for (int r = 0; r < reps && !rc; ++r) rc = status(gemm(0, g, 1, tile));
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
