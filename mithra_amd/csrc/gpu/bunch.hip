// Bunches generated and boosted on the device.
//
// A mithra_bunch is a list of Charge records (the reference's q, rnp, rnm, gb, e) held on the
// device in the particle arrays' SoA layout.  Two generators fill it:
//  * the reference's ellipsoid (Bunch::initializeEllipsoid, src/classes.cpp:104-298) with its
//    halton quiet start: the halton numbers are integer and correctly rounded arithmetic and so
//    agree with the host's bit for bit; positions and momenta go through the device's log, sin
//    and cos, an ulp or so from the host's;
//  * a seeded quiet-start flat-top slice (four beamlets lambda/4 apart, one shot-noise modulation
//    per lambda bucket), drawn from a counter-based generator so every record is a pure
//    function of (seed, index).
// Then the particle loops of Solver::lorentzBoostBunch (src/solver.cpp:263-423) and
// Solver::shiftBackInTime (src/solver.cpp:1185-1206) run on the device too, and a domain takes
// the records over without a round trip through the host.
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "common.h"

using namespace mithra;

struct mithra_bunch {
  double* part = nullptr;   // PCOUNT arrays of cap doubles (x y z xm ym zm gx gy gz q e)
  long n = 0, cap = 0;
  double* red = nullptr;    // per-workgroup partial reductions
  unsigned* flags = nullptr;  // generator scratch: accept flags, then their exclusive scan
  long flags_cap = 0;
  void* scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
};

namespace {

inline int status(hipError_t e) { return e == hipSuccess ? MITHRA_OK : MITHRA_E_HIP - (int)e; }

#define BUNCH_TRY(expr)                      \
  do {                                       \
    hipError_t _e = (expr);                  \
    if (_e != hipSuccess) return status(_e); \
  } while (0)

constexpr int kRedBlocks = 1024;

// halton(i, j) of the reference (src/stdinclude.cpp:45-73): the radical inverse of j + 1 in base
// prime[i], returned as 1 - x.  Integer steps and correctly rounded divisions only.
__device__ inline double halton(int i, unsigned j) {
  const int prime[20] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71};
  const int p = prime[i];
  int p0 = p;
  int k = (int)j + 1;
  double x = 0.0;
  while (k > 0) {
    const int a = k % p;
    x += a / (double)p0;
    k = k / p;
    p0 *= p;
  }
  return 1.0 - x;
}

// Box-Muller pair of the reference's draws: s sqrt(-2 log u) cos|sin(2 PI v)
__device__ inline double bm_cos(double s, double u, double v) { return s * sqrt(-2.0 * log(u)) * cos(2.0 * kPI * v); }
__device__ inline double bm_sin(double s, double u, double v) { return s * sqrt(-2.0 * log(u)) * sin(2.0 * kPI * v); }

// One draw of the ellipsoid generator (the `generate` lambda and its loop body,
// src/classes.cpp:200-297): position offset r, momentum offset t, and whether the truncation keeps it.
__device__ inline bool ellipsoid_draw(const mithra_ellipsoid& e, unsigned i, bool tail, double r[3], double t[3]) {
  const unsigned m = i + e.first_index;
  r[0] = bm_cos(e.sigma_position[0], halton(0, m), halton(1, m));
  r[1] = bm_sin(e.sigma_position[1], halton(0, m), halton(1, m));
  if (tail) {
    r[2] = bm_sin(2.0 * e.lambda, halton(2, m), halton(3, m));
    r[2] += (r[2] < 0.0) ? (-e.sigma_position[2]) : (e.sigma_position[2]);
  } else if (e.distribution == MITHRA_DISTRIBUTION_UNIFORM) {
    r[2] = (2.0 * halton(2, m) - 1.0) * e.sigma_position[2];
  } else {
    r[2] = bm_sin(e.sigma_position[2], halton(2, m), halton(3, m));
  }
  t[0] = bm_cos(e.sigma_gb[0], halton(4, m), halton(5, m));
  t[1] = bm_sin(e.sigma_gb[1], halton(4, m), halton(5, m));
  t[2] = bm_cos(e.sigma_gb[2], halton(6, m), halton(7, m));
  return fabs(r[0]) < e.tran_trunc && fabs(r[1]) < e.tran_trunc && fabs(r[2]) < e.long_trunc;
}

// The shot-noise z origin: the smallest longitudinal draw (src/classes.cpp:179-200), per workgroup.
__global__ __launch_bounds__(256) void k_ellipsoid_zmin(mithra_ellipsoid e, unsigned n_core, unsigned n_all,
                                                        double* __restrict__ partial) {
  __shared__ double s[256];
  double zmin = 1e100;
  for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < n_all; i += gridDim.x * blockDim.x) {
    const unsigned m = i + e.first_index;
    double z;
    if (i < n_core) {
      z = e.distribution == MITHRA_DISTRIBUTION_UNIFORM ? (2.0 * halton(2, m) - 1.0) * e.sigma_position[2]
                                                       : bm_sin(e.sigma_position[2], halton(2, m), halton(3, m));
    } else {
      z = bm_sin(2.0 * e.lambda, halton(2, m), halton(3, m));
      z += (z < 0.0) ? (-e.sigma_position[2]) : (e.sigma_position[2]);
    }
    zmin = fmin(zmin, z);
  }
  s[threadIdx.x] = zmin;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] = fmin(s[threadIdx.x], s[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void k_ellipsoid_flags(mithra_ellipsoid e, unsigned n_core, unsigned n_all,
                                                         unsigned* __restrict__ flags) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_all) return;
  double r[3], t[3];
  flags[i] = ellipsoid_draw(e, i, i >= n_core, r, t) ? 1u : 0u;
}

// insertCharge (src/classes.cpp:160-187): ng records per kept draw, beamlet ii shifted by
// lambda / 4 ii and modulated (shot noise: amplitude and phase from halton dimensions 8 and 9 of
// the draw's lambda bucket; otherwise the given bunching factor and phase).
__global__ __launch_bounds__(256) void k_ellipsoid_insert(mithra_ellipsoid e, unsigned n_core, unsigned n_all,
                                                          const unsigned* __restrict__ offset, double zmin,
                                                          double bF, int ng, double* __restrict__ part, long cap,
                                                          long base, double q) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_all) return;
  double r[3], t[3];
  if (!ellipsoid_draw(e, i, i >= n_core, r, t)) return;
  const long o = base + (long)ng * offset[i];
  const double rz = e.position[2] + r[2];
  for (int ii = 0; ii < ng; ++ii) {
    double z = rz;
    if (e.shot_noise) {
      const int bmi = (int)((rz - zmin) / e.lambda);
      const double bFi = bF * sqrt(-2.0 * log(halton(8, (unsigned)bmi)));
      z = rz - e.lambda / 4 * ii;
      z -= e.lambda / kPI * bFi * sin(2.0 * kPI / e.lambda * z + 2.0 * kPI * halton(9, (unsigned)bmi));
    } else if (e.lambda != 0.0) {
      z = rz - e.lambda / 4 * ii;
      z -= e.lambda / kPI * e.bunching * sin(2.0 * kPI / e.lambda * z + e.bunching_phase * kPI / 180.0);
    }
    const long p = o + ii;
    part[PX * cap + p] = e.position[0] + r[0];
    part[PY * cap + p] = e.position[1] + r[1];
    part[PZ * cap + p] = z;
    part[PXM * cap + p] = 0.0;
    part[PYM * cap + p] = 0.0;
    part[PZM * cap + p] = 0.0;
    part[PGX * cap + p] = e.gamma * e.beta_vector[0] + t[0];
    part[PGY * cap + p] = e.gamma * e.beta_vector[1] + t[1];
    part[PGZ * cap + p] = e.gamma * e.beta_vector[2] + t[2];
    part[PQ * cap + p] = q;
    part[PE * cap + p] = 0.0;
  }
}

// Counter-based uniform numbers in (0, 1]: SplitMix64 of (seed, stream, index).
__device__ inline double uniform01(unsigned long long seed, unsigned stream, unsigned long long i) {
  unsigned long long z = seed + 0x9E3779B97F4A7C15ull * (i * 16ull + stream + 1ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((z >> 11) + 1ull) * (1.0 / 9007199254740992.0);
}

__device__ inline double normal(unsigned long long seed, unsigned stream, unsigned long long i) {
  return sqrt(-2.0 * log(uniform01(seed, stream, i))) * cos(2.0 * kPI * uniform01(seed, stream + 1, i));
}

// The flat-top quiet start: base particle b (n / 4 of them) at a uniform z in the slice, Gaussian
// in x, y and the momenta; its four beamlets lambda / 4 apart share everything else, and bucket
// floor((z - z_start) / lambda) gives every beamlet of the bucket one modulation
// z -= lambda / pi a sin(2 pi z / lambda + phase) with a Rayleigh amplitude a of mean square 2 / Ne
// and a uniform phase (the shot noise of Ne electrons per bucket).
__global__ __launch_bounds__(256) void k_flattop(mithra_flattop f, unsigned long long seed, long nb,
                                                 double* __restrict__ part, long cap, long base, double q) {
  const long b = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const double zb = f.z_center + f.length * (uniform01(seed, 0, b) - 0.5);
  const double x = f.sigma_r * normal(seed, 1, b), y = f.sigma_r * normal(seed, 3, b);
  const double gx = f.sigma_gb_perp * normal(seed, 5, b), gy = f.sigma_gb_perp * normal(seed, 7, b);
  const double gz = sqrt(f.gamma * f.gamma - 1.0) * (1.0 + f.energy_spread * normal(seed, 9, b));
  const long bucket = (long)floor((zb - (f.z_center - 0.5 * f.length)) / f.lambda);
  const double amp = sqrt(-2.0 * log(uniform01(seed, 11, bucket))) / sqrt(f.electrons_per_bucket);
  const double phase = 2.0 * kPI * uniform01(seed, 12, bucket);
  for (int ii = 0; ii < 4; ++ii) {
    double z = zb - f.lambda / 4.0 * ii;
    z -= f.lambda / kPI * amp * sin(2.0 * kPI / f.lambda * z + phase);
    const long p = base + 4 * b + ii;
    part[PX * cap + p] = x;
    part[PY * cap + p] = y;
    part[PZ * cap + p] = z;
    part[PXM * cap + p] = 0.0;
    part[PYM * cap + p] = 0.0;
    part[PZM * cap + p] = 0.0;
    part[PGX * cap + p] = gx;
    part[PGY * cap + p] = gy;
    part[PGZ * cap + p] = gz;
    part[PQ * cap + p] = q;
    part[PE * cap + p] = 0.0;
  }
}

// Solver::lorentzBoostBunch, first loop (src/solver.cpp:290-302): the longitudinal position and
// momentum into the frame moving with beta; per-workgroup maxima of the new z.
__global__ __launch_bounds__(256) void k_boost(double* __restrict__ part, long cap, long n, double gamma, double beta,
                                               double* __restrict__ partial) {
  __shared__ double s[256];
  double zmax = -1.0e100;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    const double g0 = part[PGX * cap + p], g1 = part[PGY * cap + p], g2 = part[PGZ * cap + p];
    const double g = sqrt(1.0 + (g0 * g0 + g1 * g1 + g2 * g2));
    const double bz = g2 / g;
    const double z = part[PZ * cap + p] * gamma;
    part[PZ * cap + p] = z;
    part[PGZ * cap + p] = gamma * g * (bz - beta);
    zmax = fmax(zmax, z);
  }
  s[threadIdx.x] = zmax;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + h]);
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = s[0];
}

// Second loop (src/solver.cpp:340-346): the straight-line drift to the start point, which every
// component takes with the particle's z before the drift.
__global__ __launch_bounds__(256) void k_drift(double* __restrict__ part, long cap, long n, double beta, double zu) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double g0 = part[PGX * cap + p], g1 = part[PGY * cap + p], g2 = part[PGZ * cap + p];
  const double g = sqrt(1.0 + (g0 * g0 + g1 * g1 + g2 * g2));
  const double z = part[PZ * cap + p];
  part[PX * cap + p] += g0 / g * (z - zu) * beta;
  part[PY * cap + p] += g1 / g * (z - zu) * beta;
  part[PZ * cap + p] = z + g2 / g * (z - zu) * beta;
}

// Per-workgroup sums of z and beta_z and the minimum of z (src/solver.cpp:354-366, :396-413), each
// workgroup in a fixed order, so the moments do not depend on scheduling.
__global__ __launch_bounds__(256) void k_moments(const double* __restrict__ part, long cap, long n,
                                                 double* __restrict__ partial) {
  __shared__ double s[3][256];
  double zs = 0.0, bs = 0.0, zmin = 1e100;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    const double g0 = part[PGX * cap + p], g1 = part[PGY * cap + p], g2 = part[PGZ * cap + p];
    const double z = part[PZ * cap + p];
    zs += z;
    bs += g2 / sqrt(1 + (g0 * g0 + g1 * g1 + g2 * g2));
    zmin = fmin(zmin, z);
  }
  s[0][threadIdx.x] = zs;
  s[1][threadIdx.x] = bs;
  s[2][threadIdx.x] = zmin;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s[0][threadIdx.x] += s[0][threadIdx.x + h];
      s[1][threadIdx.x] += s[1][threadIdx.x + h];
      s[2][threadIdx.x] = fmin(s[2][threadIdx.x], s[2][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) partial[3 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_shift_z(double* __restrict__ part, long cap, long n, double shift) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) part[PZ * cap + p] -= shift;
}

// Solver::shiftBackInTime (src/solver.cpp:1195-1199): r -= c0 dt / gamma * gb.
__global__ __launch_bounds__(256) void k_time_shift(double* __restrict__ part, long cap, long n, double c0, double dt) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double g0 = part[PGX * cap + p], g1 = part[PGY * cap + p], g2 = part[PGZ * cap + p];
  const double t = c0 * dt / sqrt(1.0 + (g0 * g0 + g1 * g1 + g2 * g2));
  part[PX * cap + p] -= t * g0;
  part[PY * cap + p] -= t * g1;
  part[PZ * cap + p] -= t * g2;
}

unsigned blocks_of(long n) { return (unsigned)((n + 255) / 256); }

int red_blocks(long n) { return (int)std::max(1L, std::min((long)kRedBlocks, (n + 255) / 256)); }

// Room for `extra` more records, keeping the present ones.
int reserve(mithra_bunch* b, long extra) {
  if (b->n + extra <= b->cap) return MITHRA_OK;
  const long cap = (b->n + extra + 1023) / 1024 * 1024;
  double* p = nullptr;
  if (hipMalloc(&p, sizeof(double) * PCOUNT * (size_t)cap) != hipSuccess) return MITHRA_E_NOMEM;
  for (int a = 0; a < PCOUNT && b->n > 0; ++a) {
    const hipError_t e = hipMemcpy(p + (size_t)a * cap, b->part + (size_t)a * b->cap, sizeof(double) * b->n,
                                   hipMemcpyDeviceToDevice);
    if (e != hipSuccess) {
      (void)hipFree(p);
      return status(e);
    }
  }
  (void)hipFree(b->part);
  b->part = p;
  b->cap = cap;
  return MITHRA_OK;
}

// The per-workgroup partials of the last reduction kernel, combined on the host in workgroup
// order: width 1 (op 0 max, op 1 min) or the three moments (sum, sum, min).
int combine(mithra_bunch* b, int blocks, int width, int op, double* out) {
  std::vector<double> h((size_t)blocks * width);
  BUNCH_TRY(hipDeviceSynchronize());
  BUNCH_TRY(hipMemcpy(h.data(), b->red, sizeof(double) * h.size(), hipMemcpyDeviceToHost));
  for (int w = 0; w < width; ++w) out[w] = h[w];
  for (int k = 1; k < blocks; ++k)
    for (int w = 0; w < width; ++w) {
      const double v = h[(size_t)k * width + w];
      if (width == 3) out[w] = w < 2 ? out[w] + v : std::fmin(out[w], v);
      else out[w] = op == 0 ? std::fmax(out[w], v) : std::fmin(out[w], v);
    }
  return MITHRA_OK;
}

int ensure_flags(mithra_bunch* b, long n) {
  if (n <= b->flags_cap) return MITHRA_OK;
  (void)hipFree(b->flags);
  b->flags = nullptr;
  b->flags_cap = 0;
  if (hipMalloc(&b->flags, sizeof(unsigned) * 2 * (size_t)n) != hipSuccess) return MITHRA_E_NOMEM;
  b->flags_cap = n;
  return MITHRA_OK;
}

}  // namespace

extern "C" {

int mithra_bunch_create(mithra_bunch** out) {
  if (!out) return MITHRA_E_ARG;
  *out = nullptr;
  mithra_bunch* b = new (std::nothrow) mithra_bunch();
  if (!b) return MITHRA_E_NOMEM;
  if (hipMalloc(&b->red, sizeof(double) * 3 * kRedBlocks) != hipSuccess) {
    delete b;
    return MITHRA_E_NOMEM;
  }
  *out = b;
  return MITHRA_OK;
}

int mithra_bunch_destroy(mithra_bunch* b) {
  if (!b) return MITHRA_E_ARG;
  (void)hipDeviceSynchronize();
  (void)hipFree(b->part);
  (void)hipFree(b->red);
  (void)hipFree(b->flags);
  (void)hipFree(b->scan_tmp);
  delete b;
  return MITHRA_OK;
}

long mithra_bunch_count(const mithra_bunch* b) { return b ? b->n : -1; }

int mithra_bunch_get(mithra_bunch* b, double* q, long n) {
  if (!b || n != b->n || (n > 0 && !q)) return MITHRA_E_ARG;
  if (n == 0) return MITHRA_OK;
  BUNCH_TRY(hipDeviceSynchronize());
  static const int from[PCOUNT] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 0, 10};  // SoA column -> Charge member
  std::vector<double> col(n);
  for (int a = 0; a < PCOUNT; ++a) {
    BUNCH_TRY(hipMemcpy(col.data(), b->part + (size_t)a * b->cap, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (long p = 0; p < n; ++p) q[11 * p + from[a]] = col[p];
  }
  return MITHRA_OK;
}

int mithra_bunch_ellipsoid(mithra_bunch* b, const mithra_ellipsoid* e, long* added) {
  if (!b || !e || e->n <= 0 || e->n >= (1L << 30) || e->lambda < 0.0 || e->sigma_position[2] <= 0.0 ||
      (e->distribution != MITHRA_DISTRIBUTION_UNIFORM && e->distribution != MITHRA_DISTRIBUTION_GAUSSIAN))
    return MITHRA_E_ARG;
  if (e->bunching > 2.0 || e->bunching < 0.0) return MITHRA_E_ARG;  // src/classes.cpp:130-135
  const unsigned Np = (unsigned)((e->n + 3) / 4 * 4);                 // src/classes.cpp:107-113
  const unsigned ng = e->lambda == 0.0 ? 1 : 4;
  const unsigned n_core = Np / ng;
  const bool uniform = e->distribution == MITHRA_DISTRIBUTION_UNIFORM;
  const unsigned n_tail = uniform ? (unsigned)(Np / ng * (1.0 + 2.0 * e->lambda * std::sqrt(2.0 * kPI) /
                                                                  (2.0 * e->sigma_position[2])))
                                  : n_core;
  const unsigned n_all = std::max(n_core, n_tail);
  if (e->shot_noise && e->lambda == 0.0) return MITHRA_E_ARG;
  double zmin = 1e100, bF = 0.0;
  if (e->shot_noise) {  // src/classes.cpp:193-224
    const int blocks = red_blocks(n_all);
    hipLaunchKernelGGL(k_ellipsoid_zmin, dim3(blocks), dim3(256), 0, 0, *e, n_core, n_all, b->red);
    BUNCH_TRY(hipGetLastError());
    if (int rc = combine(b, blocks, 1, 1, &zmin)) return rc;
    zmin = zmin + e->position[2];
    const double Ne = e->charge * e->lambda / (2.0 * e->sigma_position[2]);
    bF = e->bunching == 0.0 ? 1.0 / std::sqrt(Ne) : e->bunching;
  }
  if (int rc = ensure_flags(b, n_all)) return rc;
  unsigned* flags = b->flags;
  unsigned* offset = b->flags + b->flags_cap;
  hipLaunchKernelGGL(k_ellipsoid_flags, dim3(blocks_of(n_all)), dim3(256), 0, 0, *e, n_core, n_all, flags);
  BUNCH_TRY(hipGetLastError());
  size_t bytes = 0;
  BUNCH_TRY(rocprim::exclusive_scan(nullptr, bytes, flags, offset, 0u, n_all, rocprim::plus<unsigned>()));
  if (bytes > b->scan_tmp_bytes) {
    (void)hipFree(b->scan_tmp);
    b->scan_tmp = nullptr;
    b->scan_tmp_bytes = 0;
    if (hipMalloc(&b->scan_tmp, bytes) != hipSuccess) return MITHRA_E_NOMEM;
    b->scan_tmp_bytes = bytes;
  }
  bytes = b->scan_tmp_bytes;
  BUNCH_TRY(rocprim::exclusive_scan(b->scan_tmp, bytes, flags, offset, 0u, n_all, rocprim::plus<unsigned>()));
  unsigned last[2];
  BUNCH_TRY(hipMemcpy(&last[0], offset + (n_all - 1), sizeof(unsigned), hipMemcpyDeviceToHost));
  BUNCH_TRY(hipMemcpy(&last[1], flags + (n_all - 1), sizeof(unsigned), hipMemcpyDeviceToHost));
  const long kept = (long)last[0] + last[1];
  if (int rc = reserve(b, kept * ng)) return rc;
  hipLaunchKernelGGL(k_ellipsoid_insert, dim3(blocks_of(n_all)), dim3(256), 0, 0, *e, n_core, n_all, offset, zmin,
                     bF, (int)ng, b->part, b->cap, b->n, e->charge / Np);
  BUNCH_TRY(hipGetLastError());
  BUNCH_TRY(hipDeviceSynchronize());
  b->n += kept * ng;
  if (added) *added = kept * ng;
  return MITHRA_OK;
}

int mithra_bunch_flattop(mithra_bunch* b, const mithra_flattop* f, unsigned long long seed) {
  if (!b || !f || f->n <= 0 || f->n % 4 != 0 || f->n >= (1L << 31) || f->lambda <= 0.0 || f->length <= 0.0 ||
      f->electrons_per_bucket <= 0.0 || f->gamma < 1.0)
    return MITHRA_E_ARG;
  if (int rc = reserve(b, f->n)) return rc;
  const long nb = f->n / 4;
  hipLaunchKernelGGL(k_flattop, dim3(blocks_of(nb)), dim3(256), 0, 0, *f, seed, nb, b->part, b->cap, b->n,
                     f->charge / f->n);
  BUNCH_TRY(hipGetLastError());
  BUNCH_TRY(hipDeviceSynchronize());
  b->n += f->n;
  return MITHRA_OK;
}

int mithra_bunch_boost(mithra_bunch* b, double gamma, double beta, double* zmax) {
  if (!b || !zmax || gamma < 1.0) return MITHRA_E_ARG;
  *zmax = -1.0e100;
  if (b->n == 0) return MITHRA_OK;
  const int blocks = red_blocks(b->n);
  hipLaunchKernelGGL(k_boost, dim3(blocks), dim3(256), 0, 0, b->part, b->cap, b->n, gamma, beta, b->red);
  BUNCH_TRY(hipGetLastError());
  return combine(b, blocks, 1, 0, zmax);
}

int mithra_bunch_drift(mithra_bunch* b, double beta, double zu) {
  if (!b) return MITHRA_E_ARG;
  if (b->n == 0) return MITHRA_OK;
  hipLaunchKernelGGL(k_drift, dim3(blocks_of(b->n)), dim3(256), 0, 0, b->part, b->cap, b->n, beta, zu);
  BUNCH_TRY(hipGetLastError());
  return status(hipDeviceSynchronize());
}

int mithra_bunch_moments(mithra_bunch* b, double* z_sum, double* bz_sum, double* z_min) {
  if (!b || !z_sum || !bz_sum || !z_min) return MITHRA_E_ARG;
  double m[3] = {0.0, 0.0, 1e100};
  if (b->n > 0) {
    const int blocks = red_blocks(b->n);
    hipLaunchKernelGGL(k_moments, dim3(blocks), dim3(256), 0, 0, b->part, b->cap, b->n, b->red);
    BUNCH_TRY(hipGetLastError());
    if (int rc = combine(b, blocks, 3, 0, m)) return rc;
  }
  *z_sum = m[0];
  *bz_sum = m[1];
  *z_min = m[2];
  return MITHRA_OK;
}

int mithra_bunch_shift(mithra_bunch* b, double shift) {
  if (!b) return MITHRA_E_ARG;
  if (b->n == 0) return MITHRA_OK;
  hipLaunchKernelGGL(k_shift_z, dim3(blocks_of(b->n)), dim3(256), 0, 0, b->part, b->cap, b->n, shift);
  BUNCH_TRY(hipGetLastError());
  return status(hipDeviceSynchronize());
}

int mithra_bunch_time_shift(mithra_bunch* b, double c0, double time_shift) {
  if (!b) return MITHRA_E_ARG;
  if (b->n == 0 || time_shift == 0.0) return MITHRA_OK;
  hipLaunchKernelGGL(k_time_shift, dim3(blocks_of(b->n)), dim3(256), 0, 0, b->part, b->cap, b->n, c0, time_shift);
  BUNCH_TRY(hipGetLastError());
  return status(hipDeviceSynchronize());
}

}  // extern "C"

namespace mithra {

// The bunch's records as a domain's particles: column by column, device to device.
int bunch_records(const mithra_bunch* b, const double** part, long* cap, long* n) {
  if (!b) return MITHRA_E_ARG;
  *part = b->part;
  *cap = b->cap;
  *n = b->n;
  return MITHRA_OK;
}

}  // namespace mithra
