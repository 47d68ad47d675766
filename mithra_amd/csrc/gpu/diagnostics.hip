// Device side of the diagnostics writers of the drop-in host (FdTd::fieldSample and the
// in-plane visualizations, Solver::bunchSample, Solver::screenProfile): the host writes the
// reference's files; these calls hand it the few values each file line needs instead of the
// whole domain.
//
//   sample_nodes      E/B as last evaluated, A^n and phi^n at a list of nodes (the corners the
//                     samplers interpolate between), gathered on the device.
//   bunch_sums        the charge-weighted sums of Solver::bunchSample (src/solver.cpp:1594-1607)
//                     on a fixed grid: each workgroup sums a fixed strided subset in a fixed
//                     order and the host adds the workgroup sums in order, so a domain's sums
//                     do not change from run to run.
//   screen_crossings  the particles that crossed a lab-frame screen during the last field step
//                     and the record Solver::screenProfile writes for each (src/solver.cpp:2205-2257).
//   power_map_*       Solver::powerVisualize (src/radiation.cpp:324-450): a ring of the plane's
//                     lab-frame fields updated every step, and the per-node DFT power at the
//                     steps the host writes the map.
#include <algorithm>
#include <vector>

#include "common.h"

using namespace mithra;

namespace {

inline int hip_status(hipError_t e) { return e == hipSuccess ? MITHRA_OK : MITHRA_E_HIP - (int)e; }

#define DIAG_TRY(expr)                           \
  do {                                           \
    hipError_t _e = (expr);                      \
    if (_e != hipSuccess) return hip_status(_e); \
  } while (0)

constexpr int kSumBlocks = 256;
constexpr int kSums = 13;  // q, q r[3], r^2 q [3], q gb[3], gb^2 q [3]

// Room for `bytes` of per-call staging in the domain's diagnostics buffer.
int diag_room(mithra_dev* d, size_t bytes) {
  if (bytes <= d->diag_bytes) return MITHRA_OK;
  DIAG_TRY(hipStreamSynchronize(d->stream));
  (void)hipFree(d->diag);
  d->diag = nullptr;
  d->diag_bytes = 0;
  if (hipMalloc(&d->diag, bytes) != hipSuccess) return MITHRA_E_NOMEM;
  d->diag_bytes = bytes;
  return MITHRA_OK;
}

// out_f[6 q .. 6 q + 5] = E, B of node nodes[q]; out_d[4 q .. 4 q + 3] = A^n, phi^n (0 without
// space charge).  Node indices were checked against the mesh on the host.
template <typename R>
__global__ __launch_bounds__(256) void k_sample_nodes(const long* __restrict__ nodes, long n,
                                                      const float* __restrict__ eb, FieldSet an, int ncomp,
                                                      float* __restrict__ out_f, double* __restrict__ out_d) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const long m = nodes[q];
#pragma unroll
  for (int c = 0; c < 6; ++c) out_f[6 * q + c] = eb[6 * m + c];
#pragma unroll
  for (int c = 0; c < 3; ++c) out_d[4 * q + c] = (double)((const R*)component(an, c))[m];
  out_d[4 * q + 3] = ncomp == 4 ? (double)((const R*)component(an, 3))[m] : 0.0;
}

// Solver::bunchSample's sums in its operation order per particle: q; r += q rnp; r2 += rnp rnp q;
// gb += q gb; gb2 += gb gb q.  partial[kSums * block + s].
__global__ __launch_bounds__(256) void k_bunch_sums(const double* __restrict__ part, long cap, long n,
                                                    double* __restrict__ partial) {
  __shared__ double sh[kSums][256];
  double s[kSums];
#pragma unroll
  for (int i = 0; i < kSums; ++i) s[i] = 0.0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    const double q = part[PQ * cap + p];
    s[0] += q;
#pragma unroll
    for (int l = 0; l < 3; ++l) {
      const double r = part[(PX + l) * cap + p], g = part[(PGX + l) * cap + p];
      s[1 + l] += q * r;
      s[4 + l] += r * r * q;
      s[7 + l] += q * g;
      s[10 + l] += g * g * q;
    }
  }
#pragma unroll
  for (int i = 0; i < kSums; ++i) sh[i][threadIdx.x] = s[i];
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int i = 0; i < kSums; ++i) sh[i][threadIdx.x] += sh[i][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x < kSums) partial[kSums * blockIdx.x + threadIdx.x] = sh[threadIdx.x][0];
}

// Solver::interp (src/solver.cpp:2283-2286).
__device__ inline double interp(double x0, double x1, double y0, double y1, double x) {
  return y0 + (x - x0) / (x1 - x0) * (y1 - y0);
}

struct ScreenConst {
  double gamma, beta, c0, time_bunch, time_step, dt_shift;
  int n_screens;
};

// One thread per particle, screens in order.  Crossings are appended through `count`; a record
// is written only while its slot is below `cap`, and the count keeps running, so the host
// learns how much room a second pass needs.
__global__ __launch_bounds__(256) void k_screens(const double* __restrict__ part, const unsigned* __restrict__ pid,
                                                 long cap_p, long n, const double* __restrict__ lz, ScreenConst sc,
                                                 unsigned long long* __restrict__ count, long cap,
                                                 int* __restrict__ screen, long* __restrict__ index,
                                                 double* __restrict__ rec) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double zm = part[PZM * cap_p + p], zp = part[PZ * cap_p + p];
  const double lzm = sc.gamma * (zm + sc.beta * sc.c0 * (sc.time_bunch - sc.time_step + sc.dt_shift));
  const double lzp = sc.gamma * (zp + sc.beta * sc.c0 * (sc.time_bunch + sc.dt_shift));
  for (int s = 0; s < sc.n_screens; ++s) {
    const double l = lz[s];
    if (lzm >= l || lzp < l) continue;
    const unsigned long long slot = atomicAdd(count, 1ULL);
    if ((long)slot >= cap) continue;
    const double g0 = part[PGX * cap_p + p], g1 = part[PGY * cap_p + p], g2 = part[PGZ * cap_p + p];
    const double tm = sc.gamma * (sc.time_bunch + sc.dt_shift - sc.time_step + sc.beta / sc.c0 * zm);
    const double tp = sc.gamma * (sc.time_bunch + sc.dt_shift + sc.beta / sc.c0 * zp);
    double* o = rec + 6 * (long)slot;
    o[0] = interp(lzm, lzp, part[PXM * cap_p + p], part[PX * cap_p + p], l);
    o[1] = interp(lzm, lzp, part[PYM * cap_p + p], part[PY * cap_p + p], l);
    o[2] = interp(lzm, lzp, tm, tp, l);
    o[3] = g0;
    o[4] = g1;
    o[5] = sc.gamma * (g2 + sc.beta * sqrt(1.0 + (g0 * g0 + g1 * g1 + g2 * g2)));
    screen[slot] = s;
    index[slot] = pid[p];
  }
}

__global__ void k_zero_count(unsigned long long* c) { *c = 0ULL; }

struct MapConst {
  int N0, N1, k;
  long N1N0;
  unsigned Nf;
  double dzr, gamma, beta, c0, pc;
};

// Solver::powerVisualize's time samples (src/radiation.cpp:346-377): the lab-frame Ex, Ey, Bx, By
// of the plane, interpolated between cells k and k + 1, into ring row m for every interior node.
__global__ __launch_bounds__(256) void k_map_sample(const float* __restrict__ eb, MapConst c, unsigned m,
                                                    double* __restrict__ ring) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ni_total = (long)(c.N0 - 2) * (c.N1 - 2);
  if (q >= ni_total) return;
  const int i = 1 + (int)(q / (c.N1 - 2)), j = 1 + (int)(q % (c.N1 - 2));
  const long ni = (long)i * c.N1 + j;
  const long mi = (long)c.k * c.N1N0 + ni;
  const float* lo = eb + 6 * mi;
  const float* hi = eb + 6 * (mi + c.N1N0);
  const double et0 = (1.0 - c.dzr) * lo[0] + c.dzr * hi[0];
  const double et1 = (1.0 - c.dzr) * lo[1] + c.dzr * hi[1];
  const double bt0 = (1.0 - c.dzr) * lo[3] + c.dzr * hi[3];
  const double bt1 = (1.0 - c.dzr) * lo[4] + c.dzr * hi[4];
  double* r = ring + (size_t)m * 4 * c.N1N0;
  r[ni] = c.gamma * (et0 + c.c0 * c.beta * bt1);
  r[c.N1N0 + ni] = c.gamma * (et1 - c.c0 * c.beta * bt0);
  r[2 * c.N1N0 + ni] = c.gamma * (bt0 - c.beta / c.c0 * et1);
  r[3 * c.N1N0 + ni] = c.gamma * (bt1 + c.beta / c.c0 * et0);
}

// pL[ni] = pc (Re(ew1 bw1) - Re(ew2 bw2)) with ew = sum_m f[m] ep[m], bw = sum_m f[m] em[m] in the
// reference's order (src/radiation.cpp:379-392); ep = cos + i sin, em = cos - i sin.
__global__ __launch_bounds__(256) void k_map_power(const double* __restrict__ ring, const double* __restrict__ dft,
                                                   MapConst c, double* __restrict__ out) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long ni_total = (long)(c.N0 - 2) * (c.N1 - 2);
  if (q >= ni_total) return;
  const int i = 1 + (int)(q / (c.N1 - 2)), j = 1 + (int)(q % (c.N1 - 2));
  const long ni = (long)i * c.N1 + j;
  double e1r = 0.0, e1i = 0.0, b1r = 0.0, b1i = 0.0, e2r = 0.0, e2i = 0.0, b2r = 0.0, b2i = 0.0;
  for (unsigned m = 0; m < c.Nf; ++m) {
    const double* r = ring + (size_t)m * 4 * c.N1N0;
    const double cs = dft[2 * m], sn = dft[2 * m + 1];
    const double f0 = r[ni], f1 = r[c.N1N0 + ni], f2 = r[2 * c.N1N0 + ni], f3 = r[3 * c.N1N0 + ni];
    e1r += f0 * cs;
    e1i += f0 * sn;
    b1r += f3 * cs;
    b1i += f3 * -sn;
    e2r += f1 * cs;
    e2i += f1 * sn;
    b2r += f2 * cs;
    b2i += f2 * -sn;
  }
  out[ni] = c.pc * ((e1r * b1r - e1i * b1i) - (e2r * b2r - e2i * b2i));
}

MapConst map_const(const mithra_dev* d) {
  MapConst c;
  c.N0 = d->g.N0;
  c.N1 = d->g.N1;
  c.k = d->vk;
  c.N1N0 = d->N1N0;
  c.Nf = d->vNf;
  c.dzr = d->vdzr;
  c.gamma = d->vgamma;
  c.beta = d->vbeta;
  c.c0 = d->vc0;
  c.pc = d->vpc;
  return c;
}

}  // namespace

extern "C" {

int mithra_dev_sample_nodes(mithra_dev* d, const long* nodes, long n, float* e, float* b, double* a, double* phi) {
  if (!d || n < 0 || (n > 0 && !nodes)) return MITHRA_E_ARG;
  if (n == 0) return MITHRA_OK;
  for (long q = 0; q < n; ++q)
    if (nodes[q] < 0 || nodes[q] >= d->nodes) return MITHRA_E_ARG;
  if (int rc = settle(d)) return rc;
  const size_t bn = sizeof(long) * n, bf = sizeof(float) * 6 * n, bd = sizeof(double) * 4 * n;
  if (int rc = diag_room(d, bn + bf + bd)) return rc;
  char* base = (char*)d->diag;
  long* dn = (long*)base;
  double* od = (double*)(base + bn);
  float* of = (float*)(base + bn + bd);
  DIAG_TRY(hipMemcpyAsync(dn, nodes, bn, hipMemcpyHostToDevice, d->stream));
  const FieldSet an = d->lvl[d->i_n];
  const dim3 grid((unsigned)((n + 255) / 256));
  if (d->precision == MITHRA_PRECISION_FP64)
    hipLaunchKernelGGL(k_sample_nodes<double>, grid, dim3(256), 0, d->stream, dn, n, d->eb, an, d->ncomp, of, od);
  else
    hipLaunchKernelGGL(k_sample_nodes<float>, grid, dim3(256), 0, d->stream, dn, n, d->eb, an, d->ncomp, of, od);
  DIAG_TRY(hipGetLastError());
  std::vector<float> hf(6 * n);
  std::vector<double> hd(4 * n);
  DIAG_TRY(hipMemcpyAsync(hf.data(), of, bf, hipMemcpyDeviceToHost, d->stream));
  DIAG_TRY(hipMemcpyAsync(hd.data(), od, bd, hipMemcpyDeviceToHost, d->stream));
  DIAG_TRY(hipStreamSynchronize(d->stream));
  for (long q = 0; q < n; ++q)
    for (int c = 0; c < 3; ++c) {
      if (e) e[3 * q + c] = hf[6 * q + c];
      if (b) b[3 * q + c] = hf[6 * q + 3 + c];
      if (a) a[3 * q + c] = hd[4 * q + c];
    }
  if (phi)
    for (long q = 0; q < n; ++q) phi[q] = hd[4 * q + 3];
  return MITHRA_OK;
}

int mithra_dev_bunch_sums(mithra_dev* d, double* out) {
  if (!d || !out) return MITHRA_E_ARG;
  std::fill(out, out + kSums, 0.0);
  if (d->np == 0) return MITHRA_OK;
  const size_t bytes = sizeof(double) * kSums * kSumBlocks;
  if (int rc = diag_room(d, bytes)) return rc;
  double* partial = (double*)d->diag;
  hipLaunchKernelGGL(k_bunch_sums, dim3(kSumBlocks), dim3(256), 0, d->stream, d->part, d->cap, d->np, partial);
  DIAG_TRY(hipGetLastError());
  std::vector<double> h(kSums * kSumBlocks);
  DIAG_TRY(hipMemcpyAsync(h.data(), partial, bytes, hipMemcpyDeviceToHost, d->stream));
  DIAG_TRY(hipStreamSynchronize(d->stream));
  for (int blk = 0; blk < kSumBlocks; ++blk)
    for (int s = 0; s < kSums; ++s) out[s] += h[kSums * blk + s];
  return MITHRA_OK;
}

int mithra_dev_screen_crossings(mithra_dev* d, const double* lz, int n_screens, const mithra_screen_params* p,
                                long cap, long* count, int* screen, long* index, double* rec) {
  if (!d || !p || !count || n_screens < 0 || cap < 0 || (n_screens > 0 && !lz) ||
      (cap > 0 && (!screen || !index || !rec)))
    return MITHRA_E_ARG;
  *count = 0;
  if (d->np == 0 || n_screens == 0) return MITHRA_OK;
  const long slots = std::max(cap, 1L);
  const size_t b_cnt = 256, b_lz = sizeof(double) * n_screens, b_scr = sizeof(int) * slots,
               b_idx = sizeof(long) * slots, b_rec = sizeof(double) * 6 * slots;
  auto up = [](size_t x) { return (x + 255) / 256 * 256; };
  if (int rc = diag_room(d, up(b_cnt) + up(b_lz) + up(b_scr) + up(b_idx) + up(b_rec))) return rc;
  char* base = (char*)d->diag;
  auto* dcount = (unsigned long long*)base;
  base += up(b_cnt);
  auto* dlz = (double*)base;
  base += up(b_lz);
  auto* dscr = (int*)base;
  base += up(b_scr);
  auto* didx = (long*)base;
  base += up(b_idx);
  auto* drec = (double*)base;
  DIAG_TRY(hipMemcpyAsync(dlz, lz, b_lz, hipMemcpyHostToDevice, d->stream));
  hipLaunchKernelGGL(k_zero_count, dim3(1), dim3(1), 0, d->stream, dcount);
  DIAG_TRY(hipGetLastError());
  ScreenConst sc{p->gamma, p->beta, p->c0, p->time_bunch, p->time_step, p->dt_shift, n_screens};
  hipLaunchKernelGGL(k_screens, dim3((unsigned)((d->np + 255) / 256)), dim3(256), 0, d->stream, d->part, d->pid,
                     d->cap, d->np, dlz, sc, dcount, cap, dscr, didx, drec);
  DIAG_TRY(hipGetLastError());
  unsigned long long c = 0;
  DIAG_TRY(hipMemcpyAsync(&c, dcount, sizeof c, hipMemcpyDeviceToHost, d->stream));
  DIAG_TRY(hipStreamSynchronize(d->stream));
  *count = (long)c;
  const long got = std::min((long)c, cap);
  if (got > 0) {
    DIAG_TRY(hipMemcpyAsync(screen, dscr, sizeof(int) * got, hipMemcpyDeviceToHost, d->stream));
    DIAG_TRY(hipMemcpyAsync(index, didx, sizeof(long) * got, hipMemcpyDeviceToHost, d->stream));
    DIAG_TRY(hipMemcpyAsync(rec, drec, sizeof(double) * 6 * got, hipMemcpyDeviceToHost, d->stream));
    DIAG_TRY(hipStreamSynchronize(d->stream));
  }
  return MITHRA_OK;
}

int mithra_dev_power_map_setup(mithra_dev* d, const mithra_power_map_params* p) {
  if (!d || !p || p->Nf == 0 || !p->cos_wt || !p->sin_wt) return MITHRA_E_ARG;
  if (p->k < 0 || p->k + 1 >= d->g.N2 || d->g.N0 < 3 || d->g.N1 < 3) return MITHRA_E_ARG;
  DIAG_TRY(hipStreamSynchronize(d->stream));
  (void)hipFree(d->vring);
  (void)hipFree(d->vdft);
  (void)hipFree(d->vmap);
  d->vring = d->vdft = d->vmap = nullptr;
  d->vk = -1;
  const size_t ring = sizeof(double) * p->Nf * 4 * (size_t)d->N1N0;
  if (hipMalloc(&d->vring, ring) != hipSuccess || hipMalloc(&d->vdft, sizeof(double) * 2 * p->Nf) != hipSuccess ||
      hipMalloc(&d->vmap, sizeof(double) * (size_t)d->N1N0) != hipSuccess) {
    (void)hipFree(d->vring);
    (void)hipFree(d->vdft);
    (void)hipFree(d->vmap);
    d->vring = d->vdft = d->vmap = nullptr;
    return MITHRA_E_NOMEM;
  }
  std::vector<double> dft(2 * (size_t)p->Nf);
  for (unsigned j = 0; j < p->Nf; ++j) {
    dft[2 * j] = p->cos_wt[j];
    dft[2 * j + 1] = p->sin_wt[j];
  }
  DIAG_TRY(hipMemcpy(d->vdft, dft.data(), sizeof(double) * dft.size(), hipMemcpyHostToDevice));
  DIAG_TRY(launch_fill_zero(d, d->vring, ring));  // SampleRadiationPower::fdt starts at zero
  DIAG_TRY(launch_fill_zero(d, d->vmap, sizeof(double) * (size_t)d->N1N0));
  d->vNf = p->Nf;
  d->vk = p->k;
  d->vdzr = p->dzr;
  d->vpc = p->pc;
  d->vgamma = p->gamma;
  d->vbeta = p->beta;
  d->vc0 = p->c0;
  return MITHRA_OK;
}

int mithra_dev_power_map_sample(mithra_dev* d, unsigned n_time) {
  if (!d || d->vk < 0) return MITHRA_E_STATE;
  const long n = (long)(d->g.N0 - 2) * (d->g.N1 - 2);
  hipLaunchKernelGGL(k_map_sample, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->stream, d->eb, map_const(d),
                     n_time % d->vNf, d->vring);
  return hip_status(hipGetLastError());
}

int mithra_dev_power_map_fetch(mithra_dev* d, double* p) {
  if (!d || !p) return MITHRA_E_ARG;
  if (d->vk < 0) return MITHRA_E_STATE;
  const long n = (long)(d->g.N0 - 2) * (d->g.N1 - 2);
  hipLaunchKernelGGL(k_map_power, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->stream, d->vring, d->vdft,
                     map_const(d), d->vmap);
  DIAG_TRY(hipGetLastError());
  DIAG_TRY(hipMemcpyAsync(p, d->vmap, sizeof(double) * (size_t)d->N1N0, hipMemcpyDeviceToHost, d->stream));
  return hip_status(hipStreamSynchronize(d->stream));
}

}  // extern "C"
