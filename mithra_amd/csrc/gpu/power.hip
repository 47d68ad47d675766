// Radiation power P(z) (Solver::powerSample, reference src/radiation.cpp:127-232).
//
// Pass 1, one thread per (plane, transverse node): interpolate E/B to the plane, transform
// them to the lab frame, store them in ring slot n_time % Nf and evaluate the Nf-point DFT
// of the node's ring for every frequency; a workgroup reduces its nodes to one partial sum
// per (plane, frequency).  Pass 2 adds the partials in a fixed order, so P is deterministic.
#include "common.h"

namespace mithra {

namespace {

constexpr int kPowerBlock = 256;

__device__ inline void power_nodes_body(const float* __restrict__ eb, double* __restrict__ ring,
                                        const double* __restrict__ dft, const int* __restrict__ plane_k,
                                        const double* __restrict__ plane_dzr, double* __restrict__ partial, int N0,
                                        int N1, unsigned Nz, unsigned Nl, unsigned Nf, unsigned slot, double gamma,
                                        double beta, double c0, double pc, int bx, int nbx, unsigned pz) {
  const long ni_count = (long)(N0 - 4) * (N1 - 4);
  const long t = (long)bx * blockDim.x + threadIdx.x;
  const gptr<const f2> ebg = G(reinterpret_cast<const f2*>(eb));
  const gptr<double> ringg = G(ring);
  const gptr<const double> dftg = G(dft);
  const long N1N0 = (long)N1 * N0;
  const size_t ring_slot = (size_t)4 * Nz * N1N0;  // doubles per slot
  __shared__ double red[kPowerBlock];
  for (unsigned l = 0; l < Nl; ++l) {
    double v = 0.0;
    if (t < ni_count) {
      const long i = 2 + t / (N1 - 4), j = 2 + t % (N1 - 4);
      const long node = (long)N1 * i + j;
      const long mi = (long)G(plane_k)[pz] * N1N0 + node;
      const size_t base = (size_t)pz * N1N0 + node;
      if (l == 0) {
        const double dzr = G(plane_dzr)[pz];
        const long na = kEBPairs * mi, nb = kEBPairs * (mi + N1N0);
        const f2 a0 = ebg[na], a1 = ebg[na + 1], a2 = ebg[na + 2];  // {Ex, Ey}, {Ez, Bx}, {By, Bz}
        const f2 b0 = ebg[nb], b1 = ebg[nb + 1], b2 = ebg[nb + 2];
        const double et0 = (1.0 - dzr) * (double)a0.x + dzr * (double)b0.x;
        const double et1 = (1.0 - dzr) * (double)a0.y + dzr * (double)b0.y;
        const double bt0 = (1.0 - dzr) * (double)a1.y + dzr * (double)b1.y;
        const double bt1 = (1.0 - dzr) * (double)a2.x + dzr * (double)b2.x;
        const gptr<double> r = ringg + (size_t)slot * ring_slot;
        r[0 * Nz * N1N0 + base] = gamma * (et0 + c0 * beta * bt1);
        r[1 * Nz * N1N0 + base] = gamma * (et1 - c0 * beta * bt0);
        r[2 * Nz * N1N0 + base] = gamma * (bt0 - beta / c0 * et1);
        r[3 * Nz * N1N0 + base] = gamma * (bt1 + beta / c0 * et0);
      }
      double e1r = 0.0, e1i = 0.0, b1r = 0.0, b1i = 0.0, e2r = 0.0, e2i = 0.0, b2r = 0.0, b2i = 0.0;
      const gptr<const double> w = dftg + (size_t)l * Nf * 2;
      // few workgroups (one per 256 nodes of a plane): unrolled, a thread keeps eight slots'
      // loads in flight; the sums still run over m in order
#pragma unroll 8
      for (unsigned m = 0; m < Nf; ++m) {
        const gptr<const double> r = ringg + (size_t)m * ring_slot;
        const double f0 = r[0 * Nz * N1N0 + base], f1 = r[1 * Nz * N1N0 + base];
        const double f2 = r[2 * Nz * N1N0 + base], f3 = r[3 * Nz * N1N0 + base];
        const double cs = w[2 * m], sn = w[2 * m + 1];
        e1r += f0 * cs; e1i += f0 * sn;     // ew1 += fdt0 * ep,  ep = cos + i sin
        b1r += f3 * cs; b1i += f3 * -sn;    // bw1 += fdt3 * em,  em = cos - i sin
        e2r += f1 * cs; e2i += f1 * sn;
        b2r += f2 * cs; b2i += f2 * -sn;
      }
      v = pc * ((e1r * b1r - e1i * b1i) - (e2r * b2r - e2i * b2i));
    }
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = kPowerBlock / 2; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) G(partial)[((size_t)pz * nbx + bx) * Nl + l] = red[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(kPowerBlock) void k_power_nodes(
    const float* __restrict__ eb, double* __restrict__ ring, const double* __restrict__ dft,
    const int* __restrict__ plane_k, const double* __restrict__ plane_dzr, double* __restrict__ partial, int N0,
    int N1, unsigned Nz, unsigned Nl, unsigned Nf, unsigned slot, double gamma, double beta, double c0, double pc) {
  power_nodes_body(eb, ring, dft, plane_k, plane_dzr, partial, N0, N1, Nz, Nl, Nf, slot, gamma, beta, c0, pc,
                   blockIdx.x, gridDim.x, blockIdx.y);
}

__device__ inline void power_reduce_body(const double* __restrict__ partial, const int* __restrict__ plane_of,
                                         double* __restrict__ row, unsigned N, unsigned Nl, int blocks, unsigned t) {
  if (t >= N * Nl) return;
  const unsigned k = t / Nl, l = t % Nl;
  const int pz = G(plane_of)[k];
  double s = 0.0;
  if (pz >= 0)
    for (int b = 0; b < blocks; ++b) s += G(partial)[((size_t)pz * blocks + b) * Nl + l];
  G(row)[t] = s;
}

__global__ void k_power_reduce(const double* __restrict__ partial, const int* __restrict__ plane_of,
                               double* __restrict__ row, unsigned N, unsigned Nl, int blocks) {
  power_reduce_body(partial, plane_of, row, N, Nl, blocks, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ __launch_bounds__(kPowerBlock) void kb_power_nodes(const BatchView* __restrict__ V, BatchTiles T) {
  for (int r = 0;; ++r) {
    const int t = r * (int)gridDim.x + blockIdx.x;
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (!v.active || v.step >= (long)v.prows) continue;  // uniform over the workgroup
    power_nodes_body(v.eb, v.ring, v.dft, v.pk, v.pdzr, v.partial, v.N0, v.N1, v.pNz, v.pNl, v.pNf,
                     (unsigned)(v.step % v.pNf), v.pgamma, v.pbeta, v.pc0, v.pc, e.y % v.pblocks, v.pblocks,
                     e.y / v.pblocks);
  }
}

__global__ __launch_bounds__(64) void kb_power_reduce(const BatchView* __restrict__ V, BatchTiles T) {
  for (int r = 0;; ++r) {
    const int t = r * (int)gridDim.x + blockIdx.x;
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (!v.active || v.step >= (long)v.prows) continue;
    power_reduce_body(v.partial, v.plane_of, v.history + (size_t)v.step * v.pN * v.pNl, v.pN, v.pNl, v.pblocks,
                      (unsigned)e.y * 64 + threadIdx.x);
  }
}

}  // namespace

hipError_t launch_power(Domain* d, unsigned n_time, unsigned row) {
  const mithra_grid& g = d->g;
  double* out = d->history + (size_t)row * d->pN * d->pNl;
  const unsigned slot = n_time % d->pNf;
  if (d->pNz > 0 && g.N0 > 4 && g.N1 > 4) {
    hipLaunchKernelGGL(k_power_nodes, dim3(d->pblocks, d->pNz), dim3(kPowerBlock), 0, d->stream, d->eb, d->ring, d->dft,
                       d->pk, d->pdzr, d->partial, g.N0, g.N1, d->pNz, d->pNl, d->pNf, slot, d->pgamma, d->pbeta,
                       d->pc0, d->pc);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const unsigned n = d->pN * d->pNl;
  hipLaunchKernelGGL(k_power_reduce, dim3((n + 63) / 64), dim3(64), 0, d->stream, d->partial,
                     (const int*)(d->pk + d->pNz), out, d->pN, d->pNl, d->pblocks);
  return hipGetLastError();
}


hipError_t launch_batch_power(BatchView* V, BatchTiles nodes, BatchTiles reduce, hipStream_t s) {
  if (nodes.n) hipLaunchKernelGGL(kb_power_nodes, dim3(batch_grid(nodes.n, 4)), dim3(kPowerBlock), 0, s, V, nodes);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (reduce.n) hipLaunchKernelGGL(kb_power_reduce, dim3(batch_grid(reduce.n, 8)), dim3(64), 0, s, V, reduce);
  return hipGetLastError();
}

}  // namespace mithra
