// Particle kernels: the bunch sub-steps (undulator field, E/B gather, Boris push, drift)
// and the charge-conserving zigzag current deposition.
#include <cmath>

#include "common.h"

namespace mithra {

namespace {

struct PushConst {
  // mesh
  double xmin, xmax, ymin, ymax, zmin, zp0, zp1, dx, dy, dz, Lz;
  int N0, N1, N2;
  // frame and Boris constants
  double gamma, beta, c0, dt_shift, r1, r2, dtb;
  // time of the first sub-step and the bunch time step added after each sub-step
  double tb0, dtb_time;
  int n_sub, n_und, save_rnm;
  double dist0;  // undulator[0].dist_: entry test of the emission flag
};

__device__ inline double pmod(double a, double b) {  // src/stdinclude.cpp:88-93
  double x = fmod(a, b);
  x += (x < 0.0) ? b : 0.0;
  return x;
}

// Solver::undulatorField for static undulators with Solver::staticUndulator
// (src/solver.cpp:1798-1824, src/beam.cc:14-76): adds the boosted-frame fields at r.
__device__ void undulator_field(const PushConst& c, const mithra_undulator* __restrict__ und, double x, double y,
                                double z, double tb, double et[3], double bt[3]) {
  for (int u = 0; u < c.n_und; ++u) {
    const mithra_undulator U = und[u];
    const double b0 = (U.lu != 0.0) ? kEM * c.c0 * 2 * kPI / U.lu * U.k / kEC : 0.0;
    const double ku = (U.lu != 0.0) ? 2 * kPI / U.lu : 0.0;
    const double ct = cos(U.theta), st = sin(U.theta);
    const double lz = c.gamma * (z + c.beta * c.c0 * (tb + c.dt_shift)) - U.rb;
    const double ly = x * ct + y * st;
    const double L = U.length * U.lu;
    double d1, bz;
    if (lz >= 0.0 && lz <= L) {
      d1 = b0 * cosh(ku * ly) * sin(ku * lz) * c.gamma;
      bz = b0 * sinh(ku * ly) * cos(ku * lz);
    } else {
      const double s = (lz < 0.0) ? lz : lz - L;     // distance into the fringe
      double sz = exp(-pow(ku * s, 2) / 2.0);
      if (lz < 0.0 && u != 0) {
        const mithra_undulator V = und[u - 1];
        const double r0 = V.rb + V.length * V.lu - U.rb;
        if (lz < r0 || r0 == 0.0) sz = 0.0;
        else sz *= 0.35875 + 0.48829 * cos(kPI * lz / r0) + 0.14128 * cos(2.0 * kPI * lz / r0) +
                   0.01168 * cos(3.0 * kPI * lz / r0);
      } else if (lz > L && u + 1 != c.n_und) {
        const mithra_undulator V = und[u + 1];
        const double r0 = V.rb - U.rb - L;
        if (s > r0 || r0 == 0.0) sz = 0.0;
        else sz *= 0.35875 + 0.48829 * cos(kPI * s / r0) + 0.14128 * cos(2.0 * kPI * s / r0) +
                   0.01168 * cos(3.0 * kPI * s / r0);
      }
      d1 = b0 * cosh(ku * ly) * sz * ku * s * c.gamma;
      bz = b0 * sinh(ku * ly) * sz;
    }
    bt[0] += d1 * ct;
    bt[1] += d1 * st;
    bt[2] += bz;
    d1 *= c.c0 * c.beta;
    et[1] += d1 * ct;
    et[0] += -d1 * st;
    et[2] += 0.0;
  }
}

// n_sub bunch sub-steps of Solver::bunchUpdate (src/solver.cpp:1424-1549) per particle.  The
// E/B grid does not change between sub-steps, so a particle is read from HBM once per field
// step and advanced n_sub times in registers; the time of each sub-step is accumulated exactly
// as Solver::solve() accumulates timeBunch_ (src/solver.cpp:1316-1321).
__global__ __launch_bounds__(256) void k_push(double* __restrict__ part, long cap, long n,
                                              const float4* __restrict__ eb, const mithra_undulator* __restrict__ und,
                                              PushConst c, unsigned long long* __restrict__ counters) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double* X = part;
  double x = X[PX * cap + p], y = X[PY * cap + p], z = X[PZ * cap + p];
  double gx = X[PGX * cap + p], gy = X[PGY * cap + p], gz = X[PGZ * cap + p];
  double e = X[PE * cap + p];
  if (c.save_rnm) {  // rnm = rnp (src/solver.cpp:1311-1312)
    X[PXM * cap + p] = x;
    X[PYM * cap + p] = y;
    X[PZM * cap + p] = z;
  }
  const long N1 = c.N1, N1N0 = (long)c.N1 * c.N0;
  unsigned nt = 0, ncross = 0;
  double tb = c.tb0;
  for (int s = 0; s < c.n_sub; ++s, tb += c.dtb_time) {
    double zr = pmod(z - c.zmin, c.Lz) + c.zmin;
    if (!((zr >= c.zp0) && (zr < c.zp1))) continue;
    const bool b1x = (x < c.xmax - c.dx && x > c.xmin + c.dx);
    const bool b1y = (y < c.ymax - c.dy && y > c.ymin + c.dy);
    const bool b1z = (z < c.zp1 && z >= c.zp0);
    double bt[3] = {0.0, 0.0, 0.0}, et[3] = {0.0, 0.0, 0.0};
    undulator_field(c, und, x, y, z, tb, et, bt);
    if (e == 1.0) {
      if (b1x && b1y && b1z) {
        double ip, jp, kp;
        const double dxr = modf((x - c.xmin) / c.dx, &ip);
        const double dyr = modf((y - c.ymin) / c.dy, &jp);
        const double dzr = modf((z - c.zmin) / c.dz, &kp);
        const int ci = (int)ip, cj = (int)jp, ck = (int)kp;
        // The domain test above keeps the cell inside the mesh in exact arithmetic; a quotient
        // rounded up to the last node would index one plane past the mesh (out of bounds in the
        // reference too), so such a particle skips the grid field instead.
        if (ci < 0 || ci > c.N0 - 2 || cj < 0 || cj > c.N1 - 2 || ck < 0 || ck > c.N2 - 2) goto boris;
        const long m = (long)ck * N1N0 + (long)ci * N1 + cj;
        const long off[8] = {0, N1, 1, N1 + 1, N1N0, N1N0 + N1, N1N0 + 1, N1N0 + N1 + 1};
        const double w[8] = {(1.0 - dxr) * (1.0 - dyr) * (1.0 - dzr), dxr * (1.0 - dyr) * (1.0 - dzr),
                             (1.0 - dxr) * dyr * (1.0 - dzr),         dxr * dyr * (1.0 - dzr),
                             (1.0 - dxr) * (1.0 - dyr) * dzr,         dxr * (1.0 - dyr) * dzr,
                             (1.0 - dxr) * dyr * dzr,                 dxr * dyr * dzr};
        float4 q0[8], q1[8];
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          q0[t] = eb[2 * (m + off[t])];
          q1[t] = eb[2 * (m + off[t]) + 1];
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          et[0] += w[t] * (double)q0[t].x;
          et[1] += w[t] * (double)q0[t].y;
          et[2] += w[t] * (double)q0[t].z;
        }
#pragma unroll
        for (int t = 0; t < 8; ++t) {
          bt[0] += w[t] * (double)q0[t].w;
          bt[1] += w[t] * (double)q1[t].x;
          bt[2] += w[t] * (double)q1[t].y;
        }
      } else if (!b1x && !b1y && b1z) {
        ++nt;
      }
    } else if (c.n_und > 0) {
      const double lz = c.gamma * (z + c.beta * c.c0 * (tb + c.dt_shift));
      e = (lz > -c.dist0) ? 1.0 : 0.0;
    } else {
      e = 1.0;
    }
  boris:
    // Boris rotation and drift (src/solver.cpp:1519-1541)
    double mx = gx, my = gy, mz = gz;
    mx += c.r1 * et[0];
    my += c.r1 * et[1];
    mz += c.r1 * et[2];
    double px = my * bt[2] - mz * bt[1];
    double py = mz * bt[0] - mx * bt[2];
    double pz = mx * bt[1] - my * bt[0];
    const double d1 = sqrt(1.0 + (mx * mx + my * my + mz * mz));
    double f = c.r2 / d1;
    px = f * px; py = f * py; pz = f * pz;
    px += mx; py += my; pz += mz;
    double lx = py * bt[2] - pz * bt[1];
    double ly = pz * bt[0] - px * bt[2];
    double lzv = px * bt[1] - py * bt[0];
    f = 2.0 / (d1 / c.r2 + c.r2 / d1 * (bt[0] * bt[0] + bt[1] * bt[1] + bt[2] * bt[2]));
    lx = f * lx; ly = f * ly; lzv = f * lzv;
    lx += mx; ly += my; lzv += mz;
    gx = lx; gy = ly; gz = lzv;
    gx += c.r1 * et[0];
    gy += c.r1 * et[1];
    gz += c.r1 * et[2];
    const double sdr = c.dtb / sqrt(1.0 + (gx * gx + gy * gy + gz * gz));
    const double drx = sdr * gx, dry = sdr * gy, drz = sdr * gz;
    x += drx;
    y += dry;
    z += drz;
    zr += drz;
    if (zr < c.zp0 || zr >= c.zp1) ++ncross;
  }
  X[PX * cap + p] = x;
  X[PY * cap + p] = y;
  X[PZ * cap + p] = z;
  X[PGX * cap + p] = gx;
  X[PGY * cap + p] = gy;
  X[PGZ * cap + p] = gz;
  X[PE * cap + p] = e;
  if (nt) atomicAdd(&counters[0], (unsigned long long)nt);
  if (ncross) atomicAdd(&counters[1], (unsigned long long)ncross);
}

__global__ void k_bunch_begin(double* __restrict__ part, long cap, long n) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  part[PXM * cap + p] = part[PX * cap + p];
  part[PYM * cap + p] = part[PY * cap + p];
  part[PZM * cap + p] = part[PZ * cap + p];
}

struct DepositConst {
  double xmin, xmax, ymin, ymax, zmin, zp0, zp1, dx, dy, dz;
  int N0, N1, N2, sc;
};

// A deposit at cell (i, j, k) writes nodes i..i+1, j..j+1, k..k+1.
__device__ inline bool cell_ok(const DepositConst& c, int i, int j, int k) {
  return i >= 0 && i <= c.N0 - 2 && j >= 0 && j <= c.N1 - 2 && k >= 0 && k <= c.N2 - 2;
}

__device__ inline bool inside(const DepositConst& c, double x, double y, double z) {
  return x < c.xmax - c.dx && x > c.xmin + c.dx && y < c.ymax - c.dy && y > c.ymin + c.dy && z < c.zp1 && z >= c.zp0;
}

template <typename R>
__device__ inline void add(R* p, double v) {
  unsafeAtomicAdd(p, (R)v);
}

// One particle's half-segment deposit (src/fdtd.cpp:94-137 / :140-183): weights at the
// midpoint of (end, relay point), J_x at the four y/z node pairs split equally over x, etc.
template <typename R>
__device__ inline void deposit_half(const DepositConst& c, R* jx, R* jy, R* jz, long m, double q, const double a[3],
                                    const double r[3], const double jc[3]) {
  double cc;
  const double fx = modf((0.5 * (a[0] + r[0]) - c.xmin) / c.dx, &cc);
  const double fy = modf((0.5 * (a[1] + r[1]) - c.ymin) / c.dy, &cc);
  const double fz = modf((0.5 * (a[2] + r[2]) - c.zmin) / c.dz, &cc);
  const double x1 = 1.0 - fx, x2 = fx, y1 = 1.0 - fy, y2 = fy, z1 = 1.0 - fz, z2 = fz;
  const long N1 = c.N1, N1N0 = (long)c.N1 * c.N0;
  const long o[8] = {0, N1, 1, N1 + 1, N1N0, N1N0 + N1, N1N0 + 1, N1N0 + N1 + 1};
  const double wxa[8] = {y1, y1, y2, y2, y1, y1, y2, y2}, wxb[8] = {z1, z1, z1, z1, z2, z2, z2, z2};
  const double wya[8] = {x1, x2, x1, x2, x1, x2, x1, x2}, wyb[8] = {z1, z1, z1, z1, z2, z2, z2, z2};
  const double wza[8] = {x1, x2, x1, x2, x1, x2, x1, x2}, wzb[8] = {y1, y1, y2, y2, y1, y1, y2, y2};
#pragma unroll
  for (int t = 0; t < 8; ++t) add(jx + m + o[t], q * 0.5 * wxa[t] * wxb[t] * jc[0]);
#pragma unroll
  for (int t = 0; t < 8; ++t) add(jy + m + o[t], q * 0.5 * wya[t] * wyb[t] * jc[1]);
#pragma unroll
  for (int t = 0; t < 8; ++t) add(jz + m + o[t], q * 0.5 * wza[t] * wzb[t] * jc[2]);
}

// FdTd::currentUpdate / FdTdSC::currentUpdate (src/fdtd.cpp:38-185, src/fdtdSC.cpp:38-210):
// Umeda zigzag relay point, two half-segment deposits and (with space charge) rho at rnp.
template <typename R>
__global__ __launch_bounds__(256) void k_deposit(const double* __restrict__ part, long cap, long n, FieldSet J,
                                                 DepositConst c) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double rp[3] = {part[PX * cap + p], part[PY * cap + p], part[PZ * cap + p]};
  const double rm[3] = {part[PXM * cap + p], part[PYM * cap + p], part[PZM * cap + p]};
  const bool bp = inside(c, rp[0], rp[1], rp[2]);
  const bool bm = inside(c, rm[0], rm[1], rm[2]);
  if (!(bp || bm)) return;
  const double q = part[PQ * cap + p];
  const int ip = (int)floor((rp[0] - c.xmin) / c.dx), jp = (int)floor((rp[1] - c.ymin) / c.dy),
            kp = (int)floor((rp[2] - c.zmin) / c.dz);
  const int im = (int)floor((rm[0] - c.xmin) / c.dx), jm = (int)floor((rm[1] - c.ymin) / c.dy),
            km = (int)floor((rm[2] - c.zmin) / c.dz);
  double r[3];
  r[0] = fmin(min(im, ip) * c.dx + c.dx + c.xmin, fmax(max(im, ip) * c.dx + c.xmin, 0.5 * (rm[0] + rp[0])));
  r[1] = fmin(min(jm, jp) * c.dy + c.dy + c.ymin, fmax(max(jm, jp) * c.dy + c.ymin, 0.5 * (rm[1] + rp[1])));
  r[2] = fmin(min(km, kp) * c.dz + c.dz + c.zmin, fmax(max(km, kp) * c.dz + c.zmin, 0.5 * (rm[2] + rp[2])));
  const double jcm[3] = {r[0] - rm[0], r[1] - rm[1], r[2] - rm[2]};
  const double jcp[3] = {rp[0] - r[0], rp[1] - r[1], rp[2] - r[2]};
  R* jx = (R*)J.c[0];
  R* jy = (R*)J.c[1];
  R* jz = (R*)J.c[2];
  const long N1 = c.N1, N1N0 = (long)c.N1 * c.N0;
  if (bp && cell_ok(c, ip, jp, kp)) {
    const long m = N1N0 * kp + N1 * ip + jp;
    deposit_half<R>(c, jx, jy, jz, m, q, rp, r, jcp);
    if (c.sc) {
      R* rho = (R*)J.c[3];
      double cc;
      const double fx = modf((rp[0] - c.xmin) / c.dx, &cc);
      const double fy = modf((rp[1] - c.ymin) / c.dy, &cc);
      const double fz = modf((rp[2] - c.zmin) / c.dz, &cc);
      const double x1 = 1.0 - fx, x2 = fx, y1 = 1.0 - fy, y2 = fy, z1 = 1.0 - fz, z2 = fz;
      add(rho + m, q * x1 * y1 * z1);
      add(rho + m + N1, q * x2 * y1 * z1);
      add(rho + m + 1, q * x1 * y2 * z1);
      add(rho + m + N1 + 1, q * x2 * y2 * z1);
      add(rho + m + N1N0, q * x1 * y1 * z2);
      add(rho + m + N1N0 + N1, q * x2 * y1 * z2);
      add(rho + m + N1N0 + 1, q * x1 * y2 * z2);
      add(rho + m + N1N0 + N1 + 1, q * x2 * y2 * z2);
    }
  }
  if (bm && cell_ok(c, im, jm, km)) {
    const long m = N1N0 * km + N1 * im + jm;
    deposit_half<R>(c, jx, jy, jz, m, q, rm, r, jcm);
  }
}

// Bunching factor partial sums: per workgroup sum of q cos(k z), q sin(k z), q over its particles.
__global__ __launch_bounds__(256) void k_bunching(const double* __restrict__ part, long cap, long n, double k,
                                                  double* __restrict__ partial) {
  __shared__ double s[3][256];
  double c = 0.0, sn = 0.0, q = 0.0;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    const double w = part[PQ * cap + p], z = part[PZ * cap + p];
    double sv, cv;
    sincos(k * z, &sv, &cv);
    c += w * cv;
    sn += w * sv;
    q += w;
  }
  s[0][threadIdx.x] = c;
  s[1][threadIdx.x] = sn;
  s[2][threadIdx.x] = q;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int r = 0; r < 3; ++r) s[r][threadIdx.x] += s[r][threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int r = 0; r < 3; ++r) partial[3 * blockIdx.x + r] = s[r][0];
}

}  // namespace

hipError_t launch_bunching(Domain* d, double k, double* partial, int blocks) {
  if (d->np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bunching, dim3(blocks), dim3(256), 0, d->stream, d->part, d->cap, d->np, k, partial);
  return hipGetLastError();
}

hipError_t launch_bunch_begin(Domain* d) {
  if (d->np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bunch_begin, dim3((unsigned)((d->np + 255) / 256)), dim3(256), 0, d->stream, d->part, d->cap,
                     d->np);
  return hipGetLastError();
}

hipError_t launch_push(Domain* d, const mithra_bunch_params* p, const mithra_undulator* und, int n_und,
                       double time_bunch, double bunch_time_step, int n_sub, int save_rnm) {
  if (d->np == 0 || n_sub <= 0) return hipSuccess;
  const mithra_grid& g = d->g;
  PushConst c;
  c.xmin = g.xmin; c.xmax = g.xmax; c.ymin = g.ymin; c.ymax = g.ymax; c.zmin = g.zmin;
  c.zp0 = g.zp0; c.zp1 = g.zp1; c.dx = g.dx; c.dy = g.dy; c.dz = g.dz; c.Lz = g.Lz;
  c.N0 = g.N0; c.N1 = g.N1; c.N2 = g.N2;
  c.gamma = p->gamma; c.beta = p->beta; c.c0 = p->c0; c.dt_shift = p->dt_shift;
  c.r1 = p->r1; c.r2 = p->r2; c.dtb = p->dtb;
  c.tb0 = time_bunch; c.dtb_time = bunch_time_step;
  c.n_sub = n_sub; c.n_und = n_und; c.save_rnm = save_rnm;
  c.dist0 = n_und > 0 ? und[0].dist : 0.0;
  hipLaunchKernelGGL(k_push, dim3((unsigned)((d->np + 255) / 256)), dim3(256), 0, d->stream, d->part, d->cap, d->np,
                     d->eb, (const mithra_undulator*)und, c, d->counters);
  return hipGetLastError();
}

hipError_t launch_deposit(Domain* d) {
  if (d->np == 0) return hipSuccess;
  const mithra_grid& g = d->g;
  DepositConst c;
  c.xmin = g.xmin; c.xmax = g.xmax; c.ymin = g.ymin; c.ymax = g.ymax; c.zmin = g.zmin;
  c.zp0 = g.zp0; c.zp1 = g.zp1; c.dx = g.dx; c.dy = g.dy; c.dz = g.dz;
  c.N0 = g.N0; c.N1 = g.N1; c.N2 = g.N2; c.sc = g.space_charge;
  const dim3 grid((unsigned)((d->np + 255) / 256));
  if (d->precision == MITHRA_PRECISION_FP64)
    hipLaunchKernelGGL((k_deposit<double>), grid, dim3(256), 0, d->stream, d->part, d->cap, d->np, d->lvl[d->i_np1], c);
  else
    hipLaunchKernelGGL((k_deposit<float>), grid, dim3(256), 0, d->stream, d->part, d->cap, d->np, d->lvl[d->i_np1], c);
  return hipGetLastError();
}

}  // namespace mithra
