// Particle kernels: the bunch sub-steps (undulator field, E/B gather, Boris push, drift)
// and the charge-conserving zigzag current deposition.
#include <cmath>

#include "common.h"

namespace mithra {

namespace {


__device__ inline double pmod(double a, double b) {  // src/stdinclude.cpp:88-93
  if (a >= 0.0 && a < b) return a;  // fmod(a, b) == a exactly here, and a is not negative
  double x = fmod(a, b);
  x += (x < 0.0) ? b : 0.0;
  return x;
}

// ---- laser beams of optical undulators and external fields (src/beam.cc:80-496) ------------

__device__ inline void set3(double v[3], double a, const double d[3]) {  // FieldVector::mv
  v[0] = a * d[0];
  v[1] = a * d[1];
  v[2] = a * d[2];
}
__device__ inline void add3(double v[3], double a, const double d[3]) {  // FieldVector::pmv
  v[0] += a * d[0];
  v[1] += a * d[1];
  v[2] += a * d[2];
}

// Lab-frame E (eT) and B (bT) of beam s at lab point rl and lab time t0: the propagation
// distance z = (rl - position) . direction, the delayed times tl = t0 - z / c0 and, for the
// standing waves, tlm = t0 + z / c0 (Solver::undulatorField / externalField,
// src/solver.cpp:1829-1838, :1893-1904), then the type's fields in the reference's operation
// order (Solver::planeWave ... standingSuperGaussianBeam).
__device__ void beam_fields(const BeamDev& s, const double rl[3], double t0, double c0, double eT[3], double bT[3]) {
  const double rv[3] = {rl[0] - s.position[0], rl[1] - s.position[1], rl[2] - s.position[2]};
  const double z = rv[0] * s.direction[0] + rv[1] * s.direction[1] + rv[2] * s.direction[2];
  const double tl = t0 - z / c0, tlm = t0 + z / c0;
  const double p0 = 0.0;
  for (int c = 0; c < 3; ++c) eT[c] = bT[c] = 0.0;
  const int type = s.beam_type;
  const bool standing = type >= MITHRA_SEED_STANDING_PLANEWAVE;
  if (type == MITHRA_SEED_PLANEWAVE || type == MITHRA_SEED_PLANEWAVE_TRUNCATED ||
      type == MITHRA_SEED_STANDING_PLANEWAVE || type == MITHRA_SEED_STANDING_PLANEWAVE_TRUNCATED) {
    const bool truncated = type == MITHRA_SEED_PLANEWAVE_TRUNCATED || type == MITHRA_SEED_STANDING_PLANEWAVE_TRUNCATED;
    double te, tb;  // signal of E and of B
    if (standing) {
      if (truncated) {  // the truncation test comes first for the standing wave (src/beam.cc:312-322)
        const double x = rv[0] * s.polarization[0] + rv[1] * s.polarization[1] + rv[2] * s.polarization[2];
        const double y = rv[0] * s.yv[0] + rv[1] * s.yv[1] + rv[2] * s.yv[2];
        if (pow(x / s.radius[0], 2) + pow(y / s.radius[1], 2) > 1.0) return;
      }
      const double ts = signal(s, tl, p0), tsm = signal(s, tlm, p0);
      te = ts - tsm;
      tb = ts + tsm;
      if (fabs(te) < 1.0e-6 && fabs(tb) < 1.0e-6) return;
    } else {
      const double ts = signal(s, tl, p0);
      if (fabs(ts) < 1.0e-6) return;
      if (truncated) {
        const double x = rv[0] * s.polarization[0] + rv[1] * s.polarization[1] + rv[2] * s.polarization[2];
        const double y = rv[0] * s.yv[0] + rv[1] * s.yv[1] + rv[2] * s.yv[2];
        if (pow(x / s.radius[0], 2) + pow(y / s.radius[1], 2) > 1.0) return;
      }
      te = tb = ts;
    }
    set3(eT, s.amplitude * te, s.polarization);
    set3(bT, s.amplitude * tb / c0, s.yv);
    return;
  }
  // Gaussian and super-Gaussian beams
  const double x = rv[0] * s.polarization[0] + rv[1] * s.polarization[1] + rv[2] * s.polarization[2];
  const double y = rv[0] * s.yv[0] + rv[1] * s.yv[1] + rv[2] * s.yv[2];
  const double wrp = sqrt(1.0 + z * z / (s.zR[0] * s.zR[0]));
  const double wrs = sqrt(1.0 + z * z / (s.zR[1] * s.zR[1]));
  const bool super = type == MITHRA_SEED_SUPERGAUSSIANBEAM || type == MITHRA_SEED_STANDING_SUPERGAUSSIANBEAM;
  if (super) {
    if ((fabs(x) - s.order[0] * s.radius[0]) > 4.0 * s.radius[0] * wrp ||
        (fabs(y) - s.order[1] * s.radius[1]) > 4.0 * s.radius[1] * wrs)
      return;
  } else if (fabs(x / wrp) > 4.0 * s.radius[0] || fabs(y / wrs) > 4.0 * s.radius[1]) {
    return;
  }
  {
    const double ts = signal(s, tl, p0);
    if (standing) {
      const double tsm = signal(s, tlm, p0);
      if (fabs(ts - tsm) < 1.0e-6 && fabs(ts + tsm) < 1.0e-6) return;
    } else if (fabs(ts) < 1.0e-6) {
      return;
    }
  }
  const double atanP = atan(z / s.zR[0]);
  const double atanS = atan(z / s.zR[1]);
  double ex[3] = {0.0, 0.0, 0.0}, by[3] = {0.0, 0.0, 0.0}, ez[3] = {0.0, 0.0, 0.0}, bz[3] = {0.0, 0.0, 0.0};
  if (!super) {
    const double x0 = x / wrp, y0 = y / wrs;
    const double q0 = 0.5 * (atanP + atanS) - kPI * z / s.l * (pow(x0 / s.zR[0], 2) + pow(y0 / s.zR[1], 2));
    double t = exp(-pow(x0 / s.radius[0], 2) - pow(y0 / s.radius[1], 2)) / sqrt(wrs * wrp);
    t *= s.amplitude;
    if (!standing) {  // Solver::gaussianBeam (src/beam.cc:142-197)
      double ts = signal(s, tl, q0 - kPI / 2.0);
      set3(ex, t * ts, s.polarization);
      set3(by, t * ts / c0, s.yv);
      ts = signal(s, tl, q0 + atanP);
      set3(ez, t * (-x0 / s.zR[0]) * ts, s.direction);
      ts = signal(s, tl, q0 + atanS);
      set3(bz, t * (-y0 / s.zR[1]) / c0 * ts, s.direction);
    } else {  // Solver::standingGaussianBeam (src/beam.cc:331-399)
      double ts = signal(s, tl, q0 - kPI / 2.0), tsm = signal(s, tlm, q0 - kPI / 2.0);
      set3(ex, t * (ts - tsm), s.polarization);
      set3(by, t / c0 * (ts + tsm), s.yv);
      double p1 = q0 + atanP;
      ts = signal(s, tl, p1);
      tsm = signal(s, tlm, -p1);
      set3(ez, t * (-x0 / s.zR[0]) * (ts - tsm), s.direction);
      p1 = q0 + atanS;
      ts = signal(s, tl, p1);
      tsm = signal(s, tlm, -p1);
      set3(bz, t * (-y0 / s.zR[1]) / c0 * (ts - tsm), s.direction);
    }
  } else {  // Solver::superGaussianBeam / standingSuperGaussianBeam (src/beam.cc:201-275, :403-494)
    const double af = s.amplitude / sqrt(wrs * wrp);
    for (int i = -s.order[0]; i <= s.order[0]; i++)
      for (int j = -s.order[1]; j <= s.order[1]; j++) {
        const double x0 = (x - i * s.radius[0]) / wrp;
        const double y0 = (y - j * s.radius[1]) / wrs;
        if (fabs(x0) > 4.0 * s.radius[0] || fabs(y0) > 4.0 * s.radius[1]) continue;
        const double q0 = 0.5 * (atanP + atanS) - kPI * z / s.l * (pow(x0 / s.zR[0], 2) + pow(y / s.zR[1], 2));
        const double t = af * exp(-pow(x0 / s.radius[0], 2) - pow(y0 / s.radius[1], 2));
        if (!standing) {
          double ts = signal(s, tl, q0 - kPI / 2.0);
          add3(ex, t * ts, s.polarization);
          add3(by, t * ts / c0, s.yv);
          ts = signal(s, tl, q0 + atanP);
          add3(ez, t * (-x0 / s.zR[0]) * ts, s.direction);
          ts = signal(s, tl, q0 + atanS);
          add3(bz, t * (-y0 / s.zR[1]) * ts / c0, s.direction);
        } else {
          double ts = signal(s, tl, q0 - kPI / 2.0), tsm = signal(s, tlm, q0 - kPI / 2.0);
          add3(ex, t * (ts - tsm), s.polarization);
          add3(by, t * (ts + tsm) / c0, s.yv);
          double p1 = q0 + atanP;
          ts = signal(s, tl, p1);
          tsm = signal(s, tlm, -p1);
          add3(ez, t * (-x0 / s.zR[0]) * (ts - tsm), s.direction);
          p1 = q0 + atanS;
          ts = signal(s, tl, p1);
          tsm = signal(s, tlm, -p1);
          add3(bz, t * (-y0 / s.zR[1]) * (ts - tsm) / c0, s.direction);
        }
      }
  }
  for (int c = 0; c < 3; ++c) {
    eT[c] = ex[c] + ez[c];
    bT[c] = by[c] + bz[c];
  }
}

// A beam's lab fields at the particle, transformed into the bunch frame and added
// (src/solver.cpp:1868-1876, :1938-1946).
__device__ inline void add_beam(const PushConst& c, const BeamDev& b, double x, double y, double z, double tb,
                                double et[3], double bt[3]) {
  const double rl[3] = {x, y, c.gamma * (z + c.beta * c.c0 * (tb + c.dt_shift))};
  const double t0 = c.gamma * (tb + c.dt_shift + c.beta / c.c0 * z);
  double eT[3], bT[3];
  beam_fields(b, rl, t0, c.c0, eT, bT);
  bt[0] += c.gamma * (bT[0] + c.beta / c.c0 * eT[1]);
  bt[1] += c.gamma * (bT[1] - c.beta / c.c0 * eT[0]);
  bt[2] += bT[2];
  et[0] += c.gamma * (eT[0] - c.beta * c.c0 * bT[1]);
  et[1] += c.gamma * (eT[1] + c.beta * c.c0 * bT[0]);
  et[2] += eT[2];
}

// Solver::undulatorField (src/solver.cpp:1798-1877) with Solver::staticUndulator (src/beam.cc:14-76):
// adds every undulator's boosted-frame fields at r in table order; BEAMS: some are optical.
template <bool BEAMS, typename UT>
__device__ void undulator_field(const PushConst& c, UT und, const BeamDev* beams, double x, double y, double z,
                                double tb, double et[3], double bt[3]) {
  for (int u = 0; u < c.n_und; ++u) {
    const UndDev U = *(const UndDev*)(und + u);
    if (BEAMS && U.type == MITHRA_UNDULATOR_OPTICAL) {
      add_beam(c, beams[U.beam], x, y, z, tb, et, bt);
      continue;
    }
    const double b0 = U.b0, ku = U.ku, ct = U.ct, st = U.st, L = U.L;
    const double lz = c.gamma * (z + c.beta * c.c0 * (tb + c.dt_shift)) - U.rb;
    const double ly = x * ct + y * st;
    double d1, bz;
    if (lz >= 0.0 && lz <= L) {
      d1 = b0 * cosh(ku * ly) * sin(ku * lz) * c.gamma;
      bz = b0 * sinh(ku * ly) * cos(ku * lz);
    } else {
      const double s = (lz < 0.0) ? lz : lz - L;     // distance into the fringe
      double sz = exp(-pow(ku * s, 2) / 2.0);
      if (lz < 0.0 && u != 0) {
        const UndDev V = *(const UndDev*)(und + (u - 1));
        const double r0 = V.rb + V.L - U.rb;
        if (lz < r0 || r0 == 0.0) sz = 0.0;
        else sz *= 0.35875 + 0.48829 * cos(kPI * lz / r0) + 0.14128 * cos(2.0 * kPI * lz / r0) +
                   0.01168 * cos(3.0 * kPI * lz / r0);
      } else if (lz > L && u + 1 != c.n_und) {
        const UndDev V = *(const UndDev*)(und + (u + 1));
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
template <bool BEAMS>
__device__ inline void push_body(double* __restrict__ part, long cap, long n, const float* __restrict__ eb,
                                 const UndDev* __restrict__ und, const BeamDev* __restrict__ beams, int n_opt,
                                 int n_ext, const PushConst& c, unsigned long long* __restrict__ counters, long p) {
  if (p >= n) return;
  const gptr<double> X = G(part);
  const gptr<const f2> ebg = G(reinterpret_cast<const f2*>(eb));
  const gptr<const UndDev> undg = G(und);
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
    undulator_field<BEAMS>(c, undg, beams, x, y, z, tb, et, bt);
    if (BEAMS)
      for (int q = 0; q < n_ext; ++q) add_beam(c, beams[n_opt + q], x, y, z, tb, et, bt);
    if (e == 1.0) {
      if (b1x && b1y && b1z) {
        double ip, jp, kp;
        const double dxr = modf(div_rn(x - c.xmin, c.dx, c.ix), &ip);
        const double dyr = modf(div_rn(y - c.ymin, c.dy, c.iy), &jp);
        const double dzr = modf(div_rn(z - c.zmin, c.dz, c.iz), &kp);
        const int ci = (int)ip, cj = (int)jp, ck = (int)kp;
        // The domain test above keeps the cell inside the mesh in exact arithmetic; a quotient
        // rounded up to the last node would index one plane past the mesh (out of bounds in the
        // reference too), so such a particle skips the grid field instead.
        if (ci < 0 || ci > c.N0 - 2 || cj < 0 || cj > c.N1 - 2 || ck < 0 || ck > c.N2 - 2) goto boris;
        const long m = (long)ck * N1N0 + (long)ci * N1 + cj;
        // Corner t = (x offset t & 1, y offset (t >> 1) & 1, z offset t >> 2) with weight
        // w_x w_y w_z in the reference's product order; every component sums its corners in
        // the order t = 0..7.  The corners come in two z layers of four, so only one layer's
        // E/B is held in registers at a time.
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          f2 q0[4], q1[4], q2[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const long o = m + (h ? N1N0 : 0) + ((u & 1) ? N1 : 0) + (u >> 1);
            q0[u] = ebg[kEBPairs * o];
            q1[u] = ebg[kEBPairs * o + 1];
            q2[u] = ebg[kEBPairs * o + 2];
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const double w = ((u & 1) ? dxr : 1.0 - dxr) * ((u >> 1) ? dyr : 1.0 - dyr) * (h ? dzr : 1.0 - dzr);
            et[0] += w * (double)q0[u].x;
            et[1] += w * (double)q0[u].y;
            et[2] += w * (double)q1[u].x;
            bt[0] += w * (double)q1[u].y;
            bt[1] += w * (double)q2[u].x;
            bt[2] += w * (double)q2[u].y;
          }
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

template <bool BEAMS>
__global__ __launch_bounds__(256) void k_push(double* __restrict__ part, long cap, long n,
                                              const float* __restrict__ eb, const UndDev* __restrict__ und,
                                              const BeamDev* __restrict__ beams, int n_opt, int n_ext, PushConst c,
                                              unsigned long long* __restrict__ counters) {
  push_body<BEAMS>(part, cap, n, eb, und, beams, n_opt, n_ext, c, counters,
                   (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// z extent of the bunch in cells (clamped to the mesh), for the lazy E/B evaluation: per
// workgroup min / max, then one workgroup reduces the partials into zr[0..1].
constexpr int kZBlocks = 1024;

__global__ __launch_bounds__(256) void k_zcells(const double* __restrict__ part, long cap, long n, double zmin,
                                                double dz, double iz, int kmax_cell, int2* __restrict__ partial) {
  __shared__ int s_lo[256], s_hi[256];
  int lo = 0x7fffffff, hi = -1;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
    const double zz = floor(div_rn(part[PZ * cap + p] - zmin, dz, iz));
    const int k = zz >= 0.0 ? (zz <= (double)kmax_cell ? (int)zz : kmax_cell) : 0;
    lo = min(lo, k);
    hi = max(hi, k);
  }
  s_lo[threadIdx.x] = lo;
  s_hi[threadIdx.x] = hi;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_lo[threadIdx.x] = min(s_lo[threadIdx.x], s_lo[threadIdx.x + h]);
      s_hi[threadIdx.x] = max(s_hi[threadIdx.x], s_hi[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_int2(s_lo[0], s_hi[0]);
}

__global__ __launch_bounds__(256) void k_zcells_final(const int2* __restrict__ partial, int nb, int* __restrict__ zr) {
  __shared__ int s_lo[256], s_hi[256];
  int lo = 0x7fffffff, hi = -1;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    lo = min(lo, partial[b].x);
    hi = max(hi, partial[b].y);
  }
  s_lo[threadIdx.x] = lo;
  s_hi[threadIdx.x] = hi;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_lo[threadIdx.x] = min(s_lo[threadIdx.x], s_lo[threadIdx.x + h]);
      s_hi[threadIdx.x] = max(s_hi[threadIdx.x], s_hi[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    zr[0] = s_lo[0];
    zr[1] = s_hi[0];
  }
}

__global__ void k_bunch_begin(double* __restrict__ part, long cap, long n) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  part[PXM * cap + p] = part[PX * cap + p];
  part[PYM * cap + p] = part[PY * cap + p];
  part[PZM * cap + p] = part[PZ * cap + p];
}


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

// One particle's half-segment weights (src/fdtd.cpp:94-137 / :140-183), taken at the midpoint of
// (end, relay point).  J_x is split equally over the two x nodes of each y/z pair, so each
// component has four distinct values per half segment:
//   v[0..3]  J_x at the (y, z) pairs (1,1) (2,1) (1,2) (2,2), nodes o[2u] and o[2u+1]
//   v[4..7]  J_y at the (x, z) pairs, nodes o[t] with (t & 1) + 2 (t >> 2) = v - 4
//   v[8..11] J_z at the (x, y) pairs, nodes o[t] with t & 3 = v - 8
// Each value is formed as q 0.5 w_a w_b jc, the reference's operation order.
__device__ inline void half_weights(const DepositConst& c, double q, const double a[3], const double r[3],
                                    const double jc[3], double* v) {
  double cc;
  const double fx = modf(div_rn(0.5 * (a[0] + r[0]) - c.xmin, c.dx, c.ix), &cc);
  const double fy = modf(div_rn(0.5 * (a[1] + r[1]) - c.ymin, c.dy, c.iy), &cc);
  const double fz = modf(div_rn(0.5 * (a[2] + r[2]) - c.zmin, c.dz, c.iz), &cc);
  const double x1 = 1.0 - fx, x2 = fx, y1 = 1.0 - fy, y2 = fy, z1 = 1.0 - fz, z2 = fz;
  const double h = q * 0.5;
  v[0] = h * y1 * z1 * jc[0]; v[1] = h * y2 * z1 * jc[0]; v[2] = h * y1 * z2 * jc[0]; v[3] = h * y2 * z2 * jc[0];
  v[4] = h * x1 * z1 * jc[1]; v[5] = h * x2 * z1 * jc[1]; v[6] = h * x1 * z2 * jc[1]; v[7] = h * x2 * z2 * jc[1];
  v[8] = h * x1 * y1 * jc[2]; v[9] = h * x2 * y1 * jc[2]; v[10] = h * x1 * y2 * jc[2]; v[11] = h * x2 * y2 * jc[2];
}

// Charge density weights at rnp (src/fdtdSC.cpp:141-160), node order o[0..7].
__device__ inline void rho_weights(const DepositConst& c, double q, const double rp[3], double* v) {
  double cc;
  const double fx = modf(div_rn(rp[0] - c.xmin, c.dx, c.ix), &cc);
  const double fy = modf(div_rn(rp[1] - c.ymin, c.dy, c.iy), &cc);
  const double fz = modf(div_rn(rp[2] - c.zmin, c.dz, c.iz), &cc);
  const double x1 = 1.0 - fx, x2 = fx, y1 = 1.0 - fy, y2 = fy, z1 = 1.0 - fz, z2 = fz;
  v[0] = q * x1 * y1 * z1; v[1] = q * x2 * y1 * z1; v[2] = q * x1 * y2 * z1; v[3] = q * x2 * y2 * z1;
  v[4] = q * x1 * y1 * z2; v[5] = q * x2 * y1 * z2; v[6] = q * x1 * y2 * z2; v[7] = q * x2 * y2 * z2;
}

// Wavefront-wide segmented sum: lanes holding equal keys in consecutive lanes form a run, and
// the first lane of each run ends with the run's total in v (Hillis-Steele suffix scan bounded
// by the run's last lane).  Must be reached by all 64 lanes of the wavefront.
template <int NV>
__device__ inline void run_sum(long key, double (&v)[NV]) {
  const int lane = __lane_id();
  const long next = __shfl_down(key, 1);
  const unsigned long long tails = __ballot(lane == 63 || next != key);
  const int end = lane + __builtin_ctzll(tails >> lane);  // bit 63 is always set
  for (int off = 1; off < 64; off <<= 1) {
    const bool take = lane + off <= end;
    if (__ballot(take) == 0ull) break;
#pragma unroll
    for (int t = 0; t < NV; ++t) {
      const double o = __shfl_down(v[t], off);
      if (take) v[t] += o;
    }
  }
}

__device__ inline bool run_head(long key) {
  const long prev = __shfl_up(key, 1);
  return __lane_id() == 0 || prev != key;
}

template <typename R>
__device__ inline void scatter_j(R* jx, R* jy, R* jz, long m, long N1, long N1N0, const double* v) {
  const long o[8] = {0, N1, 1, N1 + 1, N1N0, N1N0 + N1, N1N0 + 1, N1N0 + N1 + 1};
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    add(jx + m + o[t], v[t >> 1]);
    add(jy + m + o[t], v[4 + (t & 1) + 2 * (t >> 2)]);
    add(jz + m + o[t], v[8 + (t & 3)]);
  }
}

// FdTd::currentUpdate / FdTdSC::currentUpdate (src/fdtd.cpp:38-185, src/fdtdSC.cpp:38-210):
// Umeda zigzag relay point, two half-segment deposits and (with space charge) rho at rnp.
// A particle whose two halves fall in the same cell adds them before depositing; lanes of a
// wavefront depositing into the same cell (particles are kept sorted by cell) are summed
// across the wavefront, and only the first lane of each run issues the atomic adds.
template <typename R, bool SC>
__global__ __launch_bounds__(256) void k_deposit(const double* __restrict__ part, long cap, long n, FieldSet J,
                                                 DepositConst c) {
  constexpr int NP = SC ? 20 : 12;
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double vp[NP], vm[12];
#pragma unroll
  for (int t = 0; t < NP; ++t) vp[t] = 0.0;
#pragma unroll
  for (int t = 0; t < 12; ++t) vm[t] = 0.0;
  long kp_key = -1, km_key = -1;
  if (p < n) {
    const double rp[3] = {part[PX * cap + p], part[PY * cap + p], part[PZ * cap + p]};
    const double rm[3] = {part[PXM * cap + p], part[PYM * cap + p], part[PZM * cap + p]};
    const bool bp = inside(c, rp[0], rp[1], rp[2]);
    const bool bm = inside(c, rm[0], rm[1], rm[2]);
    if (bp || bm) {
      const double q = part[PQ * cap + p];
      const int ip = (int)floor(div_rn(rp[0] - c.xmin, c.dx, c.ix)), jp = (int)floor(div_rn(rp[1] - c.ymin, c.dy, c.iy)),
                kp = (int)floor(div_rn(rp[2] - c.zmin, c.dz, c.iz));
      const int im = (int)floor(div_rn(rm[0] - c.xmin, c.dx, c.ix)), jm = (int)floor(div_rn(rm[1] - c.ymin, c.dy, c.iy)),
                km = (int)floor(div_rn(rm[2] - c.zmin, c.dz, c.iz));
      double r[3];
      r[0] = fmin(min(im, ip) * c.dx + c.dx + c.xmin, fmax(max(im, ip) * c.dx + c.xmin, 0.5 * (rm[0] + rp[0])));
      r[1] = fmin(min(jm, jp) * c.dy + c.dy + c.ymin, fmax(max(jm, jp) * c.dy + c.ymin, 0.5 * (rm[1] + rp[1])));
      r[2] = fmin(min(km, kp) * c.dz + c.dz + c.zmin, fmax(max(km, kp) * c.dz + c.zmin, 0.5 * (rm[2] + rp[2])));
      const double jcm[3] = {r[0] - rm[0], r[1] - rm[1], r[2] - rm[2]};
      const double jcp[3] = {rp[0] - r[0], rp[1] - r[1], rp[2] - r[2]};
      const long N1 = c.N1, N1N0 = (long)c.N1 * c.N0;
      if (bp && cell_ok(c, ip, jp, kp)) {
        kp_key = N1N0 * kp + N1 * ip + jp;
        half_weights(c, q, rp, r, jcp, vp);
        if (SC) rho_weights(c, q, rp, vp + 12);
      }
      if (bm && cell_ok(c, im, jm, km)) {
        km_key = N1N0 * km + N1 * im + jm;
        half_weights(c, q, rm, r, jcm, vm);
      }
      if (kp_key >= 0 && kp_key == km_key) {
#pragma unroll
        for (int t = 0; t < 12; ++t) vp[t] += vm[t];
        km_key = -1;
      }
    }
  }
  R* jx = (R*)J.c[0];
  R* jy = (R*)J.c[1];
  R* jz = (R*)J.c[2];
  const long N1 = c.N1, N1N0 = (long)c.N1 * c.N0;
  run_sum<NP>(kp_key, vp);
  if (kp_key >= 0 && run_head(kp_key)) {
    scatter_j(jx, jy, jz, kp_key, N1, N1N0, vp);
    if (SC) {
      R* rho = (R*)J.c[3];
      const long o[8] = {0, N1, 1, N1 + 1, N1N0, N1N0 + N1, N1N0 + 1, N1N0 + N1 + 1};
#pragma unroll
      for (int t = 0; t < 8; ++t) add(rho + kp_key + o[t], vp[12 + t]);
    }
  }
  if (__ballot(km_key >= 0) != 0ull) {
    run_sum<12>(km_key, vm);
    if (km_key >= 0 && run_head(km_key)) scatter_j(jx, jy, jz, km_key, N1, N1N0, vm);
  }
}

// Row key of one half of every particle: the (k, i) row k*N0 + i of the cell the half deposits
// from (rnp or rnm, the cell test of k_deposit), or `rows` when that half deposits nothing.
// With span = N1 - 1 the key is the cell, (k N0 + i) span + j, and the sentinel rows span.
__global__ __launch_bounds__(256) void k_row_keys(const double* __restrict__ part, long cap, long n, DepositConst c,
                                                  int use_rm, unsigned rows, unsigned span, unsigned* __restrict__ key,
                                                  unsigned* __restrict__ val) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int o = use_rm ? PXM : PX;
  const double x = part[(o + 0) * cap + p], y = part[(o + 1) * cap + p], z = part[(o + 2) * cap + p];
  unsigned k = rows * span;
  if (inside(c, x, y, z)) {
    const int i = (int)floor(div_rn(x - c.xmin, c.dx, c.ix)), j = (int)floor(div_rn(y - c.ymin, c.dy, c.iy)),
              kk = (int)floor(div_rn(z - c.zmin, c.dz, c.iz));
    if (cell_ok(c, i, j, kk))
      k = ((unsigned)kk * (unsigned)c.N0 + (unsigned)i) * span + (span > 1 ? (unsigned)j : 0u);
  }
  key[p] = k;
  val[p] = (unsigned)p;
}

// Adds one half's contributions to the four nodes of its cell in one node plane of a tile
// (rows i0 .. i0+ti-1, all j): the cell's lower node plane (up = false) or its upper one
// (up = true).  Target t = 4 up + tt, tt = 0..3 over (row, column) offsets ((tt & 1), (tt >> 1)),
// with the value layout of half_weights / rho_weights.  The cell's row ic lies in
// i0-1 .. i0+ti-1, so row offset 1 is always inside the tile from below and row offset 0 is out
// only for ic = i0-1.
template <bool RHO>
__device__ inline void add_plane(double* tile, int TI, int N1, bool up, int i0, int ti, int ic, int jc,
                                 const double* v) {
  const int tsz = TI * N1;
  const int r0 = ic - i0;
  const double x0 = up ? v[2] : v[0], x1 = up ? v[3] : v[1];
  const double y0 = up ? v[6] : v[4], y1 = up ? v[7] : v[5];
  if (r0 >= 0) {
    double* at = tile + r0 * N1 + jc;
    atomicAdd(at, x0);
    atomicAdd(at + tsz, y0);
    atomicAdd(at + 2 * tsz, v[8]);
    atomicAdd(at + 1, x1);
    atomicAdd(at + 1 + tsz, y0);
    atomicAdd(at + 1 + 2 * tsz, v[10]);
    if (RHO) {
      atomicAdd(at + 3 * tsz, up ? v[16] : v[12]);
      atomicAdd(at + 1 + 3 * tsz, up ? v[18] : v[14]);
    }
  }
  if (r0 + 1 < ti) {
    double* at = tile + (r0 + 1) * N1 + jc;
    atomicAdd(at, x0);
    atomicAdd(at + tsz, y1);
    atomicAdd(at + 2 * tsz, v[9]);
    atomicAdd(at + 1, x1);
    atomicAdd(at + 1 + tsz, y1);
    atomicAdd(at + 1 + 2 * tsz, v[11]);
    if (RHO) {
      atomicAdd(at + 3 * tsz, up ? v[17] : v[13]);
      atomicAdd(at + 1 + 3 * tsz, up ? v[19] : v[15]);
    }
  }
}

// One listed particle half of the row-tiled deposit: positions and charge as loaded.
struct HalfItem {
  double rp[3], rm[3], q;
  int half, kk;
  bool ok;
};

template <typename PT>
__device__ inline void load_item(HalfItem& it, PT part, long cap, long p) {
  it.rp[0] = part[PX * cap + p];
  it.rp[1] = part[PY * cap + p];
  it.rp[2] = part[PZ * cap + p];
  it.rm[0] = part[PXM * cap + p];
  it.rm[1] = part[PYM * cap + p];
  it.rm[2] = part[PZM * cap + p];
  it.q = part[PQ * cap + p];
}

// Deposits one listed half into the two-plane tile ring of the marching deposit: its lower
// node plane kc when kc >= kb, its upper one kc+1 when kc+1 < ke (slot = plane & 1, `psz`
// doubles per slot).
template <bool SC>
__device__ inline void deposit_item(const HalfItem& it, const DepositConst& c, double* tile, int psz, int TI, int kb,
                                    int ke, int i0, int ti, int lo, int hi) {
  const double* rp = it.rp;
  const double* rm = it.rm;
  const int ip = (int)floor(div_rn(rp[0] - c.xmin, c.dx, c.ix)), jp = (int)floor(div_rn(rp[1] - c.ymin, c.dy, c.iy)),
            kp = (int)floor(div_rn(rp[2] - c.zmin, c.dz, c.iz));
  const int im = (int)floor(div_rn(rm[0] - c.xmin, c.dx, c.ix)), jm = (int)floor(div_rn(rm[1] - c.ymin, c.dy, c.iy)),
            km = (int)floor(div_rn(rm[2] - c.zmin, c.dz, c.iz));
  // The order lists this half under row (kk, lo..hi); skip it should its cell disagree.
  const int ic = it.half ? im : ip, jc = it.half ? jm : jp, kc = it.half ? km : kp;
  if (kc != it.kk || ic < lo || ic > hi || jc < 0 || jc > c.N1 - 2) return;
  double r[3];
  r[0] = fmin(min(im, ip) * c.dx + c.dx + c.xmin, fmax(max(im, ip) * c.dx + c.xmin, 0.5 * (rm[0] + rp[0])));
  r[1] = fmin(min(jm, jp) * c.dy + c.dy + c.ymin, fmax(max(jm, jp) * c.dy + c.ymin, 0.5 * (rm[1] + rp[1])));
  r[2] = fmin(min(km, kp) * c.dz + c.dz + c.zmin, fmax(max(km, kp) * c.dz + c.zmin, 0.5 * (rm[2] + rp[2])));
  double v[SC ? 20 : 12];
  const bool rho = SC && it.half == 0;
  if (it.half == 0) {
    const double jc3[3] = {rp[0] - r[0], rp[1] - r[1], rp[2] - r[2]};
    half_weights(c, it.q, rp, r, jc3, v);
    if (SC) rho_weights(c, it.q, rp, v + 12);
  } else {
    const double jc3[3] = {r[0] - rm[0], r[1] - rm[1], r[2] - rm[2]};
    half_weights(c, it.q, rm, r, jc3, v);
  }
  for (int up = 0; up < 2; ++up) {
    const int k = kc + up;
    if (k < kb || k >= ke) continue;
    double* t = tile + (k & 1) * psz;
    if (rho) add_plane<SC>(t, TI, c.N1, up, i0, ti, ic, jc, v);
    else add_plane<false>(t, TI, c.N1, up, i0, ti, ic, jc, v);
  }
}

// ---- fixed-order deposition ------------------------------------------------------------------
//
// The row-tiled deposit adds into LDS with atomics, so a node's sum is taken in whatever order
// the wavefronts reach it: FP64 runs agree to rounding, not bit for bit.  The fixed-order form
// sums every node in an order that depends on the particles alone.  Halves are listed by cell
// (a stable sort, so the halves of one cell keep their particle-array order).  A first pass
// computes each listed half's weights once (the arithmetic of deposit_item) into its list slot;
// a second pass gives every node one thread, which adds the weights it receives from its eight
// cells in a fixed order: cell planes k = K-1, K, rows i = I-1, I, columns j = J-1, J, in each
// cell the rnp listings, then the rnm listings, in list order.

// Weights of the halves listed in one order (half 0: rnp, with rho under space charge; half
// 1: rnm): entry s of the list gets its 12 (20) weights at w[t cap + s].  Entries past the
// listed cells (the sentinel key) are left alone.
template <bool SC>
__global__ __launch_bounds__(256) void k_half_weights(const double* __restrict__ part, long cap, long n,
                                                      DepositConst c, const unsigned* __restrict__ key,
                                                      const unsigned* __restrict__ idx, unsigned cells, int half,
                                                      double* __restrict__ w) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n || key[s] >= cells) return;
  const long p = idx[s];
  double rp[3], rm[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    rp[a] = part[(PX + a) * cap + p];
    rm[a] = part[(PXM + a) * cap + p];
  }
  const double q = part[PQ * cap + p];
  const int ip = (int)floor(div_rn(rp[0] - c.xmin, c.dx, c.ix)), jp = (int)floor(div_rn(rp[1] - c.ymin, c.dy, c.iy)),
            kp = (int)floor(div_rn(rp[2] - c.zmin, c.dz, c.iz));
  const int im = (int)floor(div_rn(rm[0] - c.xmin, c.dx, c.ix)), jm = (int)floor(div_rn(rm[1] - c.ymin, c.dy, c.iy)),
            km = (int)floor(div_rn(rm[2] - c.zmin, c.dz, c.iz));
  double r[3];
  r[0] = fmin(min(im, ip) * c.dx + c.dx + c.xmin, fmax(max(im, ip) * c.dx + c.xmin, 0.5 * (rm[0] + rp[0])));
  r[1] = fmin(min(jm, jp) * c.dy + c.dy + c.ymin, fmax(max(jm, jp) * c.dy + c.ymin, 0.5 * (rm[1] + rp[1])));
  r[2] = fmin(min(km, kp) * c.dz + c.dz + c.zmin, fmax(max(km, kp) * c.dz + c.zmin, 0.5 * (rm[2] + rp[2])));
  double v[SC ? 20 : 12];
  if (half == 0) {
    const double jc3[3] = {rp[0] - r[0], rp[1] - r[1], rp[2] - r[2]};
    half_weights(c, q, rp, r, jc3, v);
    if (SC) rho_weights(c, q, rp, v + 12);
  } else {
    const double jc3[3] = {r[0] - rm[0], r[1] - rm[1], r[2] - rm[2]};
    half_weights(c, q, rm, r, jc3, v);
  }
  const int nw = (SC && half == 0) ? 20 : 12;
#pragma unroll
  for (int t = 0; t < (SC ? 20 : 12); ++t)
    if (t < nw) w[t * cap + s] = v[t];
}

// One thread per node (I, J, K): J (and rho) = the sum of the weights its eight cells' listed
// halves give it, in the fixed order above; `accumulate` adds the sum to what the buffer holds.
// A cell row no half lists (in either order) adds nothing, which the row starts show at once:
// most rows of the mesh.  Weight of a half of cell (i, j, k) at node offset (a, b, e) =
// (I - i, J - j, K - k), with the value layout of half_weights / rho_weights (as add_plane places
// them): J_x takes v[e ? 2 + b : b], J_y v[e ? 6 + a : 4 + a], J_z v[8 + a + 2 b],
// rho v[12 + 4 e + a + 2 b].
template <typename R, bool SC>
__global__ __launch_bounds__(256) void k_deposit_fixed(FieldSet J, DepositConst c, const unsigned* __restrict__ pcell,
                                                       const unsigned* __restrict__ mcell,
                                                       const unsigned* __restrict__ prow,
                                                       const unsigned* __restrict__ mrow,
                                                       const double* __restrict__ wp, const double* __restrict__ wm,
                                                       long cap, int accumulate) {
  const long N1N0 = (long)c.N1 * c.N0;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= N1N0 * c.N2) return;
  const int K = (int)(t / N1N0);
  const long rem = t - (long)K * N1N0;
  const int I = (int)(rem / c.N1), Jn = (int)(rem - (long)I * c.N1);
  const unsigned span = (unsigned)(c.N1 - 1);
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int k = K - 1; k <= K; ++k) {
    if (k < 0 || k > c.N2 - 2) continue;
    const int e = K - k;
    for (int i = I - 1; i <= I; ++i) {
      if (i < 0 || i > c.N0 - 2) continue;
      const int a = I - i;
      const long row = (long)k * c.N0 + i;
      if (prow[row] == prow[row + 1] && mrow[row] == mrow[row + 1]) continue;
      for (int j = Jn - 1; j <= Jn; ++j) {
        if (j < 0 || j > c.N1 - 2) continue;
        const int b = Jn - j;
        const unsigned cell = ((unsigned)k * (unsigned)c.N0 + (unsigned)i) * span + (unsigned)j;
        for (int h = 0; h < 2; ++h) {
          const unsigned* cs = h ? mcell : pcell;
          const double* w = h ? wm : wp;
          const unsigned s1 = cs[cell + 1];
          for (unsigned s = cs[cell]; s < s1; ++s) {
            acc[0] += w[(e ? 2 + b : b) * cap + s];
            acc[1] += w[(e ? 6 + a : 4 + a) * cap + s];
            acc[2] += w[(8 + a + 2 * b) * cap + s];
            if (SC && h == 0) acc[3] += w[(12 + 4 * e + a + 2 * b) * cap + s];
          }
        }
      }
    }
  }
#pragma unroll
  for (int comp = 0; comp < (SC ? 4 : 3); ++comp) {
    R* out = (R*)component(J, comp);
    out[t] = accumulate ? (R)((double)out[t] + acc[comp]) : (R)acc[comp];
  }
}

// Row-tiled current deposition, marching along z (same arithmetic per particle as k_deposit).
// Workgroup b owns the J (and rho) values of rows i0 .. i0+ti-1, all j, of the node planes
// kb .. ke-1 of its chunk of KC planes.  It walks the cell planes kc = kb-1 .. ke-1 in turn; the halves
// listed under plane kc and rows i0-1 .. i0+ti-1 (two contiguous ranges, one per row order,
// two halves in flight per thread) add their lower nodes to plane kc and their upper nodes to
// plane kc+1 of a two-plane LDS ring, so each half is weighed once per workgroup.  Once plane
// kc is complete it is written out and its slot starts plane kc+2.  Every node of the mesh
// belongs to exactly one workgroup, so no global atomics are needed and the J buffer needs no
// separate reset (`accumulate` = 0 writes the deposit alone, 1 adds it to what the buffer holds).
constexpr int kDepositThreads = 512;
constexpr int kDepositChunk = 4;  // node planes per workgroup (at most)

template <typename R, bool SC, bool SRC = false>
__device__ inline void deposit_rows_body(const double* __restrict__ part, long cap, FieldSet J, const DepositConst& c,
                                         const unsigned* __restrict__ p_row, const unsigned* __restrict__ p_idx,
                                         const unsigned* __restrict__ m_row, const unsigned* __restrict__ m_idx,
                                         int TI, int tilesI, int KC, int accumulate, double src_coef,
                                         double src_rho, int skip_empty, int block) {
  extern __shared__ double tile[];  // [2][ncomp][TI][N1]
  constexpr int ncomp = SC ? 4 : 3;
  const int chunk = block / tilesI;
  const int i0 = (block - chunk * tilesI) * TI;
  const int ti = min(TI, c.N0 - i0);
  const int kb = chunk * KC, ke = min(kb + KC, c.N2);
  const int N1 = c.N1;
  const long N1N0 = (long)N1 * c.N0;
  const int tsz = TI * N1, psz = ncomp * tsz, own = ti * N1;
  const int lo = max(i0 - 1, 0), hi = min(i0 + ti - 1, c.N0 - 2);
  // slot k & 1 := node plane k (the buffer's values with `accumulate` = 1, else zero)
  const bool add = accumulate == 1;
  auto start_plane = [&](int k) {
    double* p = tile + (k & 1) * psz;
    const long base = N1N0 * k + (long)N1 * i0;
    for (int o = threadIdx.x; o < own; o += blockDim.x) {
#pragma unroll
      for (int comp = 0; comp < ncomp; ++comp)
        p[comp * tsz + o] = add ? (double)G((const R*)component(J, comp))[base + o] : 0.0;
    }
  };
  start_plane(kb);
  if (kb + 1 < ke) start_plane(kb + 1);
  __syncthreads();
  unsigned prev_total = 0;  // halves listed under the previous cell plane (this tile's rows)
  for (int kc = kb - 1; kc < ke; ++kc) {
    // the two ranges of cell plane kc: halves 0 (rnp order) and 1 (rnm order), rows lo..hi
    unsigned s0 = 0, n0 = 0, s1 = 0, n1 = 0;
    if (lo <= hi && kc >= 0 && kc <= c.N2 - 2) {
      const long r = (long)kc * c.N0;
      s0 = G(p_row)[r + lo];
      n0 = G(p_row)[r + hi + 1] - s0;
      s1 = G(m_row)[r + lo];
      n1 = G(m_row)[r + hi + 1] - s1;
    }
    const unsigned total = n0 + n1;
    auto locate = [&](unsigned t, HalfItem& it) {
      it.half = t >= n0;
      it.kk = kc;
      return (long)(it.half ? G(m_idx)[s1 + (t - n0)] : G(p_idx)[s0 + t]);
    };
    for (unsigned t = threadIdx.x; t < total; t += 2 * blockDim.x) {
      HalfItem a, b;
      const unsigned tb = t + blockDim.x;
      const long pa = locate(t, a);
      b.ok = tb < total;
      const long pb = b.ok ? locate(tb, b) : pa;
      load_item(a, G(part), cap, pa);
      load_item(b, G(part), cap, pb);
      deposit_item<SC>(a, c, tile, psz, TI, kb, ke, i0, ti, lo, hi);
      if (b.ok) deposit_item<SC>(b, c, tile, psz, TI, kb, ke, i0, ti, lo, hi);
    }
    __syncthreads();
    // Node plane kc takes halves of cell planes kc-1 and kc only.  When neither lists any, its J
    // (and rho) is +0, and in the pipelined form X + a4 J = X bit for bit when a4 (and a5, the
    // coefficient of rho) carries a minus sign (a4 (+0) = -0, and x + -0 = x for every x, -0
    // included), so that plane is left untouched.
    const bool untouched = SRC && skip_empty && prev_total + total == 0;
    prev_total = total;
    if (kc >= kb) {  // node plane kc is complete: write it, and its slot becomes plane kc+2
      double* p = tile + (kc & 1) * psz;
      const long base = N1N0 * kc + (long)N1 * i0;
      const bool next = kc + 2 < ke;
      const bool inner_plane = kc >= 1 && kc <= c.N2 - 2 && !untouched;
      // one node per thread and pass, its components unrolled: no index division per element
      for (int o = threadIdx.x; o < own; o += blockDim.x) {
        bool inner = false;
        if (SRC) {
          const int r = o / N1, jj = o - r * N1;
          inner = inner_plane && i0 + r >= 1 && i0 + r <= c.N0 - 2 && jj >= 1 && jj <= N1 - 2;
        }
#pragma unroll
        for (int comp = 0; comp < ncomp; ++comp) {
          const gptr<R> out = G((R*)component(J, comp));
          if (SRC) {
            // the pipelined step: X + a4 J (and X + a5 rho for phi) at interior nodes, J rounded as
            // the buffer would hold it
            if (inner) {
              const double xo = (double)ld_stream((gptr<const R>)out, base + o);  // read once, written once
              const double coef = comp < 3 ? src_coef : src_rho;
              st_stream(out, base + o, (R)(xo + coef * (double)(R)p[comp * tsz + o]));
            }
          } else {
            out[base + o] = (R)p[comp * tsz + o];
          }
          if (next) p[comp * tsz + o] = add ? (double)G((const R*)component(J, comp))[base + 2 * N1N0 + o] : 0.0;
        }
      }
      __syncthreads();
    }
  }
}

template <typename R, bool SC, bool SRC = false>
__global__ __launch_bounds__(kDepositThreads, SC ? 4 : 8) void k_deposit_rows(
    const double* __restrict__ part, long cap, FieldSet J, DepositConst c, const unsigned* __restrict__ p_row,
    const unsigned* __restrict__ p_idx, const unsigned* __restrict__ m_row, const unsigned* __restrict__ m_idx,
    int TI, int tilesI, int KC, int accumulate, double src_coef, double src_rho, int skip_empty) {
  deposit_rows_body<R, SC, SRC>(part, cap, J, c, p_row, p_idx, m_row, m_idx, TI, tilesI, KC, accumulate, src_coef,
                                src_rho, skip_empty, blockIdx.x);
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

// ---- batched forms: persistent workgroups walking a table of (instance, local block) -------
__device__ inline int batch_entry_p(int r) { return r * (int)gridDim.x + blockIdx.x; }

__global__ __launch_bounds__(256) void kb_push(BatchView* __restrict__ V, BatchTiles T) {
  for (int r = 0;; ++r) {
    const int t = batch_entry_p(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (!v.active) continue;
    PushConst c = v.push;
    c.tb0 = v.time_bunch;
    push_body<false>(v.part, v.cap, v.np, v.eb, v.und, nullptr, 0, 0, c, v.counters, (long)e.y * 256 + threadIdx.x);
  }
}

// Counting sort of one half of every particle by cell row (the batch's graph-safe replacement
// of the radix sort): keys + counts, a three-phase exclusive scan of the counts into the row
// starts, and a scatter that counts back down, leaving the counters at zero.  Halves that
// deposit nothing form a last bucket (key = rows), so the index list is a permutation of all
// particles, as the radix sort's is.
__device__ inline unsigned batch_slot(const BatchView& v, int slot_sel) { return slot_sel ? 1 - v.parity : v.parity; }

// Runs of equal keys in consecutive lanes: the run's first lane updates the counter for the
// whole run (length len), every lane knows its place o in the run.  All 64 lanes take part.
struct KeyRun {
  bool head;
  unsigned len, o;
};

__device__ inline KeyRun key_run(unsigned key) {
  const int lane = __lane_id();
  const unsigned prev = __shfl_up(key, 1), next = __shfl_down(key, 1);
  const unsigned long long heads = __ballot(lane == 0 || prev != key);
  const unsigned long long tails = __ballot(lane == 63 || next != key);
  const int first = 63 - __builtin_clzll(heads & (~0ull >> (63 - lane)));  // last head at or before lane
  const int last = lane + __builtin_ctzll(tails >> lane);                    // first tail at or after lane
  KeyRun r;
  r.head = lane == first;
  r.len = (unsigned)(last - first + 1);
  r.o = (unsigned)(lane - first);
  return r;
}

constexpr unsigned kNoKey = 0xffffffffu;  // lanes past the particle count

__global__ __launch_bounds__(256) void kb_sort_keys(BatchView* __restrict__ V, BatchTiles T, int use_rm,
                                                    int slot_sel) {
  for (int r = 0;; ++r) {
    const int t = batch_entry_p(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (!v.active) continue;  // uniform over the workgroup
    const long p = (long)e.y * 256 + threadIdx.x;
    unsigned k = kNoKey;
    if (p < v.np) {
      const DepositConst c = v.dep;
      const int o = use_rm ? PXM : PX;
      const gptr<const double> part = G((const double*)v.part);
      const double x = part[(o + 0) * v.cap + p], y = part[(o + 1) * v.cap + p], z = part[(o + 2) * v.cap + p];
      k = v.rows;  // the bucket of halves that deposit nothing, last
      if (inside(c, x, y, z)) {
        const int i = (int)floor(div_rn(x - c.xmin, c.dx, c.ix)), j = (int)floor(div_rn(y - c.ymin, c.dy, c.iy)),
                  kk = (int)floor(div_rn(z - c.zmin, c.dz, c.iz));
        if (cell_ok(c, i, j, kk)) k = (unsigned)kk * (unsigned)c.N0 + (unsigned)i;
      }
      G(v.ord[batch_slot(v, slot_sel)].key)[p] = k;
    }
    const KeyRun run = key_run(k);
    if (k != kNoKey && run.head) atomicAdd(&v.cnt[k], run.len);
  }
}

constexpr int kScanBlock = 1024;  // rows per scan block (256 threads x 4)

// phase 1: exclusive scan of the counts within each block of 1024 rows (rows + 1 entries, the
// last one zero), block totals into bsum
__global__ __launch_bounds__(256) void kb_scan1(BatchView* __restrict__ V, BatchTiles T, int slot_sel) {
  __shared__ unsigned s[256];
  for (int r = 0;; ++r) {
    const int t = batch_entry_p(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (!v.active) continue;  // uniform over the workgroup
    const long r0 = (long)e.y * kScanBlock + 4 * threadIdx.x;
    unsigned c[4], sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      c[q] = (r0 + q <= (long)v.rows) ? G((const unsigned*)v.cnt)[r0 + q] : 0u;
      sum += c[q];
    }
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int h = 1; h < 256; h <<= 1) {  // inclusive Hillis-Steele scan of the thread sums
      const unsigned add = threadIdx.x >= (unsigned)h ? s[threadIdx.x - h] : 0u;
      __syncthreads();
      s[threadIdx.x] += add;
      __syncthreads();
    }
    unsigned run = s[threadIdx.x] - sum;
    const gptr<unsigned> row = G(v.ord[batch_slot(v, slot_sel)].row);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (r0 + q <= (long)v.rows) row[r0 + q] = run;
      run += c[q];
    }
    if (threadIdx.x == 255) G(v.bsum)[e.y] = s[255];
    __syncthreads();
  }
}

// phase 2: exclusive scan of the block totals, one workgroup per instance
__global__ __launch_bounds__(256) void kb_scan2(BatchView* __restrict__ V, BatchTiles T) {
  __shared__ unsigned s[256];
  for (int r = 0;; ++r) {
    const int t = batch_entry_p(r);
    if (t >= T.n) break;
    const BatchView& v = V[T.t[t].x];
    if (!v.active) continue;
    const int nb = (int)(((long)v.rows + 1 + kScanBlock - 1) / kScanBlock);
    unsigned carry = 0;
    for (int b0 = 0; b0 < nb; b0 += 256) {
      const int b = b0 + threadIdx.x;
      const unsigned x = b < nb ? G((const unsigned*)v.bsum)[b] : 0u;
      s[threadIdx.x] = x;
      __syncthreads();
      for (int h = 1; h < 256; h <<= 1) {
        const unsigned add = threadIdx.x >= (unsigned)h ? s[threadIdx.x - h] : 0u;
        __syncthreads();
        s[threadIdx.x] += add;
        __syncthreads();
      }
      if (b < nb) G(v.bsum)[b] = carry + s[threadIdx.x] - x;
      carry += s[255];
      __syncthreads();
    }
  }
}

// phase 3: add the block offsets
__global__ __launch_bounds__(256) void kb_scan3(BatchView* __restrict__ V, BatchTiles T, int slot_sel) {
  for (int r = 0;; ++r) {
    const int t = batch_entry_p(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (!v.active) continue;
    const unsigned off = G((const unsigned*)v.bsum)[e.y];
    const gptr<unsigned> row = G(v.ord[batch_slot(v, slot_sel)].row);
    for (int q = threadIdx.x; q < kScanBlock; q += 256) {
      const long rr = (long)e.y * kScanBlock + q;
      if (rr <= (long)v.rows) row[rr] += off;
    }
  }
}

__global__ __launch_bounds__(256) void kb_sort_scatter(BatchView* __restrict__ V, BatchTiles T, int slot_sel) {
  for (int r = 0;; ++r) {
    const int t = batch_entry_p(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (!v.active) continue;  // uniform over the workgroup
    const long p = (long)e.y * 256 + threadIdx.x;
    const CellOrder o = v.ord[batch_slot(v, slot_sel)];
    // every particle gets a position (k <= rows), so the list is a permutation of all particles
    const unsigned k = p < v.np ? G((const unsigned*)o.key)[p] : kNoKey;
    const KeyRun run = key_run(k);
    unsigned end = 0;
    if (k != kNoKey && run.head) end = atomicSub(&v.cnt[k], run.len);  // counts back down to zero
    // the run's first lane holds the end of the run's range; lanes of one run are consecutive
    const int lane = __lane_id();
    end = __shfl(end, lane - (int)run.o);
    if (k != kNoKey) G(o.idx)[G((const unsigned*)o.row)[k] + end - run.len + run.o] = (unsigned)p;
  }
}

// The single domain's row order by counting.  The particles are walked in a guide list, the
// previous step's row order, in which the new rows are nearly sorted (a particle crosses a cell
// row now and then), so runs of equal keys in a wavefront are long and its atomics few.  Pass
// 1: the key of each listed particle (rnp, or rnm with use_rm) and the row counts.
__global__ __launch_bounds__(256) void k_guided_keys(const double* __restrict__ part, long cap, long n, DepositConst c,
                                                     int use_rm, unsigned rows, const unsigned* __restrict__ guide,
                                                     unsigned* __restrict__ key, unsigned* __restrict__ cnt) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned k = kNoKey;
  if (s < n) {
    const long p = guide ? (long)guide[s] : s;
    const int o = use_rm ? PXM : PX;
    const double x = part[(o + 0) * cap + p], y = part[(o + 1) * cap + p], z = part[(o + 2) * cap + p];
    k = rows;  // the bucket of halves that deposit nothing, last
    if (inside(c, x, y, z)) {
      const int i = (int)floor(div_rn(x - c.xmin, c.dx, c.ix)), j = (int)floor(div_rn(y - c.ymin, c.dy, c.iy)),
                kk = (int)floor(div_rn(z - c.zmin, c.dz, c.iz));
      if (cell_ok(c, i, j, kk)) k = (unsigned)kk * (unsigned)c.N0 + (unsigned)i;
    }
    key[s] = k;
  }
  const KeyRun run = key_run(k);
  if (k != kNoKey && run.head) atomicAdd(&cnt[k], run.len);
}

// Pass 3 (after the counts are scanned into the row starts): each listed particle to its place
// in its row.  The counts are taken back down to zero, ready for the next order.
__global__ __launch_bounds__(256) void k_guided_scatter(const unsigned* __restrict__ key_in,
                                                        const unsigned* __restrict__ guide, long n,
                                                        const unsigned* __restrict__ row, unsigned* __restrict__ cnt,
                                                        unsigned* __restrict__ key_out, unsigned* __restrict__ idx_out) {
  const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned k = s < n ? key_in[s] : kNoKey;
  const KeyRun run = key_run(k);
  unsigned end = 0;
  if (k != kNoKey && run.head) end = atomicSub(&cnt[k], run.len);
  // the run's first lane holds the end of the run's range; lanes of one run are consecutive
  const int lane = __lane_id();
  end = __shfl(end, lane - (int)run.o);
  if (k != kNoKey) {
    const unsigned pos = row[k] + end - run.len + run.o;  // < n whenever the counts are whole
    if (pos < (unsigned)n) {
      idx_out[pos] = guide ? guide[s] : (unsigned)s;
      key_out[pos] = k;
    }
  }
}

template <typename R>
__global__ __launch_bounds__(kDepositThreads) void kb_deposit(const BatchView* __restrict__ V, BatchTiles T) {
  for (int r = 0;; ++r) {
    const int t = batch_entry_p(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (v.active) {
      const CellOrder po = v.ord[v.parity];
      const CellOrder mo = v.ord[1 - v.parity];
      const DepositConst c = v.dep;  // a local copy: the LDS atomics cannot alias it
      if (v.ncomp == 4)  // space charge: rho rides as the fourth component
        deposit_rows_body<R, true>(v.part, v.cap, v.lvl[v.i_np1], c, po.row, po.idx, mo.row, mo.idx, v.dTI,
                                   v.dtilesI, kDepositChunk, 0, 0.0, 0.0, 0, e.y);
      else
        deposit_rows_body<R, false>(v.part, v.cap, v.lvl[v.i_np1], c, po.row, po.idx, mo.row, mo.idx, v.dTI,
                                    v.dtilesI, kDepositChunk, 0, 0.0, 0.0, 0, e.y);
    }
    __syncthreads();  // the LDS tile is reused by the next entry
  }
}

// Reorder of the particle arrays into this step's rnp row order every `reorder` steps, after
// the deposit (as the single-domain path reorders): a gather into the spare arrays, which
// kb_advance swaps in; the order's index list becomes the identity.
__device__ inline bool reorder_step(const BatchView& v) { return v.reorder > 0 && v.step % v.reorder == 0; }

__global__ __launch_bounds__(256) void kb_reorder(BatchView* __restrict__ V, BatchTiles T) {
  for (int r = 0;; ++r) {
    const int t = batch_entry_p(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    const long p = (long)e.y * 256 + threadIdx.x;
    if (!v.active || !reorder_step(v) || p >= v.np) continue;
    const gptr<unsigned> perm = G(v.ord[v.parity].idx);
    const gptr<const double> src = G((const double*)v.part);
    const gptr<double> dst = G(v.part_alt);
    const long s = perm[p];
#pragma unroll
    for (int a = 0; a < PCOUNT; ++a) dst[a * v.cap + p] = src[a * v.cap + s];
    G(v.pid_alt)[p] = G((const unsigned*)v.pid)[s];
    perm[p] = (unsigned)p;
  }
}

// Bunching factor samples: pass 1 writes each particle block's (sum q cos, sum q sin, sum q);
// pass 2 (one workgroup per instance) adds them in block order.
__global__ __launch_bounds__(256) void kb_bunch1(BatchView* __restrict__ V, BatchTiles T) {
  __shared__ double s[3][256];
  for (int r = 0;; ++r) {
    const int t = batch_entry_p(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (!v.active || v.bevery <= 0 || v.step % v.bevery != 0 || v.step / v.bevery >= v.bslots) continue;
    const long p = (long)e.y * 256 + threadIdx.x;
    double c = 0.0, sn = 0.0, q = 0.0;
    if (p < v.np) {
      const gptr<const double> part = G((const double*)v.part);
      const double w = part[PQ * v.cap + p], z = part[PZ * v.cap + p];
      double sv, cv;
      sincos(v.bk * z, &sv, &cv);
      c = w * cv;
      sn = w * sv;
      q = w;
    }
    s[0][threadIdx.x] = c;
    s[1][threadIdx.x] = sn;
    s[2][threadIdx.x] = q;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h)
        for (int a = 0; a < 3; ++a) s[a][threadIdx.x] += s[a][threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x < 3) G(v.bpart)[3 * (long)e.y + threadIdx.x] = s[threadIdx.x][0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(64) void kb_bunch2(BatchView* __restrict__ V, BatchTiles T) {
  for (int r = 0;; ++r) {
    const int t = batch_entry_p(r);
    if (t >= T.n) break;
    const BatchView& v = V[T.t[t].x];
    if (!v.active || v.bevery <= 0 || v.step % v.bevery != 0 || v.step / v.bevery >= v.bslots) continue;
    if (threadIdx.x == 0) {
      const long nb = (v.np + 255) / 256;
      double c = 0.0, sn = 0.0, q = 0.0;
      const gptr<const double> bp = G((const double*)v.bpart);
      for (long b = 0; b < nb; ++b) {
        c += bp[3 * b];
        sn += bp[3 * b + 1];
        q += bp[3 * b + 2];
      }
      G(v.bout)[v.step / v.bevery] = q != 0.0 ? sqrt(c * c + sn * sn) / fabs(q) : 0.0;
    }
  }
}

}  // namespace

hipError_t launch_bunching(Domain* d, double k, double* partial, int blocks) {
  if (d->np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bunching, dim3(blocks), dim3(256), 0, d->stream, d->part, d->cap, d->np, k, partial);
  return hipGetLastError();
}

hipError_t launch_zrange(Domain* d) {
  d->zrange_valid = 0;
  if (d->np == 0) return hipSuccess;
  const mithra_grid& g = d->g;
  const long want = (d->np + 255) / 256;
  const int nb = (int)(want < kZBlocks ? want : kZBlocks);
  hipLaunchKernelGGL(k_zcells, dim3(nb), dim3(256), 0, d->stream, d->part, d->cap, d->np, g.zmin, g.dz, 1.0 / g.dz,
                     g.N2 - 2, (int2*)d->zpart);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_zcells_final, dim3(1), dim3(256), 0, d->stream, (const int2*)d->zpart, nb, d->zrange);
  if ((e = hipGetLastError()) == hipSuccess) d->zrange_valid = 1;
  return e;
}

hipError_t launch_bunch_begin(Domain* d) {
  d->rm_order_valid = d->order1_is_rnp;  // rnm := rnp
  if (d->np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_bunch_begin, dim3((unsigned)((d->np + 255) / 256)), dim3(256), 0, d->stream, d->part, d->cap,
                     d->np);
  return hipGetLastError();
}

hipError_t launch_push(Domain* d, const mithra_bunch_params* p, const UndDev* und, int n_und,
                       double time_bunch, double bunch_time_step, int n_sub, int save_rnm) {
  if (save_rnm) d->rm_order_valid = d->order1_is_rnp;  // rnm := rnp
  if (n_sub > 0) d->order1_is_rnp = 0;                  // rnp moves
  if (d->np == 0 || n_sub <= 0) return hipSuccess;
  const mithra_grid& g = d->g;
  PushConst c;
  c.xmin = g.xmin; c.xmax = g.xmax; c.ymin = g.ymin; c.ymax = g.ymax; c.zmin = g.zmin;
  c.zp0 = g.zp0; c.zp1 = g.zp1; c.dx = g.dx; c.dy = g.dy; c.dz = g.dz; c.Lz = g.Lz;
  c.N0 = g.N0; c.N1 = g.N1; c.N2 = g.N2;
  c.ix = 1.0 / g.dx; c.iy = 1.0 / g.dy; c.iz = 1.0 / g.dz;
  c.gamma = p->gamma; c.beta = p->beta; c.c0 = p->c0; c.dt_shift = p->dt_shift;
  c.r1 = p->r1; c.r2 = p->r2; c.dtb = p->dtb;
  c.tb0 = time_bunch; c.dtb_time = bunch_time_step;
  c.n_sub = n_sub; c.n_und = n_und; c.save_rnm = save_rnm;
  c.dist0 = n_und > 0 ? d->und_dist0 : 0.0;  // kept at upload: `und` is device memory
  if (d->n_opt + d->n_ext > 0)
    hipLaunchKernelGGL(k_push<true>, dim3((unsigned)((d->np + 255) / 256)), dim3(256), 0, d->stream, d->part, d->cap,
                       d->np, d->eb, und, d->beams, d->n_opt, d->n_ext, c, d->counters);
  else
    hipLaunchKernelGGL(k_push<false>, dim3((unsigned)((d->np + 255) / 256)), dim3(256), 0, d->stream, d->part,
                       d->cap, d->np, d->eb, und, nullptr, 0, 0, c, d->counters);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_zrange(d);
}

// Cells of the fixed-order deposit's keys: (N2 - 1) N0 rows of N1 - 1 cells (0 when they do
// not fit a 32-bit key with its sentinel).
unsigned long long fixed_order_cells(const Domain* d) {
  const unsigned long long cells = (unsigned long long)(d->g.N2 - 1) * d->g.N0 * (d->g.N1 - 1);
  return cells < 0xffffffffull ? cells : 0ull;
}

namespace {

DepositConst deposit_const(const Domain* d) {
  const mithra_grid& g = d->g;
  DepositConst c;
  c.xmin = g.xmin; c.xmax = g.xmax; c.ymin = g.ymin; c.ymax = g.ymax; c.zmin = g.zmin;
  c.zp0 = g.zp0; c.zp1 = g.zp1; c.dx = g.dx; c.dy = g.dy; c.dz = g.dz;
  c.N0 = g.N0; c.N1 = g.N1; c.N2 = g.N2; c.sc = g.space_charge;
  c.ix = 1.0 / g.dx; c.iy = 1.0 / g.dy; c.iz = 1.0 / g.dz;
  return c;
}

// guide: a list of all particles in which the keys are nearly sorted (the other order), or
// nullptr for array order.  The fixed-order deposit sorts by cell with the stable radix sort;
// the default deposit counts rows in the guide's order.
hipError_t build_order(Domain* d, CellOrder& o, int use_rm, unsigned rows, const unsigned* guide) {
  const DepositConst c = deposit_const(d);
  const unsigned grid = (unsigned)((d->np + 255) / 256);
  hipError_t e;
  if (d->deterministic) {
    const unsigned span = (unsigned)(c.N1 - 1);  // cells
    hipLaunchKernelGGL(k_row_keys, dim3(grid), dim3(256), 0, d->stream, d->part, d->cap, d->np, c, use_rm, rows,
                       span, o.key, o.idx);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return sort_order(d, o, d->np, rows, span);
  }
  hipLaunchKernelGGL(k_guided_keys, dim3(grid), dim3(256), 0, d->stream, d->part, d->cap, d->np, c, use_rm, rows,
                     guide, d->key_alt, d->row_cnt);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = scan_row_counts(d, d->row_cnt, o.row, rows)) != hipSuccess) {
    (void)launch_fill_zero(d, d->row_cnt, sizeof(unsigned) * ((size_t)rows + 1));  // counts zero again
    return e;
  }
  hipLaunchKernelGGL(k_guided_scatter, dim3(grid), dim3(256), 0, d->stream, d->key_alt, guide, d->np, o.row,
                     d->row_cnt, o.key, o.idx);
  if ((e = hipGetLastError()) != hipSuccess)
    (void)launch_fill_zero(d, d->row_cnt, sizeof(unsigned) * ((size_t)rows + 1));  // counts zero again
  return e;
}

// The fixed-order deposit's buffers: both orders' cell starts and both halves' weights.
hipError_t ensure_fixed_order(Domain* d) {
  const unsigned long long cells = fixed_order_cells(d);
  const size_t starts = (size_t)(cells + 1 + 63) / 64 * 64;
  const size_t nw = d->ncomp == 4 ? 32 : 24;  // weights per particle: rnp (12 or 20) + rnm (12)
  const size_t bytes = 2 * starts * sizeof(unsigned) + nw * (size_t)d->cap * sizeof(double);
  if (d->det_block && d->det_bytes == bytes && d->ord[0].cell && d->ord[1].cell) return hipSuccess;
  if (!d->det_block || d->det_bytes != bytes) {
    (void)hipFree(d->det_block);
    d->det_block = nullptr;
    d->det_bytes = 0;
    hipError_t e = hipMalloc(&d->det_block, bytes);
    if (e != hipSuccess) return e;
    d->det_bytes = bytes;
  }
  unsigned* b = (unsigned*)d->det_block;
  d->ord[0].cell = b;
  d->ord[1].cell = b + starts;
  d->det_w[0] = (double*)(b + 2 * starts);
  d->det_w[1] = d->det_w[0] + (d->ncomp == 4 ? 20 : 12) * (size_t)d->cap;
  d->rm_order_valid = 0;  // the starts were (re)assigned: list both orders afresh
  return hipSuccess;
}

// Fixed-order deposition of J (and rho): both orders listed by cell, the weights of every listed
// half, then one thread per node.  The reorder and the order swap follow as in launch_deposit.
hipError_t deposit_fixed_order(Domain* d, int accumulate, unsigned rows) {
  const unsigned cells = (unsigned)fixed_order_cells(d);
  hipError_t e;
  if ((e = ensure_fixed_order(d)) != hipSuccess) return e;
  if (!d->rm_order_valid && (e = build_order(d, d->ord[1], 1, rows, nullptr)) != hipSuccess) return e;
  if ((e = build_order(d, d->ord[0], 0, rows, d->ord[1].idx)) != hipSuccess) return e;
  const DepositConst c = deposit_const(d);
  const dim3 gp((unsigned)((d->np + 255) / 256));
  for (int h = 0; h < 2; ++h) {
    if (c.sc)
      hipLaunchKernelGGL((k_half_weights<true>), gp, dim3(256), 0, d->stream, d->part, d->cap, d->np, c,
                         d->ord[h].key, d->ord[h].idx, cells, h, d->det_w[h]);
    else
      hipLaunchKernelGGL((k_half_weights<false>), gp, dim3(256), 0, d->stream, d->part, d->cap, d->np, c,
                         d->ord[h].key, d->ord[h].idx, cells, h, d->det_w[h]);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  const FieldSet J = d->lvl[d->i_np1];
  const dim3 gn((unsigned)((d->nodes + 255) / 256));
  const unsigned *pc = d->ord[0].cell, *mc = d->ord[1].cell, *pr = d->ord[0].row, *mr = d->ord[1].row;
  const bool f64 = d->precision == MITHRA_PRECISION_FP64;
  if (f64 && c.sc)
    hipLaunchKernelGGL((k_deposit_fixed<double, true>), gn, dim3(256), 0, d->stream, J, c, pc, mc, pr, mr, d->det_w[0],
                       d->det_w[1], d->cap, accumulate);
  else if (f64)
    hipLaunchKernelGGL((k_deposit_fixed<double, false>), gn, dim3(256), 0, d->stream, J, c, pc, mc, pr, mr, d->det_w[0],
                       d->det_w[1], d->cap, accumulate);
  else if (c.sc)
    hipLaunchKernelGGL((k_deposit_fixed<float, true>), gn, dim3(256), 0, d->stream, J, c, pc, mc, pr, mr, d->det_w[0],
                       d->det_w[1], d->cap, accumulate);
  else
    hipLaunchKernelGGL((k_deposit_fixed<float, false>), gn, dim3(256), 0, d->stream, J, c, pc, mc, pr, mr, d->det_w[0],
                       d->det_w[1], d->cap, accumulate);
  return hipGetLastError();
}

// Rows per workgroup tile of the row-tiled deposit: its two-plane LDS ring
// (2 x ncomp x TI x N1 doubles) stays within 40 KiB, so four 512-thread workgroups (eight
// waves per SIMD, the kernel's register budget) share a CU's 160 KiB; 0 when even one row does
// not fit.
int tile_rows(const Domain* d) {
  const long per_row = 2L * d->ncomp * d->g.N1 * (long)sizeof(double);
  const long fit = (40L << 10) / per_row;
  return (int)std::min(fit, 16L);
}


hipError_t deposit_atomic(Domain* d) {
  const DepositConst c = deposit_const(d);
  const dim3 grid((unsigned)((d->np + 255) / 256));
  const FieldSet J = d->lvl[d->i_np1];
  const bool f64 = d->precision == MITHRA_PRECISION_FP64;
  if (f64 && c.sc)
    hipLaunchKernelGGL((k_deposit<double, true>), grid, dim3(256), 0, d->stream, d->part, d->cap, d->np, J, c);
  else if (f64)
    hipLaunchKernelGGL((k_deposit<double, false>), grid, dim3(256), 0, d->stream, d->part, d->cap, d->np, J, c);
  else if (c.sc)
    hipLaunchKernelGGL((k_deposit<float, true>), grid, dim3(256), 0, d->stream, d->part, d->cap, d->np, J, c);
  else
    hipLaunchKernelGGL((k_deposit<float, false>), grid, dim3(256), 0, d->stream, d->part, d->cap, d->np, J, c);
  return hipGetLastError();
}

}  // namespace

int deposit_tile_rows(const Domain* d) { return tile_rows(d); }

long deposit_blocks(int N2, int tilesI, int KC) { return (long)tilesI * ((N2 + KC - 1) / KC); }

int deposit_chunk() { return kDepositChunk; }

size_t deposit_lds_bytes(int ncomp, int TI, int N1) { return sizeof(double) * 2 * ncomp * TI * N1; }

DepositConst make_deposit_const(const Domain* d) { return deposit_const(d); }

bool deposit_tiled(const Domain* d) {
  const long rows = (long)(d->g.N2 - 1) * d->g.N0;
  return tile_rows(d) > 0 && rows < 0xFFFFFFF0L;
}

hipError_t finish_deposit(Domain* d);

hipError_t launch_deposit(Domain* d, int accumulate) {
  const mithra_grid& g = d->g;
  if (accumulate == kDepositSource && (!deposit_tiled(d) || d->np == 0 || d->deterministic))
    return hipErrorInvalidValue;
  const double src_coef = d->co.a[4], src_rho = d->co.a[5];
  if (d->deterministic && d->np > 0) {
    const unsigned rows = (unsigned)((g.N2 - 1) * (long)g.N0);
    hipError_t e;
    if ((e = ensure_orders(d, rows)) != hipSuccess) return e;
    if ((e = deposit_fixed_order(d, accumulate, rows)) != hipSuccess) return e;
    return finish_deposit(d);
  }
  if (!deposit_tiled(d)) {
    if (!accumulate) {
      for (int c = 0; c < d->ncomp; ++c) {
        hipError_t e = launch_fill_zero(d, d->lvl[d->i_np1].c[c], (size_t)d->nodes * d->real_size);
        if (e != hipSuccess) return e;
      }
    }
    return d->np ? deposit_atomic(d) : hipSuccess;
  }
  hipError_t e;
  if (d->np == 0) {  // nothing to add
    if (accumulate) return hipSuccess;
    for (int cc = 0; cc < d->ncomp; ++cc)
      if ((e = launch_fill_zero(d, d->lvl[d->i_np1].c[cc], (size_t)d->nodes * d->real_size)) != hipSuccess) return e;
    return hipSuccess;
  }
  const unsigned rows = (unsigned)((g.N2 - 1) * (long)g.N0);
  if ((e = ensure_orders(d, rows)) != hipSuccess) return e;
  if (!d->rm_order_valid && (e = build_order(d, d->ord[1], 1, rows, nullptr)) != hipSuccess) return e;
  if ((e = build_order(d, d->ord[0], 0, rows, d->ord[1].idx)) != hipSuccess) return e;
  // A single domain must fill the GPU by itself: shorter chunks, then fewer rows per tile,
  // until there are some 2048 workgroups (each costs a little redundant work at its edges).
  int TI = tile_rows(d), KC = kDepositChunk;
  while (deposit_blocks(g.N2, (g.N0 + TI - 1) / TI, KC) < 2048) {
    if (KC > 2) KC /= 2;
    else if (TI > 1) TI = (TI + 1) / 2;
    else break;
  }
  const int tilesI = (g.N0 + TI - 1) / TI;
  const DepositConst c = deposit_const(d);
  const FieldSet J = d->lvl[d->i_np1];
  const size_t lds = deposit_lds_bytes(d->ncomp, TI, g.N1);
  const dim3 grid((unsigned)deposit_blocks(g.N2, tilesI, KC));
  const unsigned *pr = d->ord[0].row, *pi = d->ord[0].idx, *mr = d->ord[1].row, *mi = d->ord[1].idx;
  const bool f64 = d->precision == MITHRA_PRECISION_FP64;
  // see deposit_rows_body: exact only when every source coefficient carries a minus sign
  const int skip = std::signbit(src_coef) && (!c.sc || std::signbit(src_rho)) ? 1 : 0;
#define MITHRA_DEPOSIT(R_, SC_, SRC_)                                                                          \
  hipLaunchKernelGGL((k_deposit_rows<R_, SC_, SRC_>), grid, dim3(kDepositThreads), lds, d->stream, d->part, d->cap, J, \
                     c, pr, pi, mr, mi, TI, tilesI, KC, accumulate, src_coef, src_rho, skip)
  const bool src = accumulate == kDepositSource;
  if (f64) {
    if (c.sc) {
      if (src) MITHRA_DEPOSIT(double, true, true);
      else MITHRA_DEPOSIT(double, true, false);
    } else {
      if (src) MITHRA_DEPOSIT(double, false, true);
      else MITHRA_DEPOSIT(double, false, false);
    }
  } else {
    if (c.sc) {
      if (src) MITHRA_DEPOSIT(float, true, true);
      else MITHRA_DEPOSIT(float, true, false);
    } else {
      if (src) MITHRA_DEPOSIT(float, false, true);
      else MITHRA_DEPOSIT(float, false, false);
    }
  }
#undef MITHRA_DEPOSIT
  if ((e = hipGetLastError()) != hipSuccess) return e;
  return finish_deposit(d);
}

// After a sorted deposit: the periodic reorder of the particle arrays into the rnp order, and
// the rnp order becomes the rnm order of the next deposit once rnm is set from rnp.
hipError_t finish_deposit(Domain* d) {
  if (d->sort_interval > 0 && d->sort_count % d->sort_interval == 0)
    if (hipError_t e = reorder_particles(d, d->ord[0]); e != hipSuccess) return e;
  ++d->sort_count;
  CellOrder t = d->ord[0];
  d->ord[0] = d->ord[1];
  d->ord[1] = t;
  d->order1_is_rnp = 1;
  d->rm_order_valid = 0;
  return hipSuccess;
}


// ---- batched launchers --------------------------------------------------------------------
hipError_t launch_batch_push(BatchView* V, BatchTiles T, hipStream_t s) {
  if (T.n) hipLaunchKernelGGL(kb_push, dim3(batch_grid(T.n, 8)), dim3(256), 0, s, V, T);
  return hipGetLastError();
}

hipError_t launch_batch_bunching(BatchView* V, BatchTiles particles, BatchTiles inst, hipStream_t s) {
  if (particles.n) hipLaunchKernelGGL(kb_bunch1, dim3(batch_grid(particles.n, 8)), dim3(256), 0, s, V, particles);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !inst.n) return e;
  hipLaunchKernelGGL(kb_bunch2, dim3(inst.n), dim3(64), 0, s, V, inst);
  return hipGetLastError();
}

hipError_t launch_batch_sort(BatchView* V, BatchTiles particles, BatchTiles rowblocks, BatchTiles inst, int use_rm,
                             int slot_sel, hipStream_t s) {
  if (!particles.n) return hipSuccess;
  const int gp = batch_grid(particles.n, 8);
  hipLaunchKernelGGL(kb_sort_keys, dim3(gp), dim3(256), 0, s, V, particles, use_rm, slot_sel);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(kb_scan1, dim3(batch_grid(rowblocks.n, 8)), dim3(256), 0, s, V, rowblocks, slot_sel);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(kb_scan2, dim3(inst.n), dim3(256), 0, s, V, inst);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(kb_scan3, dim3(batch_grid(rowblocks.n, 8)), dim3(256), 0, s, V, rowblocks, slot_sel);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(kb_sort_scatter, dim3(gp), dim3(256), 0, s, V, particles, slot_sel);
  return hipGetLastError();
}

hipError_t launch_batch_reorder(BatchView* V, BatchTiles particles, hipStream_t s) {
  if (particles.n) hipLaunchKernelGGL(kb_reorder, dim3(batch_grid(particles.n, 8)), dim3(256), 0, s, V, particles);
  return hipGetLastError();
}

hipError_t launch_batch_deposit(BatchView* V, BatchTiles T, int precision, size_t lds, hipStream_t s) {
  if (!T.n) return hipSuccess;
  const int g = batch_grid(T.n, lds > (40u << 10) ? 2 : 3);
  if (precision == MITHRA_PRECISION_FP64)
    hipLaunchKernelGGL((kb_deposit<double>), dim3(g), dim3(kDepositThreads), lds, s, V, T);
  else
    hipLaunchKernelGGL((kb_deposit<float>), dim3(g), dim3(kDepositThreads), lds, s, V, T);
  return hipGetLastError();
}

}  // namespace mithra
