/*
 * cpu_ref.cpp -- TEST INFRASTRUCTURE: CPU oracle of the FDTD/PIC time march.
 *
 * Restates, for one domain (the reference at mpirun -np 1), what these reference routines
 * compute, in the same floating-point operation order:
 *   interior NSFD / FD update   src/database.cpp:44-134, loops src/fdtd.cpp:270-303,
 *                               src/fdtdSC.cpp:317-366
 *   TF/SF seed injection        src/fdtd.cpp:307-373 with Seed::fields src/classes.cpp:740-855
 *                               and Signal::self src/classes.cpp:534-575
 *   Mur faces, edges, corners   src/database.cpp:137-288, loops src/fdtd.cpp:377-725
 *   E/B evaluation              src/fdtd.cpp:742-774, :818-845; src/fdtdSC.cpp:1110-1145
 *   bunch sub-step              src/solver.cpp:1424-1549, undulator src/solver.cpp:1798-1824,
 *                               src/beam.cc:14-76; laser beams of optical undulators and
 *                               external fields src/solver.cpp:1825-1947, src/beam.cc:80-496
 *   current deposition          src/fdtd.cpp:23-185, src/fdtdSC.cpp:23-210
 *   radiation power             src/radiation.cpp:18-232
 * Known divergence, shared with the HIP path: a particle whose periodic z coordinate leaves
 * [zp0, zp1) is counted, not duplicated (the reference re-appends it to its own list on one
 * rank, src/solver.cpp:1544-1568).  The fixtures keep their bunches clear of the z edges.
 *
 * Parity status of each part (what pins it):
 *   every routine above -- bit-exact against the reference's own dumps after 1 and 100
 *   steps for the golden jobs (NSFD/FD, Mur order 1/2, space charge on/off, Gaussian-beam and
 *   plane-wave seeds, undulator array, ellipsoid and file bunches): tests/test_oracle_golden.py
 *   plus closed-form checks in tests/test_oracle_invariants.py.
 *   oracle_signal and oracle_seed_fields, every signal type (neumann, gaussian, secant,
 *   flat-top, inverse-gaussian) and every seed type (plane-wave, truncated-plane-wave,
 *   Gaussian-beam, super-gaussian-beam with its accumulation into the caller's vector) --
 *   bit-exact against known-answer tables the reference itself computes
 *   (oracle/kat_harness.cpp -> tests/golden/kat_seed_signal.npz, tests/test_oracle_kat.py).
 */
#include "cpu_ref.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <new>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* Constants of the reference (src/stdinclude.h:43-50), PI deliberately truncated. */
const double PI = 3.1415926535;
const double EC = 1.602e-19;
const double EM = 9.109e-31;

inline double pmod(double a, double b) {  // src/stdinclude.cpp:88-93
  double x = std::fmod(a, b);
  x += (x < 0.0) ? b : 0.0;
  return x;
}

}  // namespace

struct oracle_dev {
  mithra_grid g;
  mithra_coeffs c;
  long N1N0 = 0, nodes = 0;
  std::vector<double> A[3];   // AoS3 potentials, rotated through lvl[]
  std::vector<double> F[3];   // phi levels (space charge)
  int lvl_np1 = 0, lvl_n = 1, lvl_nm1 = 2;
  std::vector<float> en, bn;  // AoS3
  std::vector<double> q;      // Charge records, 11 doubles each
  long n_transverse = 0, n_cross = 0;
  // power sampling
  bool power_ready = false;
  unsigned pN = 0, pNl = 0, pNf = 0, rows = 0;
  double pc = 0, pgamma = 0, pbeta = 0, pc0 = 0;
  std::vector<double> pz, pw;
  std::vector<std::vector<std::complex<double>>> ep, em;
  std::vector<double> fdt;    // [m][ni][4]
  std::vector<double> history;
  // laser beams of optical undulators (table order) and external fields
  std::vector<mithra_beam> optical, external;

  double* anp1() { return A[lvl_np1].data(); }
  double* an() { return A[lvl_n].data(); }
  double* anm1() { return A[lvl_nm1].data(); }
  double* fnp1() { return F[lvl_np1].data(); }
  double* fn() { return F[lvl_n].data(); }
  double* fnm1() { return F[lvl_nm1].data(); }
};

/* ---------------------------------------------------------------- analytic seed ------ */

extern "C" double oracle_signal(const mithra_seed* s, double t, double phase) {
  // Signal::self, src/classes.cpp:534-575
  const double t0 = s->t0, ss = s->s, f0 = s->f0, cep = s->cep;
  if (std::fabs(t - t0) > 10.0 * ss) return 0.0;
  switch (s->signal_type) {
    case MITHRA_SIGNAL_NEUMANN:
      return -std::cos(2 * PI * f0 * (t - t0) + cep + phase) * 2.7724 * (t - t0) / (ss * ss) *
             std::exp(-1.3863 * (t - t0) * (t - t0) / (ss * ss));
    case MITHRA_SIGNAL_GAUSSIAN:
      return std::cos(2 * PI * f0 * (t - t0) + cep + phase) * std::exp(-1.3863 * std::pow((t - t0) / ss, 2));
    case MITHRA_SIGNAL_SECANT:
      return std::cos(2 * PI * f0 * (t - t0) + cep + phase) / std::cosh((t - t0) / ss);
    case MITHRA_SIGNAL_FLATTOP:
      if (t - t0 <= -ss / 2.0)
        return std::cos(2 * PI * f0 * (t - t0) + cep + phase) *
               std::exp(-std::pow((t - t0 + ss / 2.0) * f0 / s->nR, 2));
      else if (t - t0 <= ss / 2.0)
        return std::cos(2 * PI * f0 * (t - t0) + cep + phase);
      else
        return std::cos(2 * PI * f0 * (t - t0) + cep + phase) *
               std::exp(-std::pow((t - t0 - ss / 2.0) * f0 / s->nR, 2));
    case MITHRA_SIGNAL_INVGAUSSIAN: {
      double env = std::pow((1.0 + std::pow((t - t0) / s->sigma_invg[0], 2)) *
                                (1.0 + std::pow((t - t0) / s->sigma_invg[1], 2)), 0.25);
      double c = std::cos(2 * PI * f0 * (t - t0) + cep + phase);
      if (t - t0 <= -ss / 2.0) return c * env * std::exp(-std::pow((t - t0 + ss / 2.0) * f0 / s->nR, 2));
      if (t - t0 <= ss / 2.0) return c * env;
      return c * env * std::exp(-std::pow((t - t0 - ss / 2.0) * f0 / s->nR, 2));
    }
  }
  return 0.0;
}

/* Seed::fields (src/classes.cpp:740-855).  `a` is read-modify-write for the
 * super-gaussian type exactly like the reference, which never clears its output. */
extern "C" void oracle_seed_fields(const mithra_seed* s, const double r[3], double time, double a[3]) {
  double rl[3], rv[3], yv[3];
  rl[0] = r[0];
  rl[1] = r[1];
  rl[2] = s->gamma * (r[2] + s->beta * s->c0 * (time + s->dt_shift));
  double tl = s->gamma * (time + s->dt_shift + s->beta / s->c0 * r[2]);
  for (int k = 0; k < 3; ++k) rv[k] = rl[k] - s->position[k];
  double z = rv[0] * s->direction[0] + rv[1] * s->direction[1] + rv[2] * s->direction[2];
  tl -= z / s->c0;
  double p = 0.0;
  const double* pol = s->polarization;
  const double* dir = s->direction;
  auto set = [&](double y) { for (int k = 0; k < 3; ++k) a[k] = y * pol[k]; };
  auto transverse = [&](double& x, double& y) {
    x = rv[0] * pol[0] + rv[1] * pol[1] + rv[2] * pol[2];
    yv[0] = dir[1] * pol[2] - dir[2] * pol[1];
    yv[1] = dir[2] * pol[0] - dir[0] * pol[2];
    yv[2] = dir[0] * pol[1] - dir[1] * pol[0];
    y = rv[0] * yv[0] + rv[1] * yv[1] + rv[2] * yv[2];
  };
  if (s->seed_type == MITHRA_SEED_PLANEWAVE) {
    double ts = oracle_signal(s, tl, p);
    if (std::fabs(ts) < 1.0e-6) { a[0] = a[1] = a[2] = 0.0; }
    else set(s->amplitude * ts);
  } else if (s->seed_type == MITHRA_SEED_PLANEWAVE_TRUNCATED) {
    double ts = oracle_signal(s, tl, p);
    double x, y;
    transverse(x, y);
    if (std::fabs(ts) < 1.0e-6 || std::fabs(x) > s->radius[0] || std::fabs(y) > s->radius[1]) { a[0] = a[1] = a[2] = 0.0; }
    else set(s->amplitude * ts);
  } else if (s->seed_type == MITHRA_SEED_GAUSSIANBEAM || s->seed_type == MITHRA_SEED_SUPERGAUSSIANBEAM) {
    double ts = oracle_signal(s, tl, p);
    if (std::fabs(ts) < 1.0e-6) {
      a[0] = a[1] = a[2] = 0.0;
    } else {
      double x, y;
      transverse(x, y);
      double l = s->c0 / s->f0;
      double zRp = PI * s->radius[0] * s->radius[0] / l;
      double wrp = std::sqrt(1.0 + z * z / (zRp * zRp));
      double zRs = PI * s->radius[1] * s->radius[1] / l;
      double wrs = std::sqrt(1.0 + z * z / (zRs * zRs));
      if (s->seed_type == MITHRA_SEED_GAUSSIANBEAM) {
        p = 0.5 * (std::atan(z / zRp) + std::atan(z / zRs) - PI) -
            PI * z / l * (std::pow(x / (zRp * wrp), 2) + std::pow(y / (zRs * wrs), 2));
        ts = oracle_signal(s, tl, p);
        double t = std::exp(-std::pow(x / (s->radius[0] * wrp), 2) - std::pow(y / (s->radius[1] * wrs), 2)) /
                   std::sqrt(wrs * wrp) * s->amplitude;
        set(t * ts);
      } else {
        for (int i = -s->order[0]; i <= s->order[0]; i++)
          for (int j = -s->order[1]; j <= s->order[1]; j++) {
            p = 0.5 * (std::atan(z / zRp) + std::atan(z / zRs) - PI) -
                PI * z / l * (std::pow(x / (zRp * wrp), 2) + std::pow(y / (zRs * wrs), 2));
            ts = oracle_signal(s, tl, p);
            double t = std::exp(-std::pow(x / (s->radius[0] * wrp), 2) - std::pow(y / (s->radius[1] * wrs), 2)) /
                       std::sqrt(wrs * wrp) * s->amplitude;
            for (int k = 0; k < 3; ++k) a[k] += t * ts * pol[k];
          }
      }
    }
  }
  a[2] *= s->gamma;
}

/* ---------------------------------------------------------------- laser beams -------- */

namespace {

// Signal::self of a beam's signal (the beam carries the Signal members).
double beam_signal(const mithra_beam& b, double t, double phase) {
  mithra_seed s;
  std::memset(&s, 0, sizeof s);
  s.signal_type = b.signal_type;
  s.nR = b.nR;
  s.t0 = b.t0;
  s.s = b.s;
  s.f0 = b.f0;
  s.cep = b.cep;
  s.sigma_invg[0] = b.sigma_invg[0];
  s.sigma_invg[1] = b.sigma_invg[1];
  return oracle_signal(&s, t, phase);
}

// UpdateBunchParallel's members that the beam routines read and write.
struct Ubp {
  double rv[3], z, tl, tlm, p0, p1, x, y, x0, y0, wrp, wrs, atanP, atanS, t, af, tsignal, tsignalm, tsignale,
      tsignalb;
  double yv[3], eT[3], bT[3], ex[3], ez[3], by[3], bz[3];
};

void zero(double v[3]) { v[0] = v[1] = v[2] = 0.0; }
void mv(double v[3], double a, const double d[3]) { for (int c = 0; c < 3; ++c) v[c] = a * d[c]; }
void pmv(double v[3], double a, const double d[3]) { for (int c = 0; c < 3; ++c) v[c] += a * d[c]; }
double dot(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
void cross(double z[3], const double x[3], const double y[3]) {  // src/fieldvector.h:197-203
  z[0] = x[1] * y[2] - x[2] * y[1];
  z[1] = x[2] * y[0] - x[0] * y[2];
  z[2] = x[0] * y[1] - x[1] * y[0];
}
void total(Ubp& u) {
  for (int c = 0; c < 3; ++c) {
    u.eT[c] = u.ex[c] + u.ez[c];
    u.bT[c] = u.by[c] + u.bz[c];
  }
}

// Solver::planeWave and planeWaveTruncated (src/beam.cc:80-130).
void planeWave(Ubp& u, const mithra_beam& s, double c0, bool truncated) {
  u.tsignal = beam_signal(s, u.tl, u.p0);
  if (std::fabs(u.tsignal) < 1.0e-6) { zero(u.eT); zero(u.bT); return; }
  if (truncated) {
    u.x = dot(u.rv, s.polarization);
    cross(u.yv, s.direction, s.polarization);
    u.y = dot(u.rv, u.yv);
    if (std::pow(u.x / s.radius[0], 2) + std::pow(u.y / s.radius[1], 2) > 1.0) { zero(u.eT); zero(u.bT); return; }
  }
  mv(u.eT, s.amplitude * u.tsignal, s.polarization);
  cross(u.bT, s.direction, s.polarization);
  const double b[3] = {u.bT[0], u.bT[1], u.bT[2]};
  mv(u.bT, s.amplitude * u.tsignal / c0, b);
}

// Solver::gaussianBeam (src/beam.cc:132-198).
void gaussianBeam(Ubp& u, const mithra_beam& s, double c0) {
  u.x = dot(u.rv, s.polarization);
  cross(u.yv, s.direction, s.polarization);
  u.y = dot(u.rv, u.yv);
  u.wrp = std::sqrt(1.0 + u.z * u.z / (s.zR[0] * s.zR[0]));
  u.wrs = std::sqrt(1.0 + u.z * u.z / (s.zR[1] * s.zR[1]));
  u.x0 = u.x / u.wrp;
  u.y0 = u.y / u.wrs;
  if (std::fabs(u.x0) > 4.0 * s.radius[0] || std::fabs(u.y0) > 4.0 * s.radius[1]) { zero(u.eT); zero(u.bT); return; }
  u.tsignal = beam_signal(s, u.tl, u.p0);
  if (std::fabs(u.tsignal) < 1.0e-6) { zero(u.eT); zero(u.bT); return; }
  u.atanP = std::atan(u.z / s.zR[0]);
  u.atanS = std::atan(u.z / s.zR[1]);
  u.p0 = 0.5 * (u.atanP + u.atanS) - PI * u.z / s.l * (std::pow(u.x0 / s.zR[0], 2) + std::pow(u.y0 / s.zR[1], 2));
  u.t = std::exp(-std::pow(u.x0 / s.radius[0], 2) - std::pow(u.y0 / s.radius[1], 2)) / std::sqrt(u.wrs * u.wrp);
  u.t *= s.amplitude;
  u.p1 = u.p0 - PI / 2.0;
  u.tsignal = beam_signal(s, u.tl, u.p1);
  mv(u.ex, u.t * u.tsignal, s.polarization);
  mv(u.by, u.t * u.tsignal / c0, u.yv);
  u.p1 = u.p0 + u.atanP;
  u.tsignal = beam_signal(s, u.tl, u.p1);
  mv(u.ez, u.t * (-u.x0 / s.zR[0]) * u.tsignal, s.direction);
  u.p1 = u.p0 + u.atanS;
  u.tsignal = beam_signal(s, u.tl, u.p1);
  mv(u.bz, u.t * (-u.y0 / s.zR[1]) / c0 * u.tsignal, s.direction);
  total(u);
}

// Solver::superGaussianBeam (src/beam.cc:200-276).
void superGaussianBeam(Ubp& u, const mithra_beam& s, double c0) {
  u.x = dot(u.rv, s.polarization);
  cross(u.yv, s.direction, s.polarization);
  u.y = dot(u.rv, u.yv);
  u.wrp = std::sqrt(1.0 + u.z * u.z / (s.zR[0] * s.zR[0]));
  u.wrs = std::sqrt(1.0 + u.z * u.z / (s.zR[1] * s.zR[1]));
  if ((std::fabs(u.x) - s.order[0] * s.radius[0]) > 4.0 * s.radius[0] * u.wrp ||
      (std::fabs(u.y) - s.order[1] * s.radius[1]) > 4.0 * s.radius[1] * u.wrs) {
    zero(u.eT); zero(u.bT); return;
  }
  u.tsignal = beam_signal(s, u.tl, u.p0);
  if (std::fabs(u.tsignal) < 1.0e-6) { zero(u.eT); zero(u.bT); return; }
  u.atanP = std::atan(u.z / s.zR[0]);
  u.atanS = std::atan(u.z / s.zR[1]);
  u.af = s.amplitude / std::sqrt(u.wrs * u.wrp);
  zero(u.ex); zero(u.ez); zero(u.bz); zero(u.by);
  for (int i = -s.order[0]; i <= s.order[0]; i++)
    for (int j = -s.order[1]; j <= s.order[1]; j++) {
      u.x0 = (u.x - i * s.radius[0]) / u.wrp;
      u.y0 = (u.y - j * s.radius[1]) / u.wrs;
      if (std::fabs(u.x0) > 4.0 * s.radius[0] || std::fabs(u.y0) > 4.0 * s.radius[1]) continue;
      // the parallel term uses x0, the perpendicular one y (not y0), src/beam.cc:247
      u.p0 = 0.5 * (u.atanP + u.atanS) - PI * u.z / s.l * (std::pow(u.x0 / s.zR[0], 2) + std::pow(u.y / s.zR[1], 2));
      u.t = u.af * std::exp(-std::pow(u.x0 / s.radius[0], 2) - std::pow(u.y0 / s.radius[1], 2));
      u.p1 = u.p0 - PI / 2.0;
      u.tsignal = beam_signal(s, u.tl, u.p1);
      pmv(u.ex, u.t * u.tsignal, s.polarization);
      pmv(u.by, u.t * u.tsignal / c0, u.yv);
      u.p1 = u.p0 + u.atanP;
      u.tsignal = beam_signal(s, u.tl, u.p1);
      pmv(u.ez, u.t * (-u.x0 / s.zR[0]) * u.tsignal, s.direction);
      u.p1 = u.p0 + u.atanS;
      u.tsignal = beam_signal(s, u.tl, u.p1);
      pmv(u.bz, u.t * (-u.y0 / s.zR[1]) * u.tsignal / c0, s.direction);
    }
  total(u);
}

// Solver::standingPlaneWave and standingPlaneWaveTruncated (src/beam.cc:278-327).
void standingPlaneWave(Ubp& u, const mithra_beam& s, double c0, bool truncated) {
  if (truncated) {
    u.x = dot(u.rv, s.polarization);
    cross(u.yv, s.direction, s.polarization);
    u.y = dot(u.rv, u.yv);
    if (std::pow(u.x / s.radius[0], 2) + std::pow(u.y / s.radius[1], 2) > 1.0) { zero(u.eT); zero(u.bT); return; }
  }
  u.tsignal = beam_signal(s, u.tl, u.p0);
  u.tsignalm = beam_signal(s, u.tlm, u.p0);
  u.tsignale = u.tsignal - u.tsignalm;
  u.tsignalb = u.tsignal + u.tsignalm;
  if (std::fabs(u.tsignale) < 1.0e-6 && std::fabs(u.tsignalb) < 1.0e-6) { zero(u.eT); zero(u.bT); return; }
  mv(u.eT, s.amplitude * u.tsignale, s.polarization);
  cross(u.bT, s.direction, s.polarization);
  const double b[3] = {u.bT[0], u.bT[1], u.bT[2]};
  mv(u.bT, s.amplitude * u.tsignalb / c0, b);
}

// Solver::standingGaussianBeam (src/beam.cc:329-400).
void standingGaussianBeam(Ubp& u, const mithra_beam& s, double c0) {
  u.x = dot(u.rv, s.polarization);
  cross(u.yv, s.direction, s.polarization);
  u.y = dot(u.rv, u.yv);
  u.wrp = std::sqrt(1.0 + u.z * u.z / (s.zR[0] * s.zR[0]));
  u.wrs = std::sqrt(1.0 + u.z * u.z / (s.zR[1] * s.zR[1]));
  u.x0 = u.x / u.wrp;
  u.y0 = u.y / u.wrs;
  if (std::fabs(u.x0) > 4.0 * s.radius[0] || std::fabs(u.y0) > 4.0 * s.radius[1]) { zero(u.eT); zero(u.bT); return; }
  u.tsignal = beam_signal(s, u.tl, u.p0);
  u.tsignalm = beam_signal(s, u.tlm, u.p0);
  u.tsignale = u.tsignal - u.tsignalm;
  u.tsignalb = u.tsignal + u.tsignalm;
  if (std::fabs(u.tsignale) < 1.0e-6 && std::fabs(u.tsignalb) < 1.0e-6) { zero(u.eT); zero(u.bT); return; }
  u.atanP = std::atan(u.z / s.zR[0]);
  u.atanS = std::atan(u.z / s.zR[1]);
  u.p0 = 0.5 * (u.atanP + u.atanS) - PI * u.z / s.l * (std::pow(u.x0 / s.zR[0], 2) + std::pow(u.y0 / s.zR[1], 2));
  u.t = std::exp(-std::pow(u.x0 / s.radius[0], 2) - std::pow(u.y0 / s.radius[1], 2)) / std::sqrt(u.wrs * u.wrp);
  u.t *= s.amplitude;
  u.p1 = u.p0 - PI / 2.0;
  u.tsignal = beam_signal(s, u.tl, u.p1);
  u.p1 = u.p0 - PI / 2.0;
  u.tsignalm = beam_signal(s, u.tlm, u.p1);
  mv(u.ex, u.t * (u.tsignal - u.tsignalm), s.polarization);
  mv(u.by, u.t / c0 * (u.tsignal + u.tsignalm), u.yv);
  u.p1 = u.p0 + u.atanP;
  u.tsignal = beam_signal(s, u.tl, u.p1);
  u.p1 = -u.p1;
  u.tsignalm = beam_signal(s, u.tlm, u.p1);
  mv(u.ez, u.t * (-u.x0 / s.zR[0]) * (u.tsignal - u.tsignalm), s.direction);
  u.p1 = u.p0 + u.atanS;
  u.tsignal = beam_signal(s, u.tl, u.p1);
  u.p1 = -u.p1;
  u.tsignalm = beam_signal(s, u.tlm, u.p1);
  mv(u.bz, u.t * (-u.y0 / s.zR[1]) / c0 * (u.tsignal - u.tsignalm), s.direction);
  total(u);
}

// Solver::standingSuperGaussianBeam (src/beam.cc:402-496), with the beam's wavelength in the
// phase where the reference reads the never-set UpdateBunchParallel::l (src/beam.cc:466).
void standingSuperGaussianBeam(Ubp& u, const mithra_beam& s, double c0) {
  u.x = dot(u.rv, s.polarization);
  cross(u.yv, s.direction, s.polarization);
  u.y = dot(u.rv, u.yv);
  u.wrp = std::sqrt(1.0 + u.z * u.z / (s.zR[0] * s.zR[0]));
  u.wrs = std::sqrt(1.0 + u.z * u.z / (s.zR[1] * s.zR[1]));
  if ((std::fabs(u.x) - s.order[0] * s.radius[0]) > 4.0 * s.radius[0] * u.wrp ||
      (std::fabs(u.y) - s.order[1] * s.radius[1]) > 4.0 * s.radius[1] * u.wrs) {
    zero(u.eT); zero(u.bT); return;
  }
  u.tsignal = beam_signal(s, u.tl, u.p0);
  u.tsignalm = beam_signal(s, u.tlm, u.p0);
  u.tsignale = u.tsignal - u.tsignalm;
  u.tsignalb = u.tsignal + u.tsignalm;
  if (std::fabs(u.tsignale) < 1.0e-6 && std::fabs(u.tsignalb) < 1.0e-6) { zero(u.eT); zero(u.bT); return; }
  u.atanP = std::atan(u.z / s.zR[0]);
  u.atanS = std::atan(u.z / s.zR[1]);
  u.af = s.amplitude / std::sqrt(u.wrs * u.wrp);
  zero(u.ex); zero(u.ez); zero(u.bz); zero(u.by);
  for (int i = -s.order[0]; i <= s.order[0]; i++)
    for (int j = -s.order[1]; j <= s.order[1]; j++) {
      u.x0 = (u.x - i * s.radius[0]) / u.wrp;
      u.y0 = (u.y - j * s.radius[1]) / u.wrs;
      if (std::fabs(u.x0) > 4.0 * s.radius[0] || std::fabs(u.y0) > 4.0 * s.radius[1]) continue;
      u.p0 = 0.5 * (u.atanP + u.atanS) - PI * u.z / s.l * (std::pow(u.x0 / s.zR[0], 2) + std::pow(u.y / s.zR[1], 2));
      u.t = u.af * std::exp(-std::pow(u.x0 / s.radius[0], 2) - std::pow(u.y0 / s.radius[1], 2));
      u.p1 = u.p0 - PI / 2.0;
      u.tsignal = beam_signal(s, u.tl, u.p1);
      u.p1 = u.p0 - PI / 2.0;
      u.tsignalm = beam_signal(s, u.tlm, u.p1);
      pmv(u.ex, u.t * (u.tsignal - u.tsignalm), s.polarization);
      pmv(u.by, u.t * (u.tsignal + u.tsignalm) / c0, u.yv);
      u.p1 = u.p0 + u.atanP;
      u.tsignal = beam_signal(s, u.tl, u.p1);
      u.p1 = -u.p1;
      u.tsignalm = beam_signal(s, u.tlm, u.p1);
      pmv(u.ez, u.t * (-u.x0 / s.zR[0]) * (u.tsignal - u.tsignalm), s.direction);
      u.p1 = u.p0 + u.atanS;
      u.tsignal = beam_signal(s, u.tl, u.p1);
      u.p1 = -u.p1;
      u.tsignalm = beam_signal(s, u.tlm, u.p1);
      pmv(u.bz, u.t * (-u.y0 / s.zR[1]) * (u.tsignal - u.tsignalm) / c0, s.direction);
    }
  total(u);
}

// One beam's bunch-frame fields added to et, bt: the lab point and time, the beam's own
// distance and delays, the type's fields, the boost (src/solver.cpp:1825-1877, :1886-1946).
// The backward-wave delay tlm is the beam's own for optical undulators as for external
// fields (the reference computes it only in Solver::externalField).
void add_beam(const mithra_bunch_params* p, const mithra_beam& s, const double r[3], double tb, double et[3],
              double bt[3]) {
  const double gamma = p->gamma, beta = p->beta, c0 = p->c0;
  double rl[3] = {r[0], r[1], gamma * (r[2] + beta * c0 * (tb + p->dt_shift))};
  const double t0 = gamma * (tb + p->dt_shift + beta / c0 * r[2]);
  Ubp u;
  std::memset(&u, 0, sizeof u);
  for (int c = 0; c < 3; ++c) u.rv[c] = rl[c] - s.position[c];
  u.z = dot(u.rv, s.direction);
  u.tl = t0 - u.z / c0;
  u.tlm = t0 + u.z / c0;
  u.p0 = 0.0;
  switch (s.beam_type) {
    case MITHRA_SEED_PLANEWAVE: planeWave(u, s, c0, false); break;
    case MITHRA_SEED_PLANEWAVE_TRUNCATED: planeWave(u, s, c0, true); break;
    case MITHRA_SEED_GAUSSIANBEAM: gaussianBeam(u, s, c0); break;
    case MITHRA_SEED_SUPERGAUSSIANBEAM: superGaussianBeam(u, s, c0); break;
    case MITHRA_SEED_STANDING_PLANEWAVE: standingPlaneWave(u, s, c0, false); break;
    case MITHRA_SEED_STANDING_PLANEWAVE_TRUNCATED: standingPlaneWave(u, s, c0, true); break;
    case MITHRA_SEED_STANDING_GAUSSIANBEAM: standingGaussianBeam(u, s, c0); break;
    case MITHRA_SEED_STANDING_SUPERGAUSSIANBEAM: standingSuperGaussianBeam(u, s, c0); break;
  }
  bt[0] += gamma * (u.bT[0] + beta / c0 * u.eT[1]);
  bt[1] += gamma * (u.bT[1] - beta / c0 * u.eT[0]);
  bt[2] += u.bT[2];
  et[0] += gamma * (u.eT[0] - beta * c0 * u.bT[1]);
  et[1] += gamma * (u.eT[1] + beta * c0 * u.bT[0]);
  et[2] += u.eT[2];
}

}  // namespace

extern "C" void oracle_beam_field(const mithra_bunch_params* p, const mithra_beam* beam, const double r[3],
                                  double tb, double et[3], double bt[3]) {
  add_beam(p, *beam, r, tb, et, bt);
}

/* Solver::undulatorField with Solver::staticUndulator (static undulators; an optical one is
 * skipped here and added by the caller from its beam, in table order). */
static void undulator_field(const mithra_bunch_params* p, const mithra_undulator* und, int n_und,
                            const mithra_beam* optical, const double r[3], double tb, double et[3], double bt[3]);

extern "C" void oracle_undulator_field(const mithra_bunch_params* p, const mithra_undulator* und, int n_und,
                                       const double r[3], double tb, double et[3], double bt[3]) {
  undulator_field(p, und, n_und, nullptr, r, tb, et, bt);
}

static void undulator_field(const mithra_bunch_params* p, const mithra_undulator* und, int n_und,
                            const mithra_beam* optical, const double r[3], double tb, double et[3], double bt[3]) {
  const double gamma = p->gamma, beta = p->beta, c0 = p->c0;
  int beam = 0;
  for (int u = 0; u < n_und; ++u) {
    const mithra_undulator& U = und[u];
    if (U.type == MITHRA_UNDULATOR_OPTICAL) {
      if (optical) add_beam(p, optical[beam++], r, tb, et, bt);
      continue;
    }
    double b0 = (U.lu != 0.0) ? EM * c0 * 2 * PI / U.lu * U.k / EC : 0.0;
    double ku = (U.lu != 0.0) ? 2 * PI / U.lu : 0.0;
    double ct = std::cos(U.theta), st = std::sin(U.theta);
    double lz = gamma * (r[2] + beta * c0 * (tb + p->dt_shift)) - U.rb;
    double ly = r[0] * ct + r[1] * st;
    double L = U.length * U.lu;
    if (lz >= 0.0 && lz <= L) {
      double d1 = b0 * std::cosh(ku * ly) * std::sin(ku * lz) * gamma;
      bt[0] += d1 * ct;
      bt[1] += d1 * st;
      bt[2] += b0 * std::sinh(ku * ly) * std::cos(ku * lz);
      d1 *= c0 * beta;
      et[1] += d1 * ct;
      et[0] += -d1 * st;
      et[2] += 0.0;
    } else if (lz < 0.0) {
      double sz = std::exp(-std::pow(ku * lz, 2) / 2.0);
      if (u != 0) {
        const mithra_undulator& V = und[u - 1];
        double r0 = V.rb + V.length * V.lu - U.rb;
        if (lz < r0 || r0 == 0.0) sz = 0.0;
        else sz *= 0.35875 + 0.48829 * std::cos(PI * lz / r0) + 0.14128 * std::cos(2.0 * PI * lz / r0) +
                   0.01168 * std::cos(3.0 * PI * lz / r0);
      }
      double d1 = b0 * std::cosh(ku * ly) * sz * ku * lz * gamma;
      bt[0] += d1 * ct;
      bt[1] += d1 * st;
      bt[2] += b0 * std::sinh(ku * ly) * sz;
      d1 *= c0 * beta;
      et[1] += d1 * ct;
      et[0] += -d1 * st;
      et[2] += 0.0;
    } else if (lz > L) {
      double t0 = lz - L;
      double sz = std::exp(-std::pow(ku * t0, 2) / 2.0);
      if (u + 1 != n_und) {
        const mithra_undulator& V = und[u + 1];
        double r0 = V.rb - U.rb - L;
        if (t0 > r0 || r0 == 0.0) sz = 0.0;
        else sz *= 0.35875 + 0.48829 * std::cos(PI * t0 / r0) + 0.14128 * std::cos(2.0 * PI * t0 / r0) +
                   0.01168 * std::cos(3.0 * PI * t0 / r0);
      }
      double d1 = b0 * std::cosh(ku * ly) * sz * ku * t0 * gamma;
      bt[0] += d1 * ct;
      bt[1] += d1 * st;
      bt[2] += b0 * std::sinh(ku * ly) * sz;
      d1 *= c0 * beta;
      et[1] += d1 * ct;
      et[0] += -d1 * st;
      et[2] += 0.0;
    }
  }
}

extern "C" double oracle_halton(unsigned i, unsigned j) {  // src/stdinclude.cpp:45-73
  static const unsigned prime[20] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71};
  if (i >= 20) return -1.0;
  int p0, p, k, a;
  double x = 0.0;
  p = prime[i];
  p0 = p;
  k = j + 1;
  while (k > 0) {
    a = k % p;
    x += a / (double)p0;
    k = int(k / p);
    p0 *= p;
  }
  return 1.0 - x;
}

/* ---------------------------------------------------------------- device-like API ---- */

extern "C" int oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

extern "C" int oracle_create(const mithra_grid* grid, const mithra_coeffs* coeffs, oracle_dev** out) {
  if (!grid || !coeffs || !out || grid->N0 < 3 || grid->N1 < 3 || grid->N2 < 3) return MITHRA_E_ARG;
  oracle_dev* d = new (std::nothrow) oracle_dev();
  if (!d) return MITHRA_E_NOMEM;
  d->g = *grid;
  d->c = *coeffs;
  d->N1N0 = (long)grid->N1 * grid->N0;
  d->nodes = d->N1N0 * grid->N2;
  for (int l = 0; l < 3; ++l) {
    d->A[l].assign(3 * d->nodes, 0.0);
    if (grid->space_charge) d->F[l].assign(d->nodes, 0.0);
  }
  d->en.assign(3 * d->nodes, 0.0f);
  d->bn.assign(3 * d->nodes, 0.0f);
  *out = d;
  return MITHRA_OK;
}

extern "C" int oracle_destroy(oracle_dev* d) { delete d; return MITHRA_OK; }

static int level_index(oracle_dev* d, int level) {
  return level == 0 ? d->lvl_n : level == 1 ? d->lvl_nm1 : level == 2 ? d->lvl_np1 : -1;
}

extern "C" int oracle_set_potentials(oracle_dev* d, int level, const double* a, const double* phi) {
  int li = level_index(d, level);
  if (li < 0) return MITHRA_E_ARG;
  if (a) std::memcpy(d->A[li].data(), a, sizeof(double) * 3 * d->nodes);
  if (phi && d->g.space_charge) std::memcpy(d->F[li].data(), phi, sizeof(double) * d->nodes);
  return MITHRA_OK;
}

extern "C" int oracle_get_potentials(oracle_dev* d, int level, double* a, double* phi) {
  int li = level_index(d, level);
  if (li < 0) return MITHRA_E_ARG;
  if (a) std::memcpy(a, d->A[li].data(), sizeof(double) * 3 * d->nodes);
  if (phi && d->g.space_charge) std::memcpy(phi, d->F[li].data(), sizeof(double) * d->nodes);
  return MITHRA_OK;
}

extern "C" int oracle_get_fields(oracle_dev* d, float* e, float* b) {
  if (e) std::memcpy(e, d->en.data(), sizeof(float) * 3 * d->nodes);
  if (b) std::memcpy(b, d->bn.data(), sizeof(float) * 3 * d->nodes);
  return MITHRA_OK;
}

extern "C" int oracle_set_particles(oracle_dev* d, const double* q, long n) {
  if (n < 0 || (n > 0 && !q)) return MITHRA_E_ARG;
  d->q.assign(q, q + 11 * n);
  return MITHRA_OK;
}

extern "C" int oracle_get_particles(oracle_dev* d, double* q, long n) {
  if (n != (long)(d->q.size() / 11)) return MITHRA_E_ARG;
  std::memcpy(q, d->q.data(), sizeof(double) * d->q.size());
  return MITHRA_OK;
}

extern "C" long oracle_particle_count(const oracle_dev* d) { return (long)(d->q.size() / 11); }

/* ---------------------------------------------------------------- field update ------- */

namespace {

/* AdvanceField::advanceMagneticPotentialNSFD / FD for one scalar component
 * (src/database.cpp:44-134); stride is 3 for A and 1 for phi, s the source coefficient. */
inline double nsfd(const double* a, double alpha, double beta, const double* an, const double* anm1,
                   double src, double as, long m, long sX, long sY, long sZ) {
  return a[0] * an[m] - anm1[m] +
         alpha * (a[1] * (an[m + sX] + an[m - sX] + beta * (an[m + sX + sZ] + an[m + sX - sZ] + an[m - sX + sZ] + an[m - sX - sZ])) +
                  a[2] * (an[m + sY] + an[m - sY] + beta * (an[m + sY + sZ] + an[m + sY - sZ] + an[m - sY + sZ] + an[m - sY - sZ]))) +
         a[3] * (an[m + sZ] + an[m - sZ]) + as * src;
}

inline double fd(const double* a, const double* an, const double* anm1, double src, double as, long m, long sX,
                 long sY, long sZ) {
  return a[0] * an[m] - anm1[m] + a[1] * (an[m + sX] + an[m - sX]) + a[2] * (an[m + sY] + an[m - sY]) +
         a[3] * (an[m + sZ] + an[m - sZ]) + as * src;
}

/* Mur face (AdvanceField::advanceBoundaryF, src/database.cpp:137-156).  s is the inward
 * neighbour offset; (t1, t2) the two tangential neighbour offsets whose coefficient is B[3]
 * ((u1, u2) get B[4]).  The argument permutation of the six face loops of src/fdtd.cpp:377-472
 * is: v6..v9 = s+t1, s-t1, s+u1, s-u1 ; v10..v13 = +t1, -t1, +u1, -u1. */
inline double face(const double* B, const double* P, const double* N, const double* M, long m, long s, long t,
                   long u) {
  return B[0] * (M[m] + P[m + s]) + B[1] * M[m + s] + B[2] * (N[m] + N[m + s]) +
         B[3] * (N[m + s + t] + N[m + s - t] + N[m + t] + N[m - t]) +
         B[4] * (N[m + s + u] + N[m + s - u] + N[m + u] + N[m - u]);
}

/* Mur edge (AdvanceField::advanceEdgeF, src/database.cpp:179-205).  o1, o2 are the two
 * inward offsets (o3 = o1 + o2) and w the offset along the edge, matching the permutations
 * of src/fdtd.cpp:474-628: triples (P,N,M) at o1, o2, o3 and the eight N terms at
 * -w, -w+o1, -w+o2, -w+o3, +w, +w+o1, +w+o2, +w+o3. */
inline double edge(const double* B, const double* P, const double* N, const double* M, long m, long o1, long o2,
                   long w) {
  const long o3 = o1 + o2;
  return B[0] * (P[m + o1] + M[m + o2]) + B[1] * (M[m + o1] + P[m + o2]) + B[2] * (M[m] + P[m + o3]) +
         B[3] * (N[m] + N[m + o1] + N[m + o2] + N[m + o3]) +
         B[4] * (N[m - w] + N[m - w + o1] + N[m - w + o2] + N[m - w + o3] + N[m + w] + N[m + w + o1] +
                 N[m + w + o2] + N[m + w + o3]) -
         M[m + o3];
}

/* Mur corner (AdvanceField::advanceCornerF, src/database.cpp:230-262).  sx, sy, sz are the
 * inward offsets; the seven neighbour triples of src/fdtd.cpp:631-724 are, in order,
 * x, y, z, x+y, z+x, z+y, z+x+y. */
inline double corner(const double* h, const double* P, const double* N, const double* M, long m, long sx, long sy,
                     long sz) {
  const long o[7] = {sx, sy, sz, sx + sy, sz + sx, sz + sy, sz + sx + sy};
  double s = N[m] * h[16] + M[m] * h[8];
  for (int t = 0; t < 7; ++t) s = s + P[m + o[t]] * h[1 + t] + N[m + o[t]] * h[16] + M[m + o[t]] * h[9 + t];
  return -s / h[0];
}

}  // namespace

static void boundaries(oracle_dev* d, double* P, const double* N, const double* M, int stride) {
  const mithra_grid& g = d->g;
  const long N0 = g.N0, N1 = g.N1, np = g.N2, N1N0 = d->N1N0;
  const long X = N1 * stride, Y = stride, Z = N1N0 * stride;
  const int ncomp = stride;
  for (int c = 0; c < ncomp; ++c) {
    double* p = P + c;
    const double* n = N + c;
    const double* mm = M + c;
    // faces, src/fdtd.cpp:377-472
#pragma omp parallel for collapse(2)
    for (long j = 1; j < N1 - 1; j++)
      for (long k = 1; k < np - 1; k++) {
        long m0 = (N1N0 * k + j) * stride;
        p[m0] = face(d->c.bB, p, n, mm, m0, X, Y, Z);
        long m1 = (N1N0 * k + N1N0 - N1 + j) * stride;
        p[m1] = face(d->c.bB, p, n, mm, m1, -X, Y, Z);
      }
#pragma omp parallel for collapse(2)
    for (long i = 1; i < N0 - 1; i++)
      for (long k = 1; k < np - 1; k++) {
        long m0 = (N1N0 * k + N1 * i) * stride;
        p[m0] = face(d->c.cB, p, n, mm, m0, Y, X, Z);
        long m1 = (N1N0 * k + N1 * i + N1 - 1) * stride;
        p[m1] = face(d->c.cB, p, n, mm, m1, -Y, X, Z);
      }
#pragma omp parallel for collapse(2)
    for (long i = 1; i < N0 - 1; i++)
      for (long j = 1; j < N1 - 1; j++) {
        long m0 = (N1 * i + j) * stride;
        p[m0] = face(d->c.dB, p, n, mm, m0, Z, X, Y);
        long m1 = (N1N0 * (np - 1) + N1 * i + j) * stride;
        p[m1] = face(d->c.dB, p, n, mm, m1, -Z, X, Y);
      }
    if (g.truncation_order != 2) continue;
    // edges along z, x, y (src/fdtd.cpp:474-628)
    for (long k = 1; k < np - 1; k++) {
      long b = N1N0 * k;
      p[b * stride] = edge(d->c.eE, p, n, mm, b * stride, X, Y, Z);
      p[(b + N1 * (N0 - 1)) * stride] = edge(d->c.eE, p, n, mm, (b + N1 * (N0 - 1)) * stride, -X, Y, Z);
      p[(b + N1 - 1) * stride] = edge(d->c.eE, p, n, mm, (b + N1 - 1) * stride, X, -Y, Z);
      p[(b + N1 * (N0 - 1) + N1 - 1) * stride] =
          edge(d->c.eE, p, n, mm, (b + N1 * (N0 - 1) + N1 - 1) * stride, -X, -Y, Z);
    }
    for (long i = 1; i < N0 - 1; i++) {
      long b0 = N1 * i, b1 = N1 * i + N1N0 * (np - 1);
      p[b0 * stride] = edge(d->c.fE, p, n, mm, b0 * stride, Y, Z, X);
      p[(b0 + N1 - 1) * stride] = edge(d->c.fE, p, n, mm, (b0 + N1 - 1) * stride, -Y, Z, X);
      p[b1 * stride] = edge(d->c.fE, p, n, mm, b1 * stride, Y, -Z, X);
      p[(b1 + N1 - 1) * stride] = edge(d->c.fE, p, n, mm, (b1 + N1 - 1) * stride, -Y, -Z, X);
    }
    for (long j = 1; j < N1 - 1; j++) {
      long b0 = j, b1 = N1N0 - N1 + j, b2 = N1N0 * (np - 1) + j, b3 = N1N0 * (np - 1) + N1 * (N0 - 1) + j;
      p[b0 * stride] = edge(d->c.gE, p, n, mm, b0 * stride, Z, X, Y);
      p[b1 * stride] = edge(d->c.gE, p, n, mm, b1 * stride, Z, -X, Y);
      p[b2 * stride] = edge(d->c.gE, p, n, mm, b2 * stride, -Z, X, Y);
      p[b3 * stride] = edge(d->c.gE, p, n, mm, b3 * stride, -Z, -X, Y);
    }
    // corners (src/fdtd.cpp:631-724)
    const long zb[2] = {0, N1N0 * (np - 1)};
    for (int kz = 0; kz < 2; ++kz) {
      long sz = kz == 0 ? Z : -Z;
      long b = zb[kz];
      p[b * stride] = corner(d->c.hC, p, n, mm, b * stride, X, Y, sz);
      p[(b + N1N0 - N1) * stride] = corner(d->c.hC, p, n, mm, (b + N1N0 - N1) * stride, -X, Y, sz);
      p[(b + N1 - 1) * stride] = corner(d->c.hC, p, n, mm, (b + N1 - 1) * stride, X, -Y, sz);
      p[(b + N1N0 - N1 + N1 - 1) * stride] = corner(d->c.hC, p, n, mm, (b + N1N0 - 1) * stride, -X, -Y, sz);
    }
  }
}

static void tfsf(oracle_dev* d, double time, const mithra_seed* s) {
  // src/fdtd.cpp:307-373 (one domain: rank 0 == size-1)
  const mithra_grid& g = d->g;
  const long N0 = g.N0, N1 = g.N1, np = g.N2, N1N0 = d->N1N0;
  double* P = d->anp1();
  double atemp[3] = {0.0, 0.0, 0.0};
  auto rc = [&](long m, double r[3]) {  // Solver::rc, src/solver.cpp:2301-2315
    long k = m / N1N0, i = (m % N1N0) / N1, j = m - (N1N0 * k + N1 * i);
    r[0] = g.xmin + i * g.dx;
    r[1] = g.ymin + j * g.dy;
    r[2] = g.zmin + k * g.dz;
  };
  auto apply = [&](long m, long src, double coef, double sign) {
    double r[3];
    rc(src, r);
    oracle_seed_fields(s, r, time, atemp);
    for (int c = 0; c < 3; ++c) {
      if (sign < 0) P[3 * m + c] -= coef * atemp[c];
      else P[3 * m + c] += coef * atemp[c];
    }
  };
  const long KI = 2, KF = np - 2;
  for (long j = 2; j < N1 - 2; j++)
    for (long k = KI; k < KF; k++) {
      long m;
      m = N1N0 * k + N1 * 1 + j;        apply(m, m + N1, d->c.a[1], -1);
      m = N1N0 * k + N1 * 2 + j;        apply(m, m - N1, d->c.a[1], +1);
      m = N1N0 * k + N1 * (N0 - 2) + j; apply(m, m - N1, d->c.a[1], -1);
      m = N1N0 * k + N1 * (N0 - 3) + j; apply(m, m + N1, d->c.a[1], +1);
    }
  for (long i = 2; i < N0 - 2; i++)
    for (long k = KI; k < KF; k++) {
      long m;
      m = N1N0 * k + N1 * i + 1;        apply(m, m + 1, d->c.a[2], -1);
      m = N1N0 * k + N1 * i + 2;        apply(m, m - 1, d->c.a[2], +1);
      m = N1N0 * k + N1 * i + N1 - 2;   apply(m, m - 1, d->c.a[2], -1);
      m = N1N0 * k + N1 * i + N1 - 3;   apply(m, m + 1, d->c.a[2], +1);
    }
  for (long i = 2; i < N0 - 2; i++)
    for (long j = 2; j < N1 - 2; j++) {
      long m;
      m = N1N0 * 1 + N1 * i + j;        apply(m, m + N1N0, d->c.a[3], -1);
      m = N1N0 * 2 + N1 * i + j;        apply(m, m - N1N0, d->c.a[3], +1);
    }
  for (long i = 2; i < N0 - 2; i++)
    for (long j = 2; j < N1 - 2; j++) {
      long m;
      m = N1N0 * (np - 2) + N1 * i + j; apply(m, m - N1N0, d->c.a[3], -1);
      m = N1N0 * (np - 3) + N1 * i + j; apply(m, m + N1N0, d->c.a[3], +1);
    }
}

static void eval_eb(oracle_dev* d, long m) {
  // FdTd::fieldEvaluate (src/fdtd.cpp:818-845) / FdTdSC::fieldEvaluate (src/fdtdSC.cpp:1110-1145)
  const mithra_grid& g = d->g;
  const long N1 = g.N1, N1N0 = d->N1N0;
  const double* an = d->an();
  const double* ap = d->anp1();
  const double dt = g.dt, dx2 = 2.0 * g.dx, dy2 = 2.0 * g.dy, dz2 = 2.0 * g.dz;
  float* e = &d->en[3 * m];
  float* b = &d->bn[3 * m];
  /* The source computes every component as float(float(ap/-dt) - an/-dt) [then
   * float(. - grad phi)] through FieldVector<float>::dv / mdv (src/fieldvector.h:87-108).
   * The reference as compiled here (g++ 11.4 -O3, reference makefile:2-5) vectorises x and y
   * into one double expression rounded once to float, while z keeps every float rounding;
   * the golden dumps carry exactly that, so the oracle follows the compiled reference. */
  if (g.space_charge) {
    const double* fn = d->fn();
    e[0] = (float)((ap[3 * m + 0] / -dt - an[3 * m + 0] / -dt) - (fn[m + N1] - fn[m - N1]) / dx2);
    e[1] = (float)((ap[3 * m + 1] / -dt - an[3 * m + 1] / -dt) - (fn[m + 1] - fn[m - 1]) / dy2);
    e[2] = (float)(ap[3 * m + 2] / -dt);
    e[2] -= an[3 * m + 2] / -dt;
    e[2] -= (fn[m + N1N0] - fn[m - N1N0]) / dz2;
  } else {
    e[0] = (float)(ap[3 * m + 0] / -dt - an[3 * m + 0] / -dt);
    e[1] = (float)(ap[3 * m + 1] / -dt - an[3 * m + 1] / -dt);
    e[2] = (float)(ap[3 * m + 2] / -dt);
    e[2] -= an[3 * m + 2] / -dt;
  }
  auto A = [&](const double* a, long o, int c) { return a[3 * (m + o) + c]; };
  b[0] = 0.5 * ((A(an, 1, 2) - A(an, -1, 2)) / dy2 - (A(an, N1N0, 1) - A(an, -N1N0, 1)) / dz2 +
                (A(ap, 1, 2) - A(ap, -1, 2)) / dy2 - (A(ap, N1N0, 1) - A(ap, -N1N0, 1)) / dz2);
  b[1] = 0.5 * ((A(an, N1N0, 0) - A(an, -N1N0, 0)) / dz2 - (A(an, N1, 2) - A(an, -N1, 2)) / dx2 +
                (A(ap, N1N0, 0) - A(ap, -N1N0, 0)) / dz2 - (A(ap, N1, 2) - A(ap, -N1, 2)) / dx2);
  b[2] = 0.5 * ((A(an, N1, 1) - A(an, -N1, 1)) / dx2 - (A(an, 1, 0) - A(an, -1, 0)) / dy2 +
                (A(ap, N1, 1) - A(ap, -N1, 1)) / dx2 - (A(ap, 1, 0) - A(ap, -1, 0)) / dy2);
}

/* E/B on every node a particle or sampler reads; planes k=0, np-1 copy k=1, np-2
 * (src/fdtd.cpp:742-774). */
static void eval_all(oracle_dev* d) {
  const mithra_grid& g = d->g;
  const long N0 = g.N0, N1 = g.N1, np = g.N2, N1N0 = d->N1N0;
#pragma omp parallel for collapse(2)
  for (long k = 1; k < np - 1; k++)
    for (long i = 1; i < N0 - 1; i++)
      for (long j = 1; j < N1 - 1; j++) eval_eb(d, N1N0 * k + N1 * i + j);
  for (long i = 1; i < N0 - 1; i++)
    for (long j = 1; j < N1 - 1; j++) {
      long m = N1N0 + N1 * i + j;
      for (int c = 0; c < 3; ++c) {
        d->en[3 * (m - N1N0) + c] = d->en[3 * m + c];
        d->bn[3 * (m - N1N0) + c] = d->bn[3 * m + c];
      }
      m = N1N0 * (np - 2) + N1 * i + j;
      for (int c = 0; c < 3; ++c) {
        d->en[3 * (m + N1N0) + c] = d->en[3 * m + c];
        d->bn[3 * (m + N1N0) + c] = d->bn[3 * m + c];
      }
    }
}

extern "C" int oracle_field_evaluate(oracle_dev* d) {
  eval_all(d);
  return MITHRA_OK;
}

extern "C" int oracle_field_update(oracle_dev* d, double time, const mithra_seed* seed) {
  const mithra_grid& g = d->g;
  const long N0 = g.N0, N1 = g.N1, np = g.N2, N1N0 = d->N1N0;
  double* P = d->anp1();
  const double* N = d->an();
  const double* M = d->anm1();
  const double* a = d->c.a;
  const double alpha = d->c.alpha, beta = d->c.beta;
  const bool ns = g.solver == MITHRA_SOLVER_NSFD;
  // interior (src/fdtd.cpp:270-303, src/fdtdSC.cpp:317-366); J / rho are read in place
#pragma omp parallel for collapse(2)
  for (long i = 1; i < N0 - 1; i++)
    for (long j = 1; j < N1 - 1; j++)
      for (long k = 1; k < np - 1; k++) {
        long m = N1N0 * k + N1 * i + j;
        for (int c = 0; c < 3; ++c) {
          long l = 3 * m + c;
          double src = P[l];
          P[l] = ns ? nsfd(a, alpha, beta, N, M, src, a[4], l, 3 * N1, 3, 3 * N1N0)
                    : fd(a, N, M, src, a[4], l, 3 * N1, 3, 3 * N1N0);
        }
        if (g.space_charge) {
          double* fp = d->fnp1();
          double src = fp[m];
          fp[m] = ns ? nsfd(a, alpha, beta, d->fn(), d->fnm1(), src, a[5], m, N1, 1, N1N0)
                     : fd(a, d->fn(), d->fnm1(), src, a[5], m, N1, 1, N1N0);
        }
      }
  if (seed && std::fabs(seed->amplitude) > 1.0e-50) {
    if (seed->seed_type < MITHRA_SEED_PLANEWAVE || seed->seed_type > MITHRA_SEED_SUPERGAUSSIANBEAM)
      return MITHRA_E_UNSUPPORTED;
    tfsf(d, time, seed);
  }
  boundaries(d, P, N, M, 3);
  if (g.space_charge) boundaries(d, d->fnp1(), d->fn(), d->fnm1(), 1);
  eval_all(d);
  return MITHRA_OK;
}

/* ---------------------------------------------------------------- particles ---------- */

extern "C" int oracle_set_beams(oracle_dev* d, const mithra_beam* optical, int n_optical, const mithra_beam* external,
                                int n_external) {
  if (!d || n_optical < 0 || n_external < 0) return MITHRA_E_ARG;
  d->optical.assign(optical, optical + n_optical);
  d->external.assign(external, external + n_external);
  return MITHRA_OK;
}

extern "C" int oracle_bunch_begin(oracle_dev* d) {
  const long n = (long)(d->q.size() / 11);
#pragma omp parallel for
  for (long p = 0; p < n; ++p)
    for (int c = 0; c < 3; ++c) d->q[11 * p + 4 + c] = d->q[11 * p + 1 + c];
  return MITHRA_OK;
}

extern "C" int oracle_bunch_update(oracle_dev* d, const mithra_bunch_params* bp, const mithra_undulator* und,
                                   int n_und, double tb) {
  const mithra_grid& g = d->g;
  const long N1 = g.N1, N1N0 = d->N1N0;
  const long n = (long)(d->q.size() / 11);
  long nt = 0, ncross = 0;
#pragma omp parallel for reduction(+ : nt, ncross)
  for (long p = 0; p < n; ++p) {
    double* Q = &d->q[11 * p];
    double* r = Q + 1;
    double* gb = Q + 7;
    double zr = pmod(r[2] - g.zmin, g.Lz) + g.zmin;
    if (!((zr >= g.zp0) && (zr < g.zp1))) continue;
    bool b1x = (r[0] < g.xmax - g.dx && r[0] > g.xmin + g.dx);
    bool b1y = (r[1] < g.ymax - g.dy && r[1] > g.ymin + g.dy);
    bool b1z = (r[2] < g.zp1 && r[2] >= g.zp0);
    double bt[3] = {0.0, 0.0, 0.0}, et[3] = {0.0, 0.0, 0.0};
    undulator_field(bp, und, n_und, d->optical.data(), r, tb, et, bt);
    for (const mithra_beam& b : d->external) add_beam(bp, b, r, tb, et, bt);
    if (Q[10] == 1.0) {
      if (b1x && b1y && b1z) {
        double d1, dxr, dyr, dzr;
        dxr = std::modf((r[0] - g.xmin) / g.dx, &d1); long i = (int)d1;
        dyr = std::modf((r[1] - g.ymin) / g.dy, &d1); long j = (int)d1;
        dzr = std::modf((r[2] - g.zmin) / g.dz, &d1); long k = (int)d1;
        long m = k * N1N0 + i * N1 + j;
        const long off[8] = {0, N1, 1, N1 + 1, N1N0, N1N0 + N1, N1N0 + 1, N1N0 + N1 + 1};
        const double w[8] = {(1.0 - dxr) * (1.0 - dyr) * (1.0 - dzr), dxr * (1.0 - dyr) * (1.0 - dzr),
                             (1.0 - dxr) * dyr * (1.0 - dzr),         dxr * dyr * (1.0 - dzr),
                             (1.0 - dxr) * (1.0 - dyr) * dzr,         dxr * (1.0 - dyr) * dzr,
                             (1.0 - dxr) * dyr * dzr,                 dxr * dyr * dzr};
        for (int t = 0; t < 8; ++t)
          for (int c = 0; c < 3; ++c) et[c] += w[t] * d->en[3 * (m + off[t]) + c];
        for (int t = 0; t < 8; ++t)
          for (int c = 0; c < 3; ++c) bt[c] += w[t] * d->bn[3 * (m + off[t]) + c];
      } else if (!b1x && !b1y && b1z) {
        nt++;
      }
    } else if (n_und > 0) {
      double lz = bp->gamma * (r[2] + bp->beta * bp->c0 * (tb + bp->dt_shift));
      Q[10] = (lz > -und[0].dist) ? 1.0 : 0.0;
    } else {
      Q[10] = 1.0;
    }
    // Boris rotation, src/solver.cpp:1519-1541
    double gbm[3], gbp[3], gbpl[3];
    for (int c = 0; c < 3; ++c) gbm[c] = gb[c];
    for (int c = 0; c < 3; ++c) gbm[c] += bp->r1 * et[c];
    gbp[0] = gbm[1] * bt[2] - gbm[2] * bt[1];
    gbp[1] = gbm[2] * bt[0] - gbm[0] * bt[2];
    gbp[2] = gbm[0] * bt[1] - gbm[1] * bt[0];
    double d1 = std::sqrt(1.0 + (gbm[0] * gbm[0] + gbm[1] * gbm[1] + gbm[2] * gbm[2]));
    double f = bp->r2 / d1;
    for (int c = 0; c < 3; ++c) gbp[c] = f * gbp[c];
    for (int c = 0; c < 3; ++c) gbp[c] += gbm[c];
    gbpl[0] = gbp[1] * bt[2] - gbp[2] * bt[1];
    gbpl[1] = gbp[2] * bt[0] - gbp[0] * bt[2];
    gbpl[2] = gbp[0] * bt[1] - gbp[1] * bt[0];
    f = 2.0 / (d1 / bp->r2 + bp->r2 / d1 * (bt[0] * bt[0] + bt[1] * bt[1] + bt[2] * bt[2]));
    for (int c = 0; c < 3; ++c) gbpl[c] = f * gbpl[c];
    for (int c = 0; c < 3; ++c) gbpl[c] += gbm[c];
    for (int c = 0; c < 3; ++c) gb[c] = gbpl[c];
    for (int c = 0; c < 3; ++c) gb[c] += bp->r1 * et[c];
    double s = bp->dtb / std::sqrt(1.0 + (gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2]));
    double dr[3];
    for (int c = 0; c < 3; ++c) dr[c] = s * gb[c];
    for (int c = 0; c < 3; ++c) r[c] += dr[c];
    zr += dr[2];
    if (zr < g.zp0 || zr >= g.zp1) ncross++;
  }
  d->n_transverse += nt;
  d->n_cross += ncross;
  return MITHRA_OK;
}

extern "C" int oracle_bunch_counters(oracle_dev* d, long* nt, long* nc) {
  if (nt) *nt = d->n_transverse;
  if (nc) *nc = d->n_cross;
  d->n_transverse = d->n_cross = 0;
  return MITHRA_OK;
}

/* ---------------------------------------------------------------- currents ----------- */

extern "C" int oracle_field_shift(oracle_dev* d) {
  int t = d->lvl_nm1;
  d->lvl_nm1 = d->lvl_n;
  d->lvl_n = d->lvl_np1;
  d->lvl_np1 = t;
  return MITHRA_OK;
}

extern "C" int oracle_current_reset(oracle_dev* d) {
  std::fill(d->A[d->lvl_np1].begin(), d->A[d->lvl_np1].end(), 0.0);
  if (d->g.space_charge) std::fill(d->F[d->lvl_np1].begin(), d->F[d->lvl_np1].end(), 0.0);
  return MITHRA_OK;
}

extern "C" int oracle_current_update(oracle_dev* d) {
  const mithra_grid& g = d->g;
  const long N1 = g.N1, N1N0 = d->N1N0;
  double* J = d->anp1();
  double* R = g.space_charge ? d->fnp1() : nullptr;
  const double dx = g.dx, dy = g.dy, dz = g.dz;
  const long n = (long)(d->q.size() / 11);
  auto inside = [&](const double* r) {
    return r[0] < g.xmax - dx && r[0] > g.xmin + dx && r[1] < g.ymax - dy && r[1] > g.ymin + dy &&
           r[2] < g.zp1 && r[2] >= g.zp0;
  };
  for (long p = 0; p < n; ++p) {
    const double* Q = &d->q[11 * p];
    const double* rp = Q + 1;
    const double* rm = Q + 4;
    bool bp = inside(rp), bm = inside(rm);
    if (!(bp || bm)) continue;
    const double q = Q[0];
    int ip = (int)std::floor((rp[0] - g.xmin) / dx);
    int jp = (int)std::floor((rp[1] - g.ymin) / dy);
    int kp = (int)std::floor((rp[2] - g.zmin) / dz);
    int im = (int)std::floor((rm[0] - g.xmin) / dx);
    int jm = (int)std::floor((rm[1] - g.ymin) / dy);
    int km = (int)std::floor((rm[2] - g.zmin) / dz);
    double r[3];
    r[0] = std::min(std::min(im, ip) * dx + dx + g.xmin, std::max(std::max(im, ip) * dx + g.xmin, 0.5 * (rm[0] + rp[0])));
    r[1] = std::min(std::min(jm, jp) * dy + dy + g.ymin, std::max(std::max(jm, jp) * dy + g.ymin, 0.5 * (rm[1] + rp[1])));
    r[2] = std::min(std::min(km, kp) * dz + dz + g.zmin, std::max(std::max(km, kp) * dz + g.zmin, 0.5 * (rm[2] + rp[2])));
    double jcm[3], jcp[3];
    for (int c = 0; c < 3; ++c) jcm[c] = r[c] - rm[c];
    for (int c = 0; c < 3; ++c) jcp[c] = rp[c] - r[c];
    auto deposit = [&](long m, const double* a, const double* jc) {
      double cc;
      double fx = std::modf((0.5 * (a[0] + r[0]) - g.xmin) / dx, &cc);
      double fy = std::modf((0.5 * (a[1] + r[1]) - g.ymin) / dy, &cc);
      double fz = std::modf((0.5 * (a[2] + r[2]) - g.zmin) / dz, &cc);
      double x1 = 1.0 - fx, x2 = fx, y1 = 1.0 - fy, y2 = fy, z1 = 1.0 - fz, z2 = fz;
      const long o[8] = {0, N1, 1, N1 + 1, N1N0, N1N0 + N1, N1N0 + 1, N1N0 + N1 + 1};
      const double wx[8] = {y1 * z1, y1 * z1, y2 * z1, y2 * z1, y1 * z2, y1 * z2, y2 * z2, y2 * z2};
      const double* wy_a[8] = {&x1, &x2, &x1, &x2, &x1, &x2, &x1, &x2};
      const double* wy_b[8] = {&z1, &z1, &z1, &z1, &z2, &z2, &z2, &z2};
      const double* wz_a[8] = {&x1, &x2, &x1, &x2, &x1, &x2, &x1, &x2};
      const double* wz_b[8] = {&y1, &y1, &y2, &y2, &y1, &y1, &y2, &y2};
      (void)wx;
      const double* wx_a[8] = {&y1, &y1, &y2, &y2, &y1, &y1, &y2, &y2};
      const double* wx_b[8] = {&z1, &z1, &z1, &z1, &z2, &z2, &z2, &z2};
      for (int t = 0; t < 8; ++t) J[3 * (m + o[t]) + 0] += q * 0.5 * *wx_a[t] * *wx_b[t] * jc[0];
      for (int t = 0; t < 8; ++t) J[3 * (m + o[t]) + 1] += q * 0.5 * *wy_a[t] * *wy_b[t] * jc[1];
      for (int t = 0; t < 8; ++t) J[3 * (m + o[t]) + 2] += q * 0.5 * *wz_a[t] * *wz_b[t] * jc[2];
    };
    if (bp) {
      long m = N1N0 * kp + N1 * ip + jp;
      deposit(m, rp, jcp);
      if (R) {
        double cc;
        double fx = std::modf((rp[0] - g.xmin) / dx, &cc);
        double fy = std::modf((rp[1] - g.ymin) / dy, &cc);
        double fz = std::modf((rp[2] - g.zmin) / dz, &cc);
        double x1 = 1.0 - fx, x2 = fx, y1 = 1.0 - fy, y2 = fy, z1 = 1.0 - fz, z2 = fz;
        R[m] += q * x1 * y1 * z1;
        R[m + N1] += q * x2 * y1 * z1;
        R[m + 1] += q * x1 * y2 * z1;
        R[m + N1 + 1] += q * x2 * y2 * z1;
        R[m + N1N0] += q * x1 * y1 * z2;
        R[m + N1N0 + N1] += q * x2 * y1 * z2;
        R[m + N1N0 + 1] += q * x1 * y2 * z2;
        R[m + N1N0 + N1 + 1] += q * x2 * y2 * z2;
      }
    }
    if (bm) {
      long m = N1N0 * km + N1 * im + jm;
      deposit(m, rm, jcm);
    }
  }
  return MITHRA_OK;
}

/* ---------------------------------------------------------------- radiation power ---- */

extern "C" int oracle_power_setup(oracle_dev* d, const mithra_power_params* p, unsigned rows) {
  if (!p || p->Nf == 0) return MITHRA_E_ARG;
  const mithra_grid& g = d->g;
  d->pN = p->N; d->pNl = p->Nl; d->pNf = p->Nf; d->pc = p->pc;
  d->pgamma = p->gamma; d->pbeta = p->beta; d->pc0 = p->c0;
  d->pz.assign(p->z, p->z + p->N);
  d->pw.assign(p->w, p->w + p->Nl);
  d->ep.assign(p->Nl, std::vector<std::complex<double>>(p->Nf));
  d->em.assign(p->Nl, std::vector<std::complex<double>>(p->Nf));
  const std::complex<double> I(0.0, 1.0);
  for (unsigned i = 0; i < p->Nl; i++)
    for (unsigned j = 0; j < p->Nf; j++) {  // src/radiation.cpp:110-117
      d->ep[i][j] = std::cos(d->pw[i] * j * g.dt) + I * std::sin(d->pw[i] * j * g.dt);
      d->em[i][j] = std::cos(d->pw[i] * j * g.dt) - I * std::sin(d->pw[i] * j * g.dt);
    }
  unsigned Nz = 0;
  for (unsigned k = 0; k < p->N; ++k)
    if (p->z[k] < g.zp1 && p->z[k] >= g.zp0) ++Nz;
  d->fdt.assign((size_t)p->Nf * Nz * d->N1N0 * 4, 0.0);
  d->rows = rows;
  d->history.assign((size_t)rows * p->N * p->Nl, 0.0);
  d->power_ready = true;
  return MITHRA_OK;
}

extern "C" int oracle_power_sample(oracle_dev* d, unsigned n_time, unsigned row) {
  // Solver::powerSample, src/radiation.cpp:127-232 (E/B already evaluated on every node)
  if (!d->power_ready || row >= d->rows) return MITHRA_E_STATE;
  const mithra_grid& g = d->g;
  const long N0 = g.N0, N1 = g.N1, N1N0 = d->N1N0;
  const unsigned Nl = d->pNl, Nf = d->pNf;
  std::vector<double> pL((size_t)d->pN * Nl, 0.0);
  unsigned kz = 0;
  const size_t slot = (size_t)kz;
  (void)slot;
  const size_t per_m = (d->fdt.size() / 4) / Nf;  // Nz * N1N0
  for (unsigned k = 0; k < d->pN; ++k) {
    if (!(d->pz[k] < g.zp1 && d->pz[k] >= g.zp0)) continue;
    double c;
    double dzr = std::modf((d->pz[k] - g.zmin) / g.dz, &c);
    long kk = (int)c;
    unsigned m = n_time % Nf;
    for (long i = 2; i < N0 - 2; i += 1)
      for (long j = 2; j < N1 - 2; j += 1) {
        long mi = kk * N1 * N0 + i * N1 + j;
        size_t ni = (size_t)kz * N1 * N0 + i * N1 + j;
        double et0 = (1.0 - dzr) * d->en[3 * mi + 0] + dzr * d->en[3 * (mi + N1N0) + 0];
        double et1 = (1.0 - dzr) * d->en[3 * mi + 1] + dzr * d->en[3 * (mi + N1N0) + 1];
        double bt0 = (1.0 - dzr) * d->bn[3 * mi + 0] + dzr * d->bn[3 * (mi + N1N0) + 0];
        double bt1 = (1.0 - dzr) * d->bn[3 * mi + 1] + dzr * d->bn[3 * (mi + N1N0) + 1];
        double* f = &d->fdt[((size_t)m * per_m + ni) * 4];
        f[0] = d->pgamma * (et0 + d->pc0 * d->pbeta * bt1);
        f[1] = d->pgamma * (et1 - d->pc0 * d->pbeta * bt0);
        f[2] = d->pgamma * (bt0 - d->pbeta / d->pc0 * et1);
        f[3] = d->pgamma * (bt1 + d->pbeta / d->pc0 * et0);
        for (unsigned l = 0; l < Nl; l++) {
          std::complex<double> ew1(0.0, 0.0), bw1 = ew1, ew2 = ew1, bw2 = ew1;
          for (unsigned mm = 0; mm < Nf; mm++) {
            const double* h = &d->fdt[((size_t)mm * per_m + ni) * 4];
            ew1 += h[0] * d->ep[l][mm];
            bw1 += h[3] * d->em[l][mm];
            ew2 += h[1] * d->ep[l][mm];
            bw2 += h[2] * d->em[l][mm];
          }
          pL[k * Nl + l] += d->pc * (std::real(ew1 * bw1) - std::real(ew2 * bw2));
        }
      }
    kz += 1;
  }
  std::memcpy(&d->history[(size_t)row * d->pN * Nl], pL.data(), sizeof(double) * pL.size());
  return MITHRA_OK;
}

extern "C" int oracle_power_fetch(oracle_dev* d, unsigned first, unsigned count, double* out) {
  if (!d->power_ready || first + count > d->rows) return MITHRA_E_STATE;
  size_t w = (size_t)d->pN * d->pNl;
  std::memcpy(out, &d->history[(size_t)first * w], sizeof(double) * w * count);
  return MITHRA_OK;
}

/* Bunching factor |sum q exp(i k z)| / |sum q| (test infrastructure twin of mithra_dev_bunching). */
extern "C" int oracle_bunching(oracle_dev* d, double k, double* b) {
  double c = 0.0, s = 0.0, q = 0.0;
  const long n = (long)(d->q.size() / 11);
  for (long p = 0; p < n; ++p) {
    const double w = d->q[11 * p], z = d->q[11 * p + 3];
    c += w * std::cos(k * z);
    s += w * std::sin(k * z);
    q += w;
  }
  *b = q != 0.0 ? std::sqrt(c * c + s * s) / std::fabs(q) : 0.0;
  return MITHRA_OK;
}
