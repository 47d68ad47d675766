/*
 * kat_harness.cpp -- TEST INFRASTRUCTURE ONLY (part of the oracle, never shipped).
 *
 * Known-answer tables of the reference's analytic seed, computed by the reference itself
 * (compiled from /root/reference/src by oracle/Makefile into oracle/_ref/): Signal::self
 * (src/classes.cpp:534-575) for every signal type over times that cross each branch, and
 * Seed::fields (src/classes.cpp:740-855) for every seed type and signal over one lattice of
 * points at four times.  Seed::fields is called the way the TF/SF loop calls it
 * (src/fdtd.cpp:239, :315-371): one output vector, zeroed once, passed to every call in turn,
 * which is what gives the super-gaussian type (which adds into its output) its values.
 *
 * Usage: kat_harness <out.bin>   writes a flat dump (tests/golden/dumpio.py reads it).
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "classes.h"
#include "stdinclude.h"

using namespace MITHRA;

namespace {

struct Dump {
  FILE* f = nullptr;
  explicit Dump(const char* path) {
    f = std::fopen(path, "wb");
    if (!f) { std::perror(path); std::exit(2); }
    std::fwrite("MITHRADUMP1\n", 1, 12, f);
  }
  ~Dump() { if (f) std::fclose(f); }
  void dv(const std::string& name, const std::vector<double>& v) {
    unsigned int len = (unsigned int)name.size();
    char type = 'd';
    unsigned long long n = v.size();
    std::fwrite(&len, 4, 1, f);
    std::fwrite(name.data(), 1, len, f);
    std::fwrite(&type, 1, 1, f);
    std::fwrite(&n, 8, 1, f);
    if (n) std::fwrite(v.data(), 8, n, f);
  }
};

const char* kSignals[] = {"neumann", "gaussian", "secant-hyperbolic", "flat-top", "inverse-gaussian"};
const char* kSeeds[] = {"plane-wave", "truncated-plane-wave", "gaussian-beam", "super-gaussian-beam"};

// Signal parameters as Solver::setSimulationParameters leaves them (src/solver.cpp:75-78):
// offset and pulse length divided by c0, the frequency multiplied by it.
void scale_signal(Signal& s, double c0) {
  s.t0_ /= c0;
  s.f0_ *= c0;
  s.s_ /= c0;
}

std::vector<double> signal_params(const Signal& s) {
  return {(double)s.signalType_, s.t0_, s.s_, s.f0_, s.cep_, (double)s.nR_, s.sigmaInvG_[0], s.sigmaInvG_[1]};
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: kat_harness <out.bin>\n"); return 2; }
  Dump out(argv[1]);
  const double c0 = 299.792458;  // micrometre per picosecond

  /* Signal::self: each type, two parameter sets (the second with one rising cycle), times
   * from t0 - 12 s to t0 + 12 s (beyond the 10 s cut-off) plus the branch edges, three phases. */
  int idx = 0;
  for (int set = 0; set < 2; ++set)
    for (const char* type : kSignals) {
      Signal s;
      const double offset = set ? 31.0 : 50.0, pulse = set ? 7.5 : 20.0, wavelength = set ? 1.3 : 2.98;
      const unsigned nR = set ? 1 : 3;
      std::vector<double> sig = set ? std::vector<double>{0.9, 2.6} : std::vector<double>{0.05, 0.013};
      s.initialize(type, offset, pulse, wavelength, set ? -35.0 : 23.0, nR, sig);
      scale_signal(s, c0);
      std::vector<double> ts, ph, val;
      const double t0 = s.t0_, ss = s.s_;
      std::vector<double> times;
      for (int q = -240; q <= 240; ++q) times.push_back(t0 + ss * q / 20.0);
      for (double e : {-10.0, -0.5, 0.5, 10.0}) times.push_back(t0 + e * ss);
      for (double t : times)
        for (double phase : {0.0, 0.37, -2.1}) {
          double tt = t, pp = phase;
          ts.push_back(t);
          ph.push_back(phase);
          val.push_back(s.self(tt, pp));
        }
      const std::string p = "signal" + std::to_string(idx++) + "_";
      out.dv(p + "params", signal_params(s));
      out.dv(p + "t", ts);
      out.dv(p + "phase", ph);
      out.dv(p + "value", val);
    }

  /* Seed::fields: every seed type with every signal, two geometries (propagation along z with
   * a diagonal polarisation; a tilted propagation direction), a lattice of boosted-frame points
   * and three times; the output vector is shared by all calls of one table, like atemp. */
  std::vector<double> pts, tv;  // the lattice, shared by every table
  for (double time : {0.0, 0.0125, 3.0, 140.0})  // the last: the pulse has passed every point
    for (int k = 0; k < 7; ++k)
      for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 4; ++j) {
          pts.push_back(-900.0 + 450.0 * i);
          pts.push_back(-700.0 + 466.0 * j);
          pts.push_back(-3000.0 + 1000.0 * k);
          tv.push_back(time);
        }
  out.dv("seed_r", pts);
  out.dv("seed_time", tv);
  idx = 0;
  for (int geo = 0; geo < 2; ++geo)
    for (const char* seedType : kSeeds)
      for (const char* sigType : kSignals) {
        Signal s;
        std::vector<double> sig = {0.05, 0.013};
        s.initialize(sigType, 50.0, 20.0, 2.98, 11.0, 2, sig);
        Seed seed;
        std::vector<double> position = {10.0, -20.0, 50.0};
        std::vector<double> direction = geo ? std::vector<double>{0.01, 0.0, 1.0} : std::vector<double>{0.0, 0.0, 1.0};
        std::vector<double> polarization = geo ? std::vector<double>{0.0, 1.0, 0.0} : std::vector<double>{1.0, 1.0, 0.0};
        std::vector<double> radius = {300.0, 260.0};
        std::vector<int> order = {1, 2};
        seed.initialize(seedType, position, direction, polarization, 2.0e-5, radius, order, s);
        // Solver::setSimulationParameters (src/solver.cpp:75-78, :119) and the boost values
        // (src/solver.cpp:212-213, :327)
        seed.c0_ = c0;
        scale_signal(seed.signal_, c0);
        seed.amplitude_ = seed.a0_ * EM * c0 / EC;
        seed.gamma_ = 100.41;
        seed.beta_ = std::sqrt(1.0 - 1.0 / (seed.gamma_ * seed.gamma_));
        seed.dt_ = -0.731;
        std::vector<double> av;
        FieldVector<double> a(0.0), r(0.0);
        for (size_t q = 0; q < tv.size(); ++q) {
          r[0] = pts[3 * q];
          r[1] = pts[3 * q + 1];
          r[2] = pts[3 * q + 2];
          double t = tv[q];
          seed.fields(r, t, a);
          av.push_back(a[0]); av.push_back(a[1]); av.push_back(a[2]);
        }
        const std::string p = "seed" + std::to_string(idx++) + "_";
        std::vector<double> desc = {(double)seed.seedType_, seed.amplitude_, seed.radius_[0], seed.radius_[1],
                                    !seed.order_.empty() ? (double)seed.order_[0] : 0.0,
                                    seed.order_.size() > 1 ? (double)seed.order_[1] : 0.0, seed.gamma_, seed.beta_,
                                    seed.dt_, seed.c0_};
        for (int c = 0; c < 3; ++c) desc.push_back(seed.position_[c]);
        for (int c = 0; c < 3; ++c) desc.push_back(seed.direction_[c]);
        for (int c = 0; c < 3; ++c) desc.push_back(seed.polarization_[c]);
        out.dv(p + "desc", desc);
        out.dv(p + "signal", signal_params(seed.signal_));
        out.dv(p + "a", av);
      }
  return 0;
}
