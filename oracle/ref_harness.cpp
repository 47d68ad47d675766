/*
 * ref_harness.cpp -- TEST INFRASTRUCTURE ONLY (part of the oracle, never shipped).
 *
 * Drives the real MITHRA reference (compiled from /root/reference/src by
 * oracle/Makefile into oracle/_ref/) through its own public API and records
 * its internal state in a flat binary format, so that golden fixtures can be
 * generated for the parity tests.  It never changes what the reference
 * computes: it subclasses the reference's FdTd / FdTdSC plugins
 * (src/fdtd.h:15, src/fdtdSC.h:15) and only wraps the virtual
 * fieldUpdate() (src/fdtd.cpp:231, src/fdtdSC.cpp:260), which
 * Solver::solve() calls first in every step (src/solver.cpp:1303).
 *
 *   pre  dump at step n : A^n, A^{n-1}, J (held in the A^{n+1} buffer,
 *                         src/fdtd.cpp:25), phi/rho, the particle list and
 *                         all update coefficients, before fieldUpdate().
 *   post dump at step n : A^{n+1} after the full field update plus E/B on
 *                         every interior node (forced through the reference's
 *                         own fieldEvaluate(), src/fdtd.cpp:818; the values are
 *                         a pure function of A so forcing them early does not
 *                         change any later result).
 *
 * Environment:
 *   MITHRA_DUMP_DIR    output directory (default ".")
 *   MITHRA_DUMP_STEPS  comma separated step numbers to dump (pre and post)
 *   MITHRA_MAX_STEPS   stop (exit 0) when nTime_ reaches this value (after the pre
 *                      dump of that step when it is listed in MITHRA_DUMP_STEPS).
 *   MITHRA_STEP_CLOCK  file to which rank 0 writes "nTime seconds" at the start of every
 *                      step (MPI_Wtime after a barrier of all ranks), for timing the
 *                      reference's own steps without its set-up; no dumps are needed.
 *
 * Usage: ref_harness <job-file>   (single MPI rank, the one-domain case)
 */

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "datainput.h"
#include "fdtd.h"
#include "fdtdSC.h"
#include "readdata.h"
#include "solver.h"

using namespace MITHRA;

namespace {

struct Dump {
  FILE* f = nullptr;
  explicit Dump(const std::string& path) {
    f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    std::fwrite("MITHRADUMP1\n", 1, 12, f);
  }
  ~Dump() { if (f) std::fclose(f); }
  void rec(const char* name, char type, const void* data, unsigned long long n) {
    unsigned int len = (unsigned int)std::strlen(name);
    std::fwrite(&len, 4, 1, f);
    std::fwrite(name, 1, len, f);
    std::fwrite(&type, 1, 1, f);
    std::fwrite(&n, 8, 1, f);
    size_t esz = (type == 'f') ? 4 : 8;
    if (n) std::fwrite(data, esz, n, f);
  }
  void d(const char* name, double v) { rec(name, 'd', &v, 1); }
  void i(const char* name, long long v) { rec(name, 'i', &v, 1); }
  void dv(const char* name, const double* p, size_t n) { rec(name, 'd', p, n); }
  void fv(const char* name, const float* p, size_t n) { rec(name, 'f', p, n); }
};

std::set<unsigned> parse_steps() {
  std::set<unsigned> s;
  const char* e = std::getenv("MITHRA_DUMP_STEPS");
  if (!e) return s;
  std::string str(e);
  size_t p = 0;
  while (p < str.size()) {
    size_t q = str.find(',', p);
    if (q == std::string::npos) q = str.size();
    if (q > p) s.insert((unsigned)std::atol(str.substr(p, q - p).c_str()));
    p = q + 1;
  }
  return s;
}

std::string dump_dir() {
  const char* e = std::getenv("MITHRA_DUMP_DIR");
  return e ? std::string(e) : std::string(".");
}

long max_steps() {
  const char* e = std::getenv("MITHRA_MAX_STEPS");
  return e ? std::atol(e) : -1;
}

FILE* step_clock() {
  const char* e = std::getenv("MITHRA_STEP_CLOCK");
  if (!e) return nullptr;
  int rank = 0;
  MPI_Comm_rank(MPI_COMM_WORLD, &rank);
  return rank == 0 ? std::fopen(e, "w") : nullptr;
}

}  // namespace

/* On one rank Solver::bunchUpdate() sends the migrating-particle buffers to itself with a
 * blocking MPI_Send before it posts the matching receive (src/solver.cpp:1552-1562).  The MPI
 * standard allows that send to block until it is matched, and MPICH's self-send path does, so
 * the unmodified reference deadlocks at its first particle step when run as a single domain.
 * Through the standard profiling interface, sends addressed to the calling rank become
 * MPI_Isend + MPI_Request_free.  The receive that consumes the message is posted in the same
 * scope while the send buffer is still alive, so the data and every computed number are
 * unchanged; sends to other ranks go straight to PMPI_Send. */
extern "C" int MPI_Send(const void* buf, int count, MPI_Datatype type, int dest, int tag,
                        MPI_Comm comm) {
  int rank = -1;
  PMPI_Comm_rank(comm, &rank);
  if (dest == rank) {
    MPI_Request req;
    int rc = PMPI_Isend(buf, count, type, dest, tag, comm, &req);
    if (rc == MPI_SUCCESS) rc = PMPI_Request_free(&req);
    return rc;
  }
  return PMPI_Send(buf, count, type, dest, tag, comm);
}

template <class Base>
class Harness : public Base {
 public:
  using Base::Base;
  std::set<unsigned> steps_ = parse_steps();
  std::string dir_ = dump_dir();
  long maxSteps_ = max_steps();
  bool clock_ = std::getenv("MITHRA_STEP_CLOCK") != nullptr;
  FILE* clockFile_ = step_clock();

  void dumpCommon(Dump& o) {
    Solver& s = *this;
    o.i("nTime", s.nTime_);
    o.d("time", s.time_);
    o.d("timeBunch", s.timeBunch_);
    o.i("N0", s.N0_); o.i("N1", s.N1_); o.i("N2", s.N2_);
    o.i("np", s.np_); o.i("k0", s.k0_);
    o.d("xmin", s.xmin_); o.d("xmax", s.xmax_);
    o.d("ymin", s.ymin_); o.d("ymax", s.ymax_);
    o.d("zmin", s.zmin_); o.d("zmax", s.zmax_);
    o.d("zp0", s.zp_[0]); o.d("zp1", s.zp_[1]);
    o.d("gamma", s.gamma_); o.d("beta", s.beta_); o.d("dt_shift", s.dt_);
    o.d("c0", s.c0_); o.d("m0", s.m0_); o.d("e0", s.e0_);
    o.d("nUpdateBunch", s.nUpdateBunch_);
    o.d("bunchTimeStep", s.bunch_.timeStep_);
    o.d("meshLength0", s.mesh_.meshLength_[0]);
    o.d("meshLength1", s.mesh_.meshLength_[1]);
    o.d("meshLength2", s.mesh_.meshLength_[2]);
    o.d("totalTime", s.mesh_.totalTime_);
    o.i("solver", (long long)s.mesh_.solver_);
    o.i("truncationOrder", (long long)s.mesh_.truncationOrder_);
    o.i("spaceCharge", (long long)s.mesh_.spaceCharge_);
    o.d("uf_dt", s.uf_.dt); o.d("uf_dx", s.uf_.dx); o.d("uf_dy", s.uf_.dy); o.d("uf_dz", s.uf_.dz);
    o.dv("uf_a", s.uf_.a, 6);
    o.dv("uf_bB", s.uf_.bB, 5); o.dv("uf_cB", s.uf_.cB, 5); o.dv("uf_dB", s.uf_.dB, 5);
    o.dv("uf_eE", s.uf_.eE, 5); o.dv("uf_fE", s.uf_.fE, 5); o.dv("uf_gE", s.uf_.gE, 5);
    o.dv("uf_hC", s.uf_.hC, 17);
    o.d("af_alpha", s.uf_.af.alpha_); o.d("af_beta", s.uf_.af.beta_);
    o.d("uc_dv", s.uc_.dv); o.d("uc_rc", s.uc_.rc);
    o.d("ub_r1", s.ub_.r1); o.d("ub_r2", s.ub_.r2); o.d("ub_dtb", s.ub_.dtb);
    /* seed after unit scaling and boost parameters */
    o.i("seed_type", (long long)s.seed_.seedType_);
    o.d("seed_amplitude", s.seed_.amplitude_);
    o.dv("seed_position", &s.seed_.position_[0], 3);
    o.dv("seed_direction", &s.seed_.direction_[0], 3);
    o.dv("seed_polarization", &s.seed_.polarization_[0], 3);
    o.dv("seed_radius", s.seed_.radius_.data(), s.seed_.radius_.size());
    o.i("seed_signal_type", (long long)s.seed_.signal_.signalType_);
    o.d("seed_t0", s.seed_.signal_.t0_); o.d("seed_s", s.seed_.signal_.s_);
    o.d("seed_f0", s.seed_.signal_.f0_); o.d("seed_cep", s.seed_.signal_.cep_);
    o.i("seed_nR", (long long)s.seed_.signal_.nR_);
    o.dv("seed_sigma_invg", s.seed_.signal_.sigmaInvG_.data(), s.seed_.signal_.sigmaInvG_.size());
    std::vector<double> order(s.seed_.order_.begin(), s.seed_.order_.end());
    order.resize(2, 0.0);
    o.dv("seed_order", order.data(), 2);
    o.d("seed_gamma", s.seed_.gamma_); o.d("seed_beta", s.seed_.beta_);
    o.d("seed_dt", s.seed_.dt_); o.d("seed_c0", s.seed_.c0_);
    /* undulators after sorting/shifting */
    std::vector<double> und;
    for (auto& u : s.undulator_) {
      und.push_back((double)u.type_); und.push_back(u.k_); und.push_back(u.lu_);
      und.push_back(u.rb_); und.push_back((double)u.length_); und.push_back(u.dist_);
      und.push_back(u.theta_);
    }
    o.dv("undulators", und.data(), und.size());
    /* laser beams of the optical undulators (table order), then of the external fields, after
       the scaling of setSimulationParameters: 27 values each */
    std::vector<double> beams;
    auto beam = [&](double kind, int seedType, const Signal& g, const std::vector<int>& order, double amp,
                    const FieldVector<Double>& pos, const FieldVector<Double>& dir, const FieldVector<Double>& pol,
                    const std::vector<Double>& radius, const std::vector<Double>& zR, double l) {
      double v[27] = {kind, (double)seedType, (double)g.signalType_, (double)order[0], (double)order[1],
                      (double)g.nR_, amp, pos[0], pos[1], pos[2], dir[0], dir[1], dir[2], pol[0], pol[1], pol[2],
                      radius[0], radius[1], zR[0], zR[1], l, g.t0_, g.s_, g.f0_, g.cep_, g.sigmaInvG_[0],
                      g.sigmaInvG_[1]};
      beams.insert(beams.end(), v, v + 27);
    };
    for (auto& u : s.undulator_)
      if (u.type_ == OPTICAL)
        beam(0.0, u.seedType_, u.signal_, u.order_, u.amplitude_, u.position_, u.direction_, u.polarization_,
             u.radius_, u.zR_, u.l_);
    for (auto& e : s.extField_)
      beam(1.0, e.seedType_, e.signal_, e.order_, e.amplitude_, e.position_, e.direction_, e.polarization_,
           e.radius_, e.zR_, e.l_);
    if (!beams.empty()) o.dv("beams", beams.data(), beams.size());
    /* radiation power sampling parameters */
    std::vector<double> rp;
    for (size_t jf = 0; jf < s.rp_.size(); ++jf) {
      if (!s.FEL_[jf].radiationPower_.sampling_) continue;
      auto& r = s.rp_[jf];
      rp.push_back((double)r.N); rp.push_back((double)r.Nz); rp.push_back((double)r.Nl);
      rp.push_back((double)r.Nf); rp.push_back(r.pc);
      for (unsigned l = 0; l < r.Nl; ++l) rp.push_back(r.w[l]);
      for (unsigned k = 0; k < r.N; ++k) rp.push_back(s.FEL_[jf].radiationPower_.z_[k]);
    }
    o.dv("power", rp.data(), rp.size());
  }

  void dumpPre() {
    Solver& s = *this;
    char name[64];
    std::snprintf(name, sizeof name, "/pre_%06u.bin", s.nTime_);
    Dump o(dir_ + name);
    dumpCommon(o);
    size_t n = (size_t)s.N1N0_ * s.np_;
    o.dv("an", &(*s.an_)[0][0], 3 * n);
    o.dv("anm1", &(*s.anm1_)[0][0], 3 * n);
    o.dv("jn", &(*s.anp1_)[0][0], 3 * n);
    if (s.mesh_.spaceCharge_) {
      o.dv("fn", s.fn_->data(), n);
      o.dv("fnm1", s.fnm1_->data(), n);
      o.dv("rn", s.fnp1_->data(), n);
    }
    std::vector<double> q;
    q.reserve(s.chargeVectorn_.size() * 11);
    for (auto& c : s.chargeVectorn_) {
      q.push_back(c.q);
      for (int k = 0; k < 3; ++k) q.push_back(c.rnp[k]);
      for (int k = 0; k < 3; ++k) q.push_back(c.rnm[k]);
      for (int k = 0; k < 3; ++k) q.push_back(c.gb[k]);
      q.push_back(c.e);
    }
    o.dv("particles", q.data(), q.size());
  }

  void dumpPost() {
    Solver& s = *this;
    /* Force E/B on every interior node through the reference's own evaluator. */
    for (int k = 1; k < s.np_ - 1; ++k)
      for (int i = 1; i < s.N0_ - 1; ++i)
        for (int j = 1; j < s.N1_ - 1; ++j) {
          long m = (long)s.N1N0_ * k + (long)s.N1_ * i + j;
          if (!s.pic_[m]) this->fieldEvaluate(m);
        }
    char name[64];
    std::snprintf(name, sizeof name, "/post_%06u.bin", s.nTime_);
    Dump o(dir_ + name);
    o.i("nTime", s.nTime_);
    size_t n = (size_t)s.N1N0_ * s.np_;
    o.dv("anp1", &(*s.anp1_)[0][0], 3 * n);
    if (s.mesh_.spaceCharge_) o.dv("fnp1", s.fnp1_->data(), n);
    o.fv("en", &s.en_[0][0], 3 * n);
    o.fv("bn", &s.bn_[0][0], 3 * n);
  }

  void fieldUpdate() override {
    Solver& s = *this;
    if (clock_) {
      MPI_Barrier(MPI_COMM_WORLD);
      if (clockFile_) std::fprintf(clockFile_, "%u %.9f\n", s.nTime_, MPI_Wtime());
    }
    bool dump = steps_.count(s.nTime_) > 0;
    if (dump) dumpPre();
    if (maxSteps_ >= 0 && (long)s.nTime_ >= maxSteps_) {
      if (clockFile_) std::fclose(clockFile_);
      std::fflush(nullptr);
      MPI_Finalize();
      std::exit(0);
    }
    Base::fieldUpdate();
    if (dump) dumpPost();
  }
};

int main(int argc, char** argv) {
  MPI_Init(&argc, &argv);
  if (argc < 2) { std::fprintf(stderr, "usage: ref_harness <job-file>\n"); return 2; }
  std::list<std::string> jobFile = read_file(argv[1]);
  cleanJobFile(jobFile);
  Mesh mesh; mesh.initialize();
  Bunch bunch; Seed seed;
  std::vector<Undulator> undulator;
  std::vector<ExtField> extField;
  std::vector<FreeElectronLaser> FEL;
  ParseDarius parser(jobFile, mesh, bunch, seed, undulator, extField, FEL);
  parser.setJobParameters();
  Solver* solver;
  if (mesh.spaceCharge_) solver = new Harness<FdTdSC>(mesh, bunch, seed, undulator, extField, FEL);
  else                   solver = new Harness<FdTd>(mesh, bunch, seed, undulator, extField, FEL);
  solver->solve();
  MPI_Finalize();
  return 0;
}
