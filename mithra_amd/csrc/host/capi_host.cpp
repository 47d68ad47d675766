// C-ABI of the drop-in host (include/mithra_host.h).
#include <cstring>
#include <memory>
#include <string>

#include "mithra_host.h"
#include "solver.h"

using namespace mithra_host;

struct mithra_job {
  Mesh mesh;
  Bunch bunch;
  Seed seed;
  std::vector<Undulator> undulator;
  std::vector<ExtField> extField;
  std::vector<FreeElectronLaser> FEL;
  std::unique_ptr<Solver> solver;
  std::string error;
};

namespace {

template <typename F>
int guarded(mithra_job* job, F&& f) {
  if (!job) return MITHRA_E_ARG;
  try {
    return f();
  } catch (const JobError& e) {
    job->error = e.what();
    return MITHRA_E_ARG;
  } catch (const std::exception& e) {
    job->error = e.what();
    return MITHRA_E_STATE;
  }
}

int need_init(mithra_job* job) {
  if (!job->solver || !job->solver->initialized_) {
    job->error = "mithra_job_initialize has not run";
    return MITHRA_E_STATE;
  }
  return MITHRA_OK;
}

}  // namespace

extern "C" {

int mithra_job_open(const char* path, mithra_job** out) {
  if (!path || !out) return MITHRA_E_ARG;
  std::unique_ptr<mithra_job> job(new mithra_job());
  int rc = guarded(job.get(), [&]() {
    std::list<std::string> lines = read_file(path);
    cleanJobFile(lines);
    ParseDarius parser(lines, job->mesh, job->bunch, job->seed, job->undulator, job->extField, job->FEL);
    parser.setJobParameters();
    if (job->mesh.spaceCharge_)
      job->solver.reset(new FdTdSC(job->mesh, job->bunch, job->seed, job->undulator, job->extField, job->FEL));
    else
      job->solver.reset(new FdTd(job->mesh, job->bunch, job->seed, job->undulator, job->extField, job->FEL));
    return MITHRA_OK;
  });
  *out = job.release();
  return rc;
}

int mithra_job_close(mithra_job* job) {
  delete job;
  return MITHRA_OK;
}

const char* mithra_job_error(const mithra_job* job) { return job ? job->error.c_str() : "no job"; }

int mithra_job_set_external_bunch(mithra_job* job, const double* c7, long n) {
  return guarded(job, [&]() {
    if (n < 0 || (n && !c7)) return MITHRA_E_ARG;
    job->solver->external_.resize(n);
    for (long p = 0; p < n; ++p) {
      Charge& c = job->solver->external_[p];
      c.q = c7[7 * p];
      for (int k = 0; k < 3; ++k) {
        c.rnp[k] = c7[7 * p + 1 + k];
        c.gb[k] = c7[7 * p + 4 + k];
      }
    }
    return MITHRA_OK;
  });
}

int mithra_job_initialize(mithra_job* job) {
  return guarded(job, [&]() {
    job->solver->initialize();
    return MITHRA_OK;
  });
}

#define VIEW(name, type, expr)                         \
  int name(mithra_job* job, type* out) {               \
    return guarded(job, [&]() {                        \
      int rc = need_init(job);                         \
      if (!rc && out) *out = job->solver->expr;        \
      return rc;                                       \
    });                                                \
  }
VIEW(mithra_job_grid, mithra_grid, grid())
VIEW(mithra_job_coeffs, mithra_coeffs, coeffs())
VIEW(mithra_job_seed, mithra_seed, seed())
VIEW(mithra_job_bunch_params, mithra_bunch_params, bunchParams())
#undef VIEW

int mithra_job_undulators(mithra_job* job, mithra_undulator* out, int max) {
  return guarded(job, [&]() {
    int rc = need_init(job);
    if (rc) return rc;
    std::vector<mithra_undulator> u = job->solver->undulators();
    for (int i = 0; i < (int)u.size() && i < max; ++i) out[i] = u[i];
    return (int)u.size();
  });
}

int mithra_job_times(mithra_job* job, double* t) {
  return guarded(job, [&]() {
    int rc = need_init(job);
    if (rc) return rc;
    Solver& s = *job->solver;
    t[0] = s.time_;
    t[1] = s.timeBunch_;
    t[2] = s.bunch_.timeStep_;
    t[3] = s.mesh_.timeStep_;
    t[4] = s.mesh_.totalTime_;
    t[5] = s.nUpdateBunch_;
    return s.nSubSteps();
  });
}

int mithra_job_power(mithra_job* job, double* p, int max) {
  return guarded(job, [&]() {
    int rc = need_init(job);
    if (rc) return rc;
    Solver& s = *job->solver;
    if (s.powerFEL_ < 0) return 0;
    const auto& R = s.FEL_[s.powerFEL_].radiationPower_;
    std::vector<double> v = {(double)s.pN_, (double)s.pNl_, (double)s.pNf_, s.pc_};
    v.insert(v.end(), R.z_.begin(), R.z_.end());
    v.insert(v.end(), s.pw_.begin(), s.pw_.end());
    for (int i = 0; i < (int)v.size() && i < max; ++i) p[i] = v[i];
    return 1;
  });
}

long mithra_job_particle_count(mithra_job* job) {
  if (!job || !job->solver) return MITHRA_E_ARG;
  return job->solver->particleCount();
}

int mithra_job_particles(mithra_job* job, double* q, long n) {
  return guarded(job, [&]() {
    int rc = need_init(job);
    if (rc) return rc;
    if (n != job->solver->particleCount()) return MITHRA_E_ARG;
    job->solver->particles(q, n);
    return MITHRA_OK;
  });
}

int mithra_job_fill_domain(mithra_job* job, mithra_dev* dev) {
  return guarded(job, [&]() {
    int rc = need_init(job);
    if (rc) return rc;
    if (!dev) return MITHRA_E_ARG;
    job->solver->fill(dev);
    return MITHRA_OK;
  });
}

int mithra_job_device_bunch(mithra_job* job, int on) {
  return guarded(job, [&]() {
    if ((on != 0 && on != 1) || job->solver->initialized_) return MITHRA_E_ARG;
    job->solver->deviceBunch_ = on != 0;
    return MITHRA_OK;
  });
}

int mithra_job_set_external_device_bunch(mithra_job* job, mithra_bunch* bunch) {
  return guarded(job, [&]() {
    if (!bunch || job->solver->initialized_ || job->solver->dbunch_) return MITHRA_E_ARG;
    job->solver->deviceBunch_ = true;
    job->solver->dbunch_ = bunch;
    job->solver->ownsDbunch_ = false;
    return MITHRA_OK;
  });
}

int mithra_job_potentials(mithra_job* job, int level, double* a, double* phi) {
  return guarded(job, [&]() {
    int rc = need_init(job);
    if (rc) return rc;
    Solver& s = *job->solver;
    if (level != 0 && level != 1) return MITHRA_E_ARG;
    const std::vector<Double>& A = level == 0 ? s.an_ : s.anm1_;
    const std::vector<Double>& F = level == 0 ? s.fn_ : s.fnm1_;
    if (a) std::memcpy(a, A.data(), sizeof(double) * A.size());
    if (phi && !F.empty()) std::memcpy(phi, F.data(), sizeof(double) * F.size());
    return MITHRA_OK;
  });
}

long mithra_job_run(mithra_job* job, long max_steps, int precision, int quiet) {
  long steps = 0;
  int rc = guarded(job, [&]() {
    steps = job->solver->solve(max_steps, precision, quiet != 0);
    return MITHRA_OK;
  });
  return rc ? rc : steps;
}

}  // extern "C"
