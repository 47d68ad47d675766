// Set-up and main loop of the drop-in Solver (reference src/solver.cpp, src/radiation.cpp).
#include "solver.h"

#include <sys/stat.h>
#include <sys/time.h>

#include <algorithm>
#include <cerrno>
#include <cstring>
#include <iostream>
#include <sstream>

namespace mithra_host {

namespace {

void say(bool quiet, const std::string& msg) {
  if (!quiet) std::cout << msg << std::endl;
}

std::string str(Double v) {
  std::ostringstream o;
  o << v;
  return o.str();
}

void createDirectory(const std::string& filename) {  // src/stdinclude.cpp:22-35
  size_t found = filename.find_last_of('/');
  std::string path = filename.substr(0, found + 1);
  struct stat st;
  if (path.empty() || stat(path.c_str(), &st) == 0) return;
  // EEXIST: another job set up at the same time created it between the test and here
  if (mkdir(path.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH) == -1 && errno != EEXIST)
    throw JobError("Could not create the directory " + path + ". Probably the given address does not exist.");
}

}  // namespace

Solver::Solver(Mesh& mesh, Bunch& bunch, Seed& seed, std::vector<Undulator>& undulator, std::vector<ExtField>& extField,
               std::vector<FreeElectronLaser>& FEL)
    : mesh_(mesh), bunch_(bunch), seed_(seed), undulator_(undulator), extField_(extField), FEL_(FEL) {
  // src/solver.cpp:59-61
  c0_ = C0 / mesh_.lengthScale_ * mesh_.timeScale_;
  m0_ = MU_ZERO / mesh_.lengthScale_;
  e0_ = 1.0 / (c0_ * c0_ * m0_);
  std::memset(a_, 0, sizeof a_);
}

Solver::~Solver() {
  if (dev_) mithra_dev_destroy(dev_);
  if (dbunch_ && ownsDbunch_) mithra_bunch_destroy(dbunch_);
}

void Solver::check(int status, const char* what) const {
  if (status != MITHRA_OK) throw JobError(std::string(what) + ": " + mithra_strerror(status));
}

void Solver::setSimulationParameters() {
  seed_.c0_ = c0_;
  seed_.signal_.t0_ /= c0_;
  seed_.signal_.f0_ *= c0_;
  seed_.signal_.s_ /= c0_;
  seed_.l_ = c0_ / seed_.signal_.f0_;
  seed_.zR_.assign(2, 0.0);
  seed_.zR_[0] = PI * seed_.radius_[0] * seed_.radius_[0] / seed_.l_;
  seed_.zR_[1] = PI * seed_.radius_[1] * seed_.radius_[1] / seed_.l_;
  for (auto& u : undulator_) {
    u.c0_ = c0_;
    u.signal_.t0_ /= c0_;
    u.signal_.f0_ *= c0_;
    u.signal_.s_ /= c0_;
    u.l_ = c0_ / u.signal_.f0_;
    u.zR_.assign(2, 0.0);
    u.zR_[0] = PI * u.radius_[0] * u.radius_[0] / u.l_;
    u.zR_[1] = PI * u.radius_[1] * u.radius_[1] / u.l_;
  }
  for (auto& e : extField_) {
    e.c0_ = c0_;
    e.signal_.t0_ /= c0_;
    e.signal_.f0_ *= c0_;
    e.signal_.s_ /= c0_;
    e.l_ = c0_ / e.signal_.f0_;
    e.zR_.assign(2, 0.0);
    e.zR_[0] = PI * e.radius_[0] * e.radius_[0] / e.l_;
    e.zR_[1] = PI * e.radius_[1] * e.radius_[1] / e.l_;
  }
  seed_.amplitude_ = seed_.a0_ * EM * c0_ / EC;
  for (auto& u : undulator_) u.amplitude_ = u.a0_ * EM * c0_ * 2 * PI * u.signal_.f0_ / EC;
  for (auto& e : extField_) e.amplitude_ = e.a0_ * EM * c0_ * 2 * PI * e.signal_.f0_ / EC;

  Double gamma = 0.0;
  for (auto& b : bunch_.bunchInit_) {
    if (b.bunchType_ == "file") computeFileGamma(b);
    gamma += b.initialGamma_ / bunch_.bunchInit_.size();
  }
  Double gmin = 1.0e100, gmax = -1.0e100, g = 0.0;
  for (auto& u : undulator_) {
    if (u.type_ == STATIC) {
      g = gamma / std::sqrt(1.0 + u.k_ * u.k_ / 2.0);
      gmin = (gmin < g) ? gmin : g;
      gmax = (gmax > g) ? gmax : g;
    } else if (u.signal_.signalType_ == FLATTOP) {
      g = gamma / std::sqrt(1.0 + u.a0_ * u.a0_ / 2.0);
      gmin = (gmin < g) ? gmin : g;
      gmax = (gmax > g) ? gmax : g;
    } else {
      g = gamma / std::sqrt(1.0 + u.a0_ * u.a0_ / 2.0);
      gmin = (gmin < g) ? gmin : g;
      g = gamma;
      gmax = (gmax > g) ? gmax : g;
    }
  }
  if (mesh_.gamma_ == -1.0) gamma_ = (undulator_.size() == 0) ? gamma : (gmin + gmax) / 2.0;
  else gamma_ = mesh_.gamma_;
  beta_ = std::sqrt(1.0 - 1.0 / (gamma_ * gamma_));
  for (auto& u : undulator_)
    if (u.type_ == OPTICAL) u.lu_ /= (1 + beta_);
  std::sort(undulator_.begin(), undulator_.end(), [](const Undulator& i, const Undulator& j) { return i.rb_ < j.rb_; });
  for (auto it = undulator_.rbegin(); it != undulator_.rend(); ++it) it->rb_ -= undulator_[0].rb_;
  for (auto& b : bunch_.bunchInit_) {
    b.initialBeta_ = std::sqrt(1.0 - 1.0 / std::pow(b.initialGamma_, 2));
    for (int c = 0; c < 3; ++c) b.betaVector_[c] = b.initialBeta_ * b.initialDirection_[c];
    b.lambda_ = undulator_.size() > 0 ? undulator_[0].lu_ / (2.0 * gamma_ * gamma_) * b.betaVector_[2] / beta_ : 0.0;
  }
  seed_.beta_ = beta_;
  seed_.gamma_ = gamma_;
}

void Solver::computeFileGamma(BunchInitialize& b) {
  Double ignore;
  Vec3 gb;
  b.initialGamma_ = 0.0;
  b.initialDirection_ = Vec3();
  std::ifstream f(b.fileName_.c_str());
  while (f.good()) {
    f >> ignore >> ignore >> ignore >> gb[0] >> gb[1] >> gb[2];
    b.initialGamma_ += std::sqrt(1 + gb.norm2());
    for (int c = 0; c < 3; ++c) b.initialDirection_[c] += gb[c];
  }
  b.initialGamma_ /= b.numberOfParticles_;
  const Double n = b.initialDirection_.norm();
  for (int c = 0; c < 3; ++c) b.initialDirection_[c] /= n;
}

void Solver::deviceBunch() {
  // Solver::initializeBunch (src/solver.cpp:1130-1176) with the records on the device: every
  // ellipsoid initialization through the device generator, in the reference's order; a
  // handed-over `other` bunch is already there.
  if (!dbunch_) {
    check(mithra_bunch_create(&dbunch_), "device bunch");
    ownsDbunch_ = true;
  }
  for (auto& b : bunch_.bunchInit_) {
    if (b.position_.empty()) b.position_.push_back(Vec3());
    if (b.bunchType_ == "other") continue;
    if (b.bunchType_ != "ellipsoid")
      throw JobError("Device bunch generation takes ellipsoid and `other` bunches only, not " + b.bunchType_ + ".");
    if (b.generator_ == "random")
      throw JobError("Device bunch generation takes the halton generator only (random draws are not reproducible).");
    if (b.distribution_ != "uniform" && b.distribution_ != "gaussian")
      throw JobError("The longitudinal type is not correctly given to the code !!!");
    long first = 0;  // records of this initialization at earlier positions (the reference's Np0)
    for (size_t ia = 0; ia < b.position_.size(); ++ia) {
      mithra_ellipsoid e;
      std::memset(&e, 0, sizeof e);
      e.n = b.numberOfParticles_;
      e.first_index = (unsigned)first;
      e.distribution = b.distribution_ == "uniform" ? MITHRA_DISTRIBUTION_UNIFORM : MITHRA_DISTRIBUTION_GAUSSIAN;
      e.shot_noise = b.shotNoise_ ? 1 : 0;
      e.charge = b.cloudCharge_;
      e.gamma = b.initialGamma_;
      for (int k = 0; k < 3; ++k) {
        e.beta_vector[k] = b.betaVector_[k];
        e.position[k] = b.position_[ia][k];
        e.sigma_position[k] = b.sigmaPosition_[k];
        e.sigma_gb[k] = b.sigmaGammaBeta_[k];
      }
      e.tran_trunc = b.tranTrun_;
      e.long_trunc = b.longTrun_;
      e.lambda = b.lambda_;
      e.bunching = b.bF_;
      e.bunching_phase = b.bFP_;
      long added = 0;
      check(mithra_bunch_ellipsoid(dbunch_, &e, &added), "device ellipsoid bunch");
      first += added;
    }
  }
}

long Solver::particleCount() const {
  return deviceBunch_ ? mithra_bunch_count(dbunch_) : (long)chargeVectorn_.size();
}

void Solver::particles(double* q, long n) const {
  if (n != particleCount()) throw JobError("particle count mismatch");
  if (deviceBunch_) {
    check(mithra_bunch_get(dbunch_, q, n), "device bunch read-back");
    return;
  }
  long i = 0;
  for (const Charge& c : chargeVectorn_) {
    double* r = q + 11 * i++;
    r[0] = c.q;
    for (int k = 0; k < 3; ++k) {
      r[1 + k] = c.rnp[k];
      r[4 + k] = c.rnm[k];
      r[7 + k] = c.gb[k];
    }
    r[10] = c.e;
  }
}

void Solver::initializeBunch() {
  if (deviceBunch_) {
    deviceBunch();
    return;
  }
  std::list<Charge> qv;
  for (auto& b : bunch_.bunchInit_) {
    qv.clear();
    if (b.position_.empty()) b.position_.push_back(Vec3());
    const int np = (int)b.position_.size();
    if (b.bunchType_ == "manual") {
      for (int ia = 0; ia < np; ++ia) bunch_.initializeManual(b, qv, ia);
    } else if (b.bunchType_ == "ellipsoid") {
      for (int ia = 0; ia < np; ++ia) bunch_.initializeEllipsoid(b, qv, ia);
    } else if (b.bunchType_ == "3D-crystal") {
      for (int ia = 0; ia < np; ++ia) bunch_.initialize3DCrystal(b, qv, ia);
    } else if (b.bunchType_ == "file") {
      for (int ia = 0; ia < np; ++ia) bunch_.initializeFile(b, qv, ia);
    } else if (b.bunchType_ == "other") {
      for (const Charge& c : external_) qv.push_back(c);   // filled in by the calling program
    }
    chargeVectorn_.splice(chargeVectorn_.end(), qv);
  }
}

void Solver::lorentzBoostMesh() {
  mesh_.meshLength_[2] *= gamma_;
  mesh_.meshResolution_[2] *= gamma_;
  mesh_.meshCenter_[2] *= gamma_;
  mesh_.totalTime_ /= gamma_;
  mesh_.timeShift_ /= gamma_;
  if (mesh_.solver_ == NSFD) {
    Double t = 1.0 / std::sqrt(std::pow(mesh_.meshResolution_[2] / mesh_.meshResolution_[0], 2.0) +
                               std::pow(mesh_.meshResolution_[2] / mesh_.meshResolution_[1], 2.0));
    if (t < 1.02) {
      mesh_.meshResolution_[0] *= 1.02 / t;
      mesh_.meshResolution_[1] *= 1.02 / t;
    }
    mesh_.timeStep_ = mesh_.meshResolution_[2] / c0_;
  } else {
    mesh_.timeStep_ = 0.98 / (c0_ * std::sqrt(1.0 / std::pow(mesh_.meshResolution_[0], 2.0) +
                                              1.0 / std::pow(mesh_.meshResolution_[1], 2.0) +
                                              1.0 / std::pow(mesh_.meshResolution_[2], 2.0)));
  }
}

void Solver::initializeMesh() {
  N0_ = (int)(mesh_.meshLength_[0] / mesh_.meshResolution_[0]) + 2;
  N1_ = (int)(mesh_.meshLength_[1] / mesh_.meshResolution_[1]) + 2;
  N2_ = (int)(mesh_.meshLength_[2] / mesh_.meshResolution_[2]) + 2;
  N1N0_ = (long)N1_ * N0_;
  mesh_.meshLength_[0] = (N0_ - 1) * mesh_.meshResolution_[0];
  mesh_.meshLength_[1] = (N1_ - 1) * mesh_.meshResolution_[1];
  mesh_.meshLength_[2] = (N2_ - 1) * mesh_.meshResolution_[2];
  np_ = N2_;  // one device domain: the reference at mpirun -np 1
  k0_ = 0;
  const size_t n = (size_t)N1N0_ * np_;
  an_.assign(3 * n, 0.0);
  anm1_.assign(3 * n, 0.0);
  if (mesh_.spaceCharge_) {
    fn_.assign(n, 0.0);
    fnm1_.assign(n, 0.0);
  }
  xmin_ = mesh_.meshCenter_[0] - mesh_.meshLength_[0] / 2.0;
  xmax_ = mesh_.meshCenter_[0] + mesh_.meshLength_[0] / 2.0;
  ymin_ = mesh_.meshCenter_[1] - mesh_.meshLength_[1] / 2.0;
  ymax_ = mesh_.meshCenter_[1] + mesh_.meshLength_[1] / 2.0;
  zmin_ = mesh_.meshCenter_[2] - mesh_.meshLength_[2] / 2.0;
  zmax_ = mesh_.meshCenter_[2] + mesh_.meshLength_[2] / 2.0;
  zp_[0] = zmin_ + (unsigned)(0 + k0_) * mesh_.meshResolution_[2];
  zp_[1] = zmin_ + (unsigned)(np_ - 1 + k0_) * mesh_.meshResolution_[2];
  timep1_ = mesh_.timeStep_;
  time_ = 0.0;
  timem1_ = -mesh_.timeStep_;
}

void Solver::lorentzBoostBunch() {
  bunch_.timeStep_ /= gamma_;
  if (bunch_.timeStep_ == 0) bunch_.timeStep_ = mesh_.timeStep_;
  else bunch_.timeStep_ = mesh_.timeStep_ / std::ceil(mesh_.timeStep_ / bunch_.timeStep_);
  nUpdateBunch_ = mesh_.timeStep_ / bunch_.timeStep_;
  bunch_.rhythm_ /= gamma_;
  bunch_.bunchVTKRhythm_ /= gamma_;
  for (auto& t : bunch_.bunchProfileTime_) t /= gamma_;
  bunch_.bunchProfileRhythm_ /= gamma_;

  if (deviceBunch_) {
    deviceBoost();
    return;
  }
  Double zmaxG = -1.0e100;
  for (auto& q : chargeVectorn_) {
    Double g = std::sqrt(1.0 + q.gb.norm2());
    Double bz = q.gb[2] / g;
    q.rnp[2] *= gamma_;
    q.gb[2] = gamma_ * g * (bz - beta_);
    zmaxG = std::max(zmaxG, q.rnp[2]);
  }
  dt_ = 0.0;
  if (undulator_.size() > 0) {
    Double nl = (undulator_[0].type_ == STATIC) ? 2.0 : 5.0 * undulator_[0].signal_.nR_;
    if (undulator_[0].dist_ == 0.0) undulator_[0].dist_ = nl * undulator_[0].lu_;
    else if (undulator_[0].dist_ < nl * undulator_[0].lu_)
      std::cout << "Warning: the undulator is set very close to the bunch, the results may be inaccurate." << std::endl;
    dt_ = -1.0 / (beta_ * undulator_[0].c0_) * (zmaxG + undulator_[0].dist_ / gamma_);
  }
  seed_.dt_ = dt_;
  bunch_.zu_ = zmaxG;
  bunch_.beta_ = beta_;
  for (auto& q : chargeVectorn_) {
    Double g = std::sqrt(1.0 + q.gb.norm2());
    q.rnp[0] += q.gb[0] / g * (q.rnp[2] - bunch_.zu_) * beta_;
    q.rnp[1] += q.gb[1] / g * (q.rnp[2] - bunch_.zu_) * beta_;
    q.rnp[2] += q.gb[2] / g * (q.rnp[2] - bunch_.zu_) * beta_;
  }
  if (mesh_.optimizePosition_ && undulator_.size() > 0) {
    Double zG = 0.0, bzG = 0.0;
    for (auto& q : chargeVectorn_) {
      zG += q.rnp[2];
      bzG += q.gb[2] / std::sqrt(1 + q.gb.norm2());
    }
    unsigned int NqG = (unsigned int)chargeVectorn_.size();
    zG /= NqG;
    bzG /= NqG;
    Double shift = bzG * (zmaxG + undulator_[0].dist_ / gamma_ - zG) / (bzG + beta_) + zG;
    zmaxG -= shift;
    bunch_.zu_ = zmaxG;
    dt_ = -1.0 / (beta_ * undulator_[0].c0_) * (zmaxG + undulator_[0].dist_ / gamma_);
    seed_.dt_ = dt_;
    for (auto& q : chargeVectorn_) q.rnp[2] -= shift;
  }
  if (mesh_.totalDist_ > 0.0) {
    double Lu = 0.0;
    for (auto& u : undulator_) Lu += u.lu_ * u.length_ / gamma_;
    double zEnd = mesh_.totalDist_ / gamma_;
    double zMin = 1e100;
    double bz = 0;
    for (auto& q : chargeVectorn_) {
      zMin = std::min(zMin, q.rnp[2]);
      bz += q.gb[2] / std::sqrt(1 + q.gb.norm2());
    }
    unsigned int Nq = (unsigned int)chargeVectorn_.size();
    bz /= Nq;
    mesh_.totalTime_ = 1 / (c0_ * (bz + beta_)) * (zEnd - beta_ * c0_ * dt_ - zMin + bz / beta_ * Lu);
  }
}

void Solver::deviceBoost() {
  // The rest of lorentzBoostBunch above with its particle loops on the device (the time-step
  // scaling at its start has run); the sums of the means are fixed-order per-workgroup partials.
  Double zmaxG = -1.0e100;
  check(mithra_bunch_boost(dbunch_, gamma_, beta_, &zmaxG), "device boost");
  dt_ = 0.0;
  if (undulator_.size() > 0) {
    Double nl = (undulator_[0].type_ == STATIC) ? 2.0 : 5.0 * undulator_[0].signal_.nR_;
    if (undulator_[0].dist_ == 0.0) undulator_[0].dist_ = nl * undulator_[0].lu_;
    else if (undulator_[0].dist_ < nl * undulator_[0].lu_)
      std::cout << "Warning: the undulator is set very close to the bunch, the results may be inaccurate." << std::endl;
    dt_ = -1.0 / (beta_ * undulator_[0].c0_) * (zmaxG + undulator_[0].dist_ / gamma_);
  }
  seed_.dt_ = dt_;
  bunch_.zu_ = zmaxG;
  bunch_.beta_ = beta_;
  check(mithra_bunch_drift(dbunch_, beta_, bunch_.zu_), "device drift");
  const long Nq = mithra_bunch_count(dbunch_);
  if (mesh_.optimizePosition_ && undulator_.size() > 0 && Nq > 0) {
    Double zG = 0.0, bzG = 0.0, zlo = 0.0;
    check(mithra_bunch_moments(dbunch_, &zG, &bzG, &zlo), "device moments");
    zG /= (unsigned int)Nq;
    bzG /= (unsigned int)Nq;
    Double shift = bzG * (zmaxG + undulator_[0].dist_ / gamma_ - zG) / (bzG + beta_) + zG;
    zmaxG -= shift;
    bunch_.zu_ = zmaxG;
    dt_ = -1.0 / (beta_ * undulator_[0].c0_) * (zmaxG + undulator_[0].dist_ / gamma_);
    seed_.dt_ = dt_;
    check(mithra_bunch_shift(dbunch_, shift), "device shift");
  }
  if (mesh_.totalDist_ > 0.0 && Nq > 0) {
    double Lu = 0.0;
    for (auto& u : undulator_) Lu += u.lu_ * u.length_ / gamma_;
    double zEnd = mesh_.totalDist_ / gamma_;
    double zs = 0.0, bz = 0.0, zMin = 1e100;
    check(mithra_bunch_moments(dbunch_, &zs, &bz, &zMin), "device moments");
    bz /= (unsigned int)Nq;
    mesh_.totalTime_ = 1 / (c0_ * (bz + beta_)) * (zEnd - beta_ * c0_ * dt_ - zMin + bz / beta_ * Lu);
  }
}

void Solver::initializeField() {
  const Double dt = mesh_.timeStep_, dx = mesh_.meshResolution_[0], dy = mesh_.meshResolution_[1],
               dz = mesh_.meshResolution_[2];
  dv_ = -m0_ * EC / mesh_.timeStep_ / (dx * dy * dz);
  rc_ = -EC / e0_ / (dx * dy * dz);
  const Double beta = (1.0 + 0.02 / (std::pow(dz / dx, 2.0) + std::pow(dz / dy, 2.0))) / 4.0;
  const Double alpha = 1.0 - 2.0 * beta;
  alpha_ = alpha;
  betaNS_ = beta / alpha;
  a_[0] = 2.0 * (1.0 - alpha * std::pow(c0_ * dt / dx, 2) - alpha * std::pow(c0_ * dt / dy, 2) - std::pow(c0_ * dt / dz, 2));
  a_[1] = std::pow(c0_ * dt / dx, 2.0);
  a_[2] = std::pow(c0_ * dt / dy, 2.0);
  a_[3] = std::pow(c0_ * dt / dz, 2.0) - 2.0 * (beta * std::pow(c0_ * dt / dx, 2) + beta * std::pow(c0_ * dt / dy, 2));
  a_[4] = std::pow(c0_ * dt, 2.0) * dv_;
  a_[5] = std::pow(c0_ * dt, 2.0) * rc_;
  const Double to = mesh_.truncationOrder_ - 1.0;
  Double alpha1 = 0.0, alpha2 = 0.0;
  Double p = (1.0 + std::cos(alpha1) * std::cos(alpha2)) / (std::cos(alpha1) + std::cos(alpha2));
  Double q = -1.0 / (std::cos(alpha1) + std::cos(alpha2));
  // Mur faces (src/solver.cpp:745-781): normal h, tangential (u, v)
  auto face = [&](Double* B, Double h, Double u, Double v) {
    Double d = 1.0 / (2.0 * dt * h) + p / (2.0 * c0_ * dt * dt);
    B[0] = (1.0 / (2.0 * dt * h) - p / (2.0 * c0_ * dt * dt)) / d;
    B[1] = (-1.0 / (2.0 * dt * h) - p / (2.0 * c0_ * dt * dt)) / d;
    B[2] = (p / (c0_ * dt * dt) + q * to * (c0_ / (u * u) + c0_ / (v * v))) / d;
    B[3] = -q * to * (c0_ / (2.0 * u * u)) / d;
    B[4] = -q * to * (c0_ / (2.0 * v * v)) / d;
  };
  face(bB_, dx, dy, dz);
  face(cB_, dy, dx, dz);
  face(dB_, dz, dx, dy);
  // Mur edges (src/solver.cpp:784-806): the edge runs along `w`, (h1, h2) are the two normals
  auto edge = [&](Double* E, Double h1, Double h2, Double w) {
    Double d = (1.0 / h1 + 1.0 / h2) / (4.0 * dt) + 3.0 / (8.0 * c0_ * dt * dt);
    E[0] = (-(1.0 / h1 - 1.0 / h2) / (4.0 * dt) - 3.0 / (8.0 * c0_ * dt * dt)) / d;
    E[1] = ((1.0 / h1 - 1.0 / h2) / (4.0 * dt) - 3.0 / (8.0 * c0_ * dt * dt)) / d;
    E[2] = ((1.0 / h1 + 1.0 / h2) / (4.0 * dt) - 3.0 / (8.0 * c0_ * dt * dt)) / d;
    E[3] = (3.0 / (4.0 * c0_ * dt * dt) - c0_ / (4.0 * w * w)) / d;
    E[4] = c0_ / (8.0 * w * w) / d;
  };
  edge(eE_, dy, dx, dz);
  edge(fE_, dz, dy, dx);
  edge(gE_, dx, dz, dy);
  // Mur corners (src/solver.cpp:809-824)
  const Double sx[8] = {-1, 1, -1, -1, 1, 1, -1, 1}, sy[8] = {-1, -1, 1, -1, 1, -1, 1, 1},
               sz[8] = {-1, -1, -1, 1, -1, 1, 1, 1};
  for (int t = 0; t < 8; ++t) {
    hC_[t] = (sx[t] / dx + sy[t] / dy + sz[t] / dz) / (8.0 * dt) - 1.0 / (4.0 * c0_ * dt * dt);
    hC_[8 + t] = -(sx[t] / dx + sy[t] / dy + sz[t] / dz) / (8.0 * dt) - 1.0 / (4.0 * c0_ * dt * dt);
  }
  hC_[16] = 1.0 / (2.0 * c0_ * dt * dt);
  // seeded initial potentials inside the TF region (src/solver.cpp:828-839)
  if (seed_.amplitude_ > 1.0e-50) {
    for (int i = 2; i < N0_ - 2; i++)
      for (int j = 2; j < N1_ - 2; j++)
        for (int k = 2; k < np_ - 2; k++) {
          const long m = N1N0_ * k + (long)N1_ * i + j;
          Vec3 r(xmin_ + (unsigned)i * mesh_.meshResolution_[0], ymin_ + (unsigned)j * mesh_.meshResolution_[1],
                 zmin_ + (unsigned)(k + k0_) * mesh_.meshResolution_[2]);
          Vec3 a(an_[3 * m], an_[3 * m + 1], an_[3 * m + 2]);
          seed_.fields(r, time_, a);
          for (int c = 0; c < 3; ++c) an_[3 * m + c] = a[c];
          Vec3 b(anm1_[3 * m], anm1_[3 * m + 1], anm1_[3 * m + 2]);
          seed_.fields(r, timem1_, b);
          for (int c = 0; c < 3; ++c) anm1_[3 * m + c] = b[c];
        }
  }
}

void Solver::initializeBunchUpdate() {
  dtb_ = c0_ * bunch_.timeStep_;
  r1_ = -EC / (EM * c0_) * bunch_.timeStep_ / 2.0;
  r2_ = -EC / EM * bunch_.timeStep_ / 2.0;
  initializeBunchDiagnostics();
}

void Solver::initializePowerSample() {
  for (unsigned jf = 0; jf < FEL_.size(); jf++) {
    auto& R = FEL_[jf].radiationPower_;
    if (!R.sampling_) continue;
    if (powerFEL_ >= 0) throw JobError("More than one radiation-power block is not supported by this build.");
    powerFEL_ = (int)jf;
    for (auto& z : R.z_) z *= gamma_;
    R.lineBegin_ *= gamma_;
    R.lineEnd_ *= gamma_;
    Double dl = std::fabs(R.lineEnd_ - R.lineBegin_) / R.res_;
    if (R.samplingType_ == OVERLINE) {
      Double l = 0.0;
      while (std::fabs(l) < std::fabs(R.lineEnd_ - R.lineBegin_)) {
        R.z_.push_back(R.lineBegin_ + l);
        l += dl;
      }
    }
    pN_ = (unsigned)R.z_.size();
    dl = (R.lambdaMax_ - R.lambdaMin_) / R.lambdaRes_;
    for (Double rw = R.lambdaMin_; rw < R.lambdaMax_; rw += dl) R.lambda_.push_back(rw);
    pNl_ = (unsigned)R.lambda_.size();
    pNf_ = 0;
    pfile_.clear();
    pw_.assign(pNl_, 0.0);
    for (unsigned i = 0; i < pNl_; i++) {
      std::string base = R.basename_.compare(0, 1, "/") == 0 ? "" : R.directory_;
      base += R.basename_ + "-" + std::to_string(i) + ".txt";
      createDirectory(base);
      pfile_.emplace_back(new std::ofstream(base.c_str(), std::ios::trunc));
      pfile_.back()->setf(std::ios::scientific);
      pfile_.back()->precision(15);
      pfile_.back()->width(40);
      Double dt = undulator_[0].lu_ / R.lambda_[i] / (gamma_ * c0_);
      pNf_ = (unsigned(3.0 * dt / mesh_.timeStep_) > pNf_) ? unsigned(3.0 * dt / mesh_.timeStep_) : pNf_;
      pw_[i] = 2 * PI / dt;
    }
    pc_ = 2.0 * mesh_.meshResolution_[0] * mesh_.meshResolution_[1] / (m0_ * pNf_ * pNf_) *
          std::pow(mesh_.lengthScale_, 2) / std::pow(mesh_.timeScale_, 3);
  }
}

void Solver::shiftBackInTime() {
  if (mesh_.timeShift_ == 0.0) return;
  timem1_ -= mesh_.timeShift_;
  time_ -= mesh_.timeShift_;
  timep1_ -= mesh_.timeShift_;
  timeBunch_ -= mesh_.timeShift_;
  if (deviceBunch_) {
    check(mithra_bunch_time_shift(dbunch_, c0_, mesh_.timeShift_), "device time shift");
    return;
  }
  for (auto& q : chargeVectorn_) {
    Double t = c0_ * mesh_.timeShift_ / std::sqrt(1.0 + q.gb.norm2());
    for (int c = 0; c < 3; ++c) q.rnp[c] -= t * q.gb[c];
  }
}

void Solver::initialize() {
  if (initialized_) return;
  setSimulationParameters();
  initializeBunch();
  timeBunch_ = time_;
  lorentzBoostMesh();
  initializeMesh();
  lorentzBoostBunch();
  initializeField();
  if (seed_.sampling_) initializeSeedSampling();
  initializeSeedVTK();
  if (seed_.profile_) initializeSeedProfile();
  initializeBunchUpdate();
  initializePowerSample();
  initializePowerVisualize();
  initializeScreenProfile();
  shiftBackInTime();
  initialized_ = true;
}

int Solver::nSubSteps() const {
  int n = 0;
  for (Double t = 0.0; t < nUpdateBunch_; t += 1.0) ++n;
  return n;
}

mithra_grid Solver::grid() const {
  mithra_grid g;
  std::memset(&g, 0, sizeof g);
  g.N0 = N0_; g.N1 = N1_; g.N2 = np_;
  g.solver = mesh_.solver_ == NSFD ? MITHRA_SOLVER_NSFD : MITHRA_SOLVER_FD;
  g.truncation_order = (int)mesh_.truncationOrder_;
  g.space_charge = mesh_.spaceCharge_ ? 1 : 0;
  g.xmin = xmin_; g.xmax = xmax_; g.ymin = ymin_; g.ymax = ymax_; g.zmin = zmin_; g.zmax = zmax_;
  g.zp0 = zp_[0]; g.zp1 = zp_[1];
  g.dx = mesh_.meshResolution_[0]; g.dy = mesh_.meshResolution_[1]; g.dz = mesh_.meshResolution_[2];
  g.dt = mesh_.timeStep_;
  g.Lz = mesh_.meshLength_[2];
  return g;
}

mithra_coeffs Solver::coeffs() const {
  mithra_coeffs c;
  std::memcpy(c.a, a_, sizeof a_);
  std::memcpy(c.bB, bB_, sizeof bB_); std::memcpy(c.cB, cB_, sizeof cB_); std::memcpy(c.dB, dB_, sizeof dB_);
  std::memcpy(c.eE, eE_, sizeof eE_); std::memcpy(c.fE, fE_, sizeof fE_); std::memcpy(c.gE, gE_, sizeof gE_);
  std::memcpy(c.hC, hC_, sizeof hC_);
  c.alpha = alpha_;
  c.beta = betaNS_;
  return c;
}

mithra_seed Solver::seed() const {
  mithra_seed s;
  std::memset(&s, 0, sizeof s);
  s.seed_type = (int)seed_.seedType_;
  s.signal_type = (int)seed_.signal_.signalType_;
  s.order[0] = seed_.order_[0]; s.order[1] = seed_.order_[1];
  s.nR = seed_.signal_.nR_;
  s.amplitude = seed_.amplitude_;
  for (int c = 0; c < 3; ++c) {
    s.position[c] = seed_.position_[c];
    s.direction[c] = seed_.direction_[c];
    s.polarization[c] = seed_.polarization_[c];
  }
  s.radius[0] = seed_.radius_[0]; s.radius[1] = seed_.radius_[1];
  s.t0 = seed_.signal_.t0_; s.s = seed_.signal_.s_; s.f0 = seed_.signal_.f0_; s.cep = seed_.signal_.cep_;
  if (seed_.signal_.sigmaInvG_.size() >= 2) {
    s.sigma_invg[0] = seed_.signal_.sigmaInvG_[0];
    s.sigma_invg[1] = seed_.signal_.sigmaInvG_[1];
  }
  s.gamma = seed_.gamma_; s.beta = seed_.beta_; s.dt_shift = seed_.dt_; s.c0 = seed_.c0_;
  return s;
}

mithra_bunch_params Solver::bunchParams() const {
  mithra_bunch_params p;
  p.gamma = gamma_; p.beta = beta_; p.c0 = c0_; p.dt_shift = dt_;
  p.r1 = r1_; p.r2 = r2_; p.dtb = dtb_;
  return p;
}

std::vector<mithra_undulator> Solver::undulators() const {
  std::vector<mithra_undulator> v;
  for (auto& u : undulator_) {
    mithra_undulator d;
    d.type = u.type_ == STATIC ? MITHRA_UNDULATOR_STATIC : MITHRA_UNDULATOR_OPTICAL;
    d.length = u.length_;
    d.k = u.k_; d.lu = u.lu_; d.rb = u.rb_; d.dist = u.dist_; d.theta = u.theta_;
    v.push_back(d);
  }
  return v;
}

void Solver::fill(mithra_dev* dev) {
  // The pre-step state of Solver::initialize in a domain: the potentials when the seed set them
  // (src/solver.cpp:828-839; a new domain holds zeros otherwise), and the particles, handed over
  // on the device when they were generated there.
  if (seed_.amplitude_ > 1.0e-50) {
    check(mithra_dev_set_potentials(dev, 0, an_.data(), mesh_.spaceCharge_ ? fn_.data() : nullptr), "A^n upload");
    check(mithra_dev_set_potentials(dev, 1, anm1_.data(), mesh_.spaceCharge_ ? fnm1_.data() : nullptr),
          "A^n-1 upload");
  }
  setBeams(dev);
  if (deviceBunch_) {
    check(mithra_dev_take_bunch(dev, dbunch_), "particle hand-over");
    return;
  }
  std::vector<double> q(chargeVectorn_.size() * 11);
  particles(q.data(), (long)chargeVectorn_.size());
  check(mithra_dev_set_particles(dev, q.data(), (long)chargeVectorn_.size()), "particle upload");
}

namespace {

template <typename B>
mithra_beam beamOf(const B& b) {
  mithra_beam m;
  std::memset(&m, 0, sizeof m);
  m.beam_type = (int)b.seedType_;
  m.signal_type = (int)b.signal_.signalType_;
  m.order[0] = b.order_[0];
  m.order[1] = b.order_[1];
  m.nR = b.signal_.nR_;
  m.amplitude = b.amplitude_;
  for (int c = 0; c < 3; ++c) {
    m.position[c] = b.position_[c];
    m.direction[c] = b.direction_[c];
    m.polarization[c] = b.polarization_[c];
  }
  m.radius[0] = b.radius_[0];
  m.radius[1] = b.radius_[1];
  m.zR[0] = b.zR_[0];
  m.zR[1] = b.zR_[1];
  m.l = b.l_;
  m.t0 = b.signal_.t0_;
  m.s = b.signal_.s_;
  m.f0 = b.signal_.f0_;
  m.cep = b.signal_.cep_;
  m.sigma_invg[0] = b.signal_.sigmaInvG_[0];
  m.sigma_invg[1] = b.signal_.sigmaInvG_[1];
  return m;
}

}  // namespace

void Solver::setBeams(mithra_dev* dev) {
  // the laser beams of optical undulators (table order) and of EXTERNAL-FIELD waves
  std::vector<mithra_beam> optical, external;
  for (auto& u : undulator_)
    if (u.type_ == OPTICAL) optical.push_back(beamOf(u));
  for (auto& e : extField_) external.push_back(beamOf(e));
  if (optical.empty() && external.empty()) return;
  check(mithra_dev_set_beams(dev, optical.data(), (int)optical.size(), external.data(), (int)external.size()),
        "laser beams");
}

void Solver::upload(int precision) {
  if (!initialized_) initialize();
  if (dev_) {
    mithra_dev_destroy(dev_);
    dev_ = nullptr;
  }
  mithra_grid g = grid();
  mithra_coeffs c = coeffs();
  check(mithra_dev_create(&g, &c, precision, &dev_), "device allocation");
  // E/B only where the bunch and the power sampler read them, as the reference evaluates it
  check(mithra_dev_set_lazy_fields(dev_, 1), "lazy field evaluation");
  fill(dev_);
}

void Solver::writePowerRows(const std::vector<double>& rows, const std::vector<double>& tb, unsigned count) {
  // src/radiation.cpp:222-229: per frequency file, one line per step: "z_lab \t P \t" per plane
  const auto& R = FEL_[powerFEL_].radiationPower_;
  for (unsigned r = 0; r < count; ++r)
    for (unsigned l = 0; l < pNl_; ++l) {
      std::ofstream& f = *pfile_[l];
      for (unsigned k = 0; k < pN_; ++k)
        f << gamma_ * (R.z_[k] + beta_ * c0_ * (tb[r] + dt_)) << "\t" << rows[((size_t)r * pN_ + k) * pNl_ + l] << "\t";
      f << std::endl;
    }
}

long Solver::solve(long max_steps, int precision, bool quiet) {
  initialize();
  upload(precision);
  const mithra_seed sd = seed();
  const mithra_bunch_params bp = bunchParams();
  const std::vector<mithra_undulator> und = undulators();
  const int nsub = nSubSteps();
  const unsigned kRows = 256;
  std::vector<double> rows, tbs;
  if (powerFEL_ >= 0) {
    const auto& R = FEL_[powerFEL_].radiationPower_;
    mithra_power_params pp;
    pp.N = pN_; pp.Nl = pNl_; pp.Nf = pNf_; pp.pc = pc_;
    pp.gamma = gamma_; pp.beta = beta_; pp.c0 = c0_;
    pp.z = R.z_.data();
    pp.w = pw_.data();
    check(mithra_dev_power_setup(dev_, &pp, kRows), "power sampling set-up");
    rows.resize((size_t)kRows * pN_ * pNl_);
  }
  unsigned pending = 0;
  auto flush = [&]() {
    if (!pending) return;
    check(mithra_dev_power_fetch(dev_, 0, pending, rows.data()), "power fetch");
    writePowerRows(rows, tbs, pending);
    pending = 0;
    tbs.clear();
  };
  const bool diag = diagnosticsActive();
  powerVisualizeSetup();
  timeval start, now;
  gettimeofday(&start, nullptr);
  // Particle-only march before time zero when a back shift is set (src/solver.cpp:1232-1291).
  // E/B are evaluated once from the initial potentials, which is what the reference's lazy
  // evaluation sees during this phase.
  bool evaluated = false;
  while (time_ < 0.0) {
    if (!evaluated) {
      check(mithra_dev_field_evaluate(dev_), "field evaluation");
      evaluated = true;
    }
    check(mithra_dev_bunch_advance(dev_, &bp, und.data(), (int)und.size(), timeBunch_, bunch_.timeStep_, nsub), "push");
    for (int s = 0; s < nsub; ++s) { timeBunch_ += bunch_.timeStep_; ++nTimeBunch_; }
    if (diag) {
      screenProfile();
      bunchWriters();
    }
    timem1_ += mesh_.timeStep_;
    time_ += mesh_.timeStep_;
    timep1_ += mesh_.timeStep_;
    ++nTime_;
  }
  long steps = 0;
  while (time_ < mesh_.totalTime_ && (max_steps < 0 || steps < max_steps)) {
    // a step whose field writers fire evaluates E/B on every node, not only where the bunch reads
    const bool fieldDue = diag && fieldWritersDue();
    if (fieldDue) check(mithra_dev_set_lazy_fields(dev_, 0), "field evaluation mode");
    check(mithra_dev_field_update(dev_, time_, &sd), "field update");
    if (fieldDue) check(mithra_dev_set_lazy_fields(dev_, 1), "field evaluation mode");
    check(mithra_dev_bunch_advance(dev_, &bp, und.data(), (int)und.size(), timeBunch_, bunch_.timeStep_, nsub), "push");
    for (int s = 0; s < nsub; ++s) { timeBunch_ += bunch_.timeStep_; ++nTimeBunch_; }
    if (diag) {
      if (fieldDue) fieldWriters();
      bunchWriters();
      screenProfile();
    }
    if (powerFEL_ >= 0) {
      check(mithra_dev_power_sample(dev_, nTime_, pending), "power sample");
      tbs.push_back(timeBunch_);
      if (++pending == kRows) flush();
    }
    if (diag) powerVisualize();
    check(mithra_dev_field_shift(dev_), "field shift");
    check(mithra_dev_current_reset(dev_), "current reset");
    check(mithra_dev_current_update(dev_), "current update");
    timem1_ += mesh_.timeStep_;
    time_ += mesh_.timeStep_;
    timep1_ += mesh_.timeStep_;
    ++nTime_;
    ++steps;
    if (!quiet && int(time_ / mesh_.totalTime_ * 1000.0) != int(timem1_ / mesh_.totalTime_ * 1000.0)) {
      gettimeofday(&now, nullptr);
      double dtw = (now.tv_usec - start.tv_usec) / 1.0e6 + (now.tv_sec - start.tv_sec);
      std::cout << " Percentage of the total simulation completed (%)      = " << time_ / mesh_.totalTime_ * 100.0 << "\n"
                << " Average calculation time for each time step (s) = " << dtw / (double)nTime_ << std::endl;
    }
  }
  flush();
  check(mithra_dev_synchronize(dev_), "synchronize");
  long nt = 0, nc = 0;
  check(mithra_dev_bunch_counters(dev_, &nt, &nc), "counters");
  if (nt > 0)
    say(quiet, "Warning: " + std::to_string(nt) +
                   " particle sub-steps had transverse dimensions that are larger than the computational domain size.");
  if (nc > 0)
    say(quiet, "Warning: " + std::to_string(nc) +
                   " particle sub-steps crossed the periodic z boundary of the domain (the reference duplicates such "
                   "particles on one rank; this build does not).");
  for (auto& f : pfile_) f->flush();
  if (sfFile_) sfFile_->flush();
  if (sbFile_) sbFile_->flush();
  for (auto& v : scrFiles_)
    for (auto& f : v) f->flush();
  return steps;
}

}  // namespace mithra_host
