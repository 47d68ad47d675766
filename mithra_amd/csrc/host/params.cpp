// Job-file reader, grammar and bunch generators of the MITHRA drop-in host.
//
// Reader and value parsers: reference src/readdata.cpp:10-160.  Grammar: src/datainput.cpp:18-751
// (same groups, keys, defaults and error texts; an unknown key is fatal).  Generators, signals
// and the analytic seed: src/classes.cpp:85-855, src/stdinclude.cpp:45-93.
#include <algorithm>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <functional>
#include <iostream>

#include "params.h"

namespace mithra_host {

Vec3 vec3(const std::vector<Double>& v) {
  if (v.size() != 3) throw JobError("A field-vector with other than three parameters is not accepted.");
  return Vec3(v[0], v[1], v[2]);
}

Double halton(unsigned int i, unsigned int j) {
  if (i > 20) throw JobError(" dimension can not be larger than 20. ");
  static const unsigned int prime[20] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71};
  int p0, p, k, a;
  Double x = 0.0;
  p = prime[i < 20 ? i : 19];
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

Double pmod(Double a, Double b) {
  Double x = std::fmod(a, b);
  x += (x < 0.0) ? b : 0.0;
  return x;
}

/* ---------------------------------------------------------------- reader ---------------- */

std::list<std::string> read_file(const std::string& filename) {
  std::ifstream f(filename.c_str());
  std::list<std::string> lines;
  if (!f.is_open()) throw JobError("Unable to open file");
  std::string line;
  while (f.good()) {
    std::getline(f, line);
    lines.push_back(line);
  }
  return lines;
}

void cleanJobFile(std::list<std::string>& jobFile) {
  std::list<std::string> cleaned;
  for (std::string s : jobFile) {
    size_t hash = s.find('#');
    if (hash != std::string::npos) s.erase(hash);
    s.erase(std::remove_if(s.begin(), s.end(), [](char c) { return c == ' ' || c == '\t'; }), s.end());
    if (!s.empty()) cleaned.push_back(s);
  }
  jobFile.swap(cleaned);
}

namespace {

std::string parameterName(const std::string& line) { return line.substr(0, line.find('=')); }

std::string stringValue(const std::string& line) {
  std::string v = line.substr(line.find('=') + 1);
  if (v.compare(0, 1, "\"") == 0) {
    v.erase(v.find('"'), 1);
    size_t q = v.find('"');
    if (q != std::string::npos) v.erase(q, 1);
  }
  return v;
}

Double doubleValue(const std::string& line) { return std::atof(line.substr(line.find('=') + 1).c_str()); }

int intValue(const std::string& line) { return int(doubleValue(line)); }

bool boolValue(const std::string& line) {
  std::string v = line.substr(line.find('=') + 1);
  if (v == "true") return true;
  if (v == "false") return false;
  throw JobError("boolValue(std::string line) got unexpected input. Input should be \"true\" or \"false\" ");
}

template <typename T, typename Conv>
std::vector<T> vectorValue(std::string line, Conv conv) {
  line = line.substr(line.find('=') + 1);
  std::vector<T> out;
  if (line.size() < 2) return out;
  line.erase(0, 1);
  line.replace(line.size() - 1, 1, ",");
  size_t comma;
  while ((comma = line.find(',')) < line.size()) {
    out.push_back(conv(line.substr(0, comma).c_str()));
    line.erase(0, comma + 1);
  }
  return out;
}

std::vector<Double> vectorDoubleValue(const std::string& line) {
  return vectorValue<Double>(line, [](const char* s) { return std::atof(s); });
}

std::vector<unsigned int> vectorIntValue(const std::string& line) {
  return vectorValue<unsigned int>(line, [](const char* s) { return (unsigned int)std::atoi(s); });
}

SignalType signalTypeOf(const std::string& type) {
  if (type == "neumann") return NEUMANN;
  if (type == "gaussian") return GAUSSIAN;
  if (type == "secant-hyperbolic") return SECANT;
  if (type == "flat-top") return FLATTOP;
  if (type == "inverse-gaussian") return INVGAUSSIAN;
  throw JobError(type + " is an unknown signal type.");
}

SeedType seedTypeOf(const std::string& type) {
  static const char* names[] = {"plane-wave", "truncated-plane-wave", "gaussian-beam", "super-gaussian-beam",
                                "standing-plane-wave", "standing-truncated-plane-wave", "standing-gaussian-beam",
                                "standing-super-gaussian-beam"};
  for (int t = 0; t < 8; ++t)
    if (type == names[t]) return (SeedType)t;
  throw JobError(type + " is an unknown type.");
}

// Direction/polarisation normalisation and checks shared by Seed and Undulator::initialize
// (src/classes.cpp:680-720, :1040-1075).
void normalise_beam(Vec3& direction, Vec3& polarization) {
  if (direction.norm2() == 0.0) throw JobError("The direction vector has length zero.");
  Double vl = std::sqrt(direction.norm2());
  for (int c = 0; c < 3; ++c) direction[c] /= vl;
  if (polarization.norm2() == 0.0) throw JobError("The polarization vector has length zero.");
  vl = polarization.norm();
  for (int c = 0; c < 3; ++c) polarization[c] /= vl;
  Double dot = polarization[0] * direction[0] + polarization[1] * direction[1] + polarization[2] * direction[2];
  if (std::fabs(dot) > 1.0e-50) throw JobError("Polarization is not normal to the direction.");
}

// ExtField::initialize (src/classes.cpp:1188-1215) divides direction and polarization by the
// square roots of their norms, not by the norms: a unit vector is kept as given, any other is
// scaled to the square root of its length.
void normalise_external(Vec3& direction, Vec3& polarization) {
  if (direction.norm() == 0.0) throw JobError("The direction vector has length zero.");
  Double vl = std::sqrt(direction.norm());
  for (int c = 0; c < 3; ++c) direction[c] /= vl;
  if (polarization.norm() == 0.0) throw JobError("The polarization vector has length zero.");
  vl = std::sqrt(polarization.norm());
  for (int c = 0; c < 3; ++c) polarization[c] /= vl;
  Double dot = polarization[0] * direction[0] + polarization[1] * direction[1] + polarization[2] * direction[2];
  if (std::fabs(dot) > 1.0e-50) throw JobError("Polarization is not normal to the direction.");
}

bool gaussian_type(SeedType t) {
  return t == GAUSSIANBEAM || t == STANDINGGAUSSIANBEAM || t == SUPERGAUSSIANBEAM || t == STANDINGSUPERGAUSSIANBEAM;
}

void super_gaussian_setup(std::vector<int>& order, const std::vector<int>& given, std::vector<Double>& radius,
                          Double& a0) {
  order = given;
  Double d1 = 0.0, d2 = 0.0;
  for (int i = -order[0]; i <= order[0]; i++) d1 += std::exp(-i * i);
  for (int i = -order[1]; i <= order[1]; i++) d2 += std::exp(-i * i);
  // the reference divides radius_[0] twice and never radius_[1] (src/classes.cpp:733-734)
  radius[0] /= order[0] + std::sqrt(1.0 - std::log(d1));
  radius[0] /= order[0] + std::sqrt(1.0 - std::log(d2));
  a0 /= d1 * d2;
}

}  // namespace

/* ---------------------------------------------------------------- signal and seed ------- */

void Signal::initialize(const std::string& type, Double l0, Double s, Double l, Double cep, unsigned int nR,
                        const std::vector<Double>& sigmaInvG) {
  signalType_ = signalTypeOf(type);
  t0_ = l0;
  s_ = s;
  f0_ = 1 / l;
  cep_ = cep * PI / 180;
  nR_ = nR;
  sigmaInvG_ = sigmaInvG;
  if (s_ == 0.0) throw JobError(" Variance of signal is set to zero. ");
  if (signalType_ == INVGAUSSIAN && sigmaInvG_.size() >= 2 && sigmaInvG_[0] * sigmaInvG_[1] == 0.0)
    throw JobError(" sigma of the inverse-gaussian signal is set to zero. ");
}

Double Signal::self(Double t, Double phase) const {
  if (std::fabs(t - t0_) > 10.0 * s_) return 0.0;
  switch (signalType_) {
    case NEUMANN:
      return -std::cos(2 * PI * f0_ * (t - t0_) + cep_ + phase) * 2.7724 * (t - t0_) / (s_ * s_) *
             std::exp(-1.3863 * (t - t0_) * (t - t0_) / (s_ * s_));
    case GAUSSIAN:
      return std::cos(2 * PI * f0_ * (t - t0_) + cep_ + phase) * std::exp(-1.3863 * std::pow((t - t0_) / s_, 2));
    case SECANT:
      return std::cos(2 * PI * f0_ * (t - t0_) + cep_ + phase) / std::cosh((t - t0_) / s_);
    case FLATTOP:
      if (t - t0_ <= -s_ / 2.0)
        return std::cos(2 * PI * f0_ * (t - t0_) + cep_ + phase) * std::exp(-std::pow((t - t0_ + s_ / 2.0) * f0_ / nR_, 2));
      if (t - t0_ <= s_ / 2.0) return std::cos(2 * PI * f0_ * (t - t0_) + cep_ + phase);
      return std::cos(2 * PI * f0_ * (t - t0_) + cep_ + phase) * std::exp(-std::pow((t - t0_ - s_ / 2.0) * f0_ / nR_, 2));
    case INVGAUSSIAN: {
      const Double env = std::pow((1.0 + std::pow((t - t0_) / sigmaInvG_[0], 2)) * (1.0 + std::pow((t - t0_) / sigmaInvG_[1], 2)), 0.25);
      const Double c = std::cos(2 * PI * f0_ * (t - t0_) + cep_ + phase);
      if (t - t0_ <= -s_ / 2.0) return c * env * std::exp(-std::pow((t - t0_ + s_ / 2.0) * f0_ / nR_, 2));
      if (t - t0_ <= s_ / 2.0) return c * env;
      return c * env * std::exp(-std::pow((t - t0_ - s_ / 2.0) * f0_ / nR_, 2));
    }
  }
  return 0.0;
}

void Seed::initialize(const std::string& type, const std::vector<Double>& position, const std::vector<Double>& direction,
                      const std::vector<Double>& polarization, Double a0, const std::vector<Double>& radius,
                      const std::vector<int>& order, const Signal& signal) {
  seedType_ = seedTypeOf(type);
  position_ = vec3(position);
  direction_ = vec3(direction);
  polarization_ = vec3(polarization);
  normalise_beam(direction_, polarization_);
  a0_ = a0;
  radius_ = radius;
  if ((seedType_ == GAUSSIANBEAM || seedType_ == STANDINGGAUSSIANBEAM || seedType_ == SUPERGAUSSIANBEAM ||
       seedType_ == STANDINGSUPERGAUSSIANBEAM) && radius_[0] * radius_[1] == 0.0)
    throw JobError("One of the radii of the Gaussian beam is set to zero.");
  signal_ = signal;
  if (seedType_ == SUPERGAUSSIANBEAM || seedType_ == STANDINGSUPERGAUSSIANBEAM)
    super_gaussian_setup(order_, order, radius_, a0_);
}

void Seed::fields(const Vec3& r, Double time, Vec3& a) const {
  Vec3 rl, rv, yv;
  rl[0] = r[0];
  rl[1] = r[1];
  rl[2] = gamma_ * (r[2] + beta_ * c0_ * (time + dt_));
  Double tl = gamma_ * (time + dt_ + beta_ / c0_ * r[2]);
  for (int c = 0; c < 3; ++c) rv[c] = rl[c] - position_[c];
  const Double z = rv[0] * direction_[0] + rv[1] * direction_[1] + rv[2] * direction_[2];
  tl -= z / c0_;
  Double p = 0.0;
  auto zero = [&]() { a[0] = a[1] = a[2] = 0.0; };
  auto set = [&](Double y) { for (int c = 0; c < 3; ++c) a[c] = y * polarization_[c]; };
  auto transverse = [&](Double& x, Double& y) {
    x = rv[0] * polarization_[0] + rv[1] * polarization_[1] + rv[2] * polarization_[2];
    yv[0] = direction_[1] * polarization_[2] - direction_[2] * polarization_[1];
    yv[1] = direction_[2] * polarization_[0] - direction_[0] * polarization_[2];
    yv[2] = direction_[0] * polarization_[1] - direction_[1] * polarization_[0];
    y = rv[0] * yv[0] + rv[1] * yv[1] + rv[2] * yv[2];
  };
  if (seedType_ == PLANEWAVE) {
    Double ts = signal_.self(tl, p);
    if (std::fabs(ts) < 1.0e-6) zero();
    else set(amplitude_ * ts);
  } else if (seedType_ == PLANEWAVETRUNCATED) {
    Double ts = signal_.self(tl, p), x, y;
    transverse(x, y);
    if (std::fabs(ts) < 1.0e-6 || std::fabs(x) > radius_[0] || std::fabs(y) > radius_[1]) zero();
    else set(amplitude_ * ts);
  } else if (seedType_ == GAUSSIANBEAM || seedType_ == SUPERGAUSSIANBEAM) {
    Double ts = signal_.self(tl, p), x, y;
    if (std::fabs(ts) < 1.0e-6) {
      zero();
    } else {
      transverse(x, y);
      const Double l = c0_ / signal_.f0_;
      const Double zRp = PI * radius_[0] * radius_[0] / l;
      const Double wrp = std::sqrt(1.0 + z * z / (zRp * zRp));
      const Double zRs = PI * radius_[1] * radius_[1] / l;
      const Double wrs = std::sqrt(1.0 + z * z / (zRs * zRs));
      const int ni = seedType_ == GAUSSIANBEAM ? 0 : order_[0], nj = seedType_ == GAUSSIANBEAM ? 0 : order_[1];
      for (int i = -ni; i <= ni; i++)
        for (int j = -nj; j <= nj; j++) {
          p = 0.5 * (std::atan(z / zRp) + std::atan(z / zRs) - PI) -
              PI * z / l * (std::pow(x / (zRp * wrp), 2) + std::pow(y / (zRs * wrs), 2));
          ts = signal_.self(tl, p);
          const Double t = std::exp(-std::pow(x / (radius_[0] * wrp), 2) - std::pow(y / (radius_[1] * wrs), 2)) /
                           std::sqrt(wrs * wrp) * amplitude_;
          if (seedType_ == GAUSSIANBEAM) set(t * ts);
          else for (int c = 0; c < 3; ++c) a[c] += t * ts * polarization_[c];
        }
    }
  }
  a[2] *= gamma_;
}

/* ---------------------------------------------------------------- grammar --------------- */

ParseDarius::It ParseDarius::next(It& it) {
  ++it;
  if (it == jobFile_.end()) throw JobError("The job file ends inside a group.");
  return it;
}

void ParseDarius::setJobParameters() {
  It it = jobFile_.begin();
  if (it == jobFile_.end()) return;
  do {
    if (*it == "MESH") readMesh(it);
    else if (*it == "BUNCH") readBunch(it);
    else if (*it == "FIELD") readField(it);
    else if (*it == "UNDULATOR") readUndulator(it);
    else if (*it == "EXTERNAL-FIELD") readExtField(it);
    else if (*it == "FEL-OUTPUT") readFEL(it);
    else throw JobError(*it + " is not a defined group.");
    ++it;
  } while (it != jobFile_.end());
}

void ParseDarius::readMesh(It& it) {
  if (*next(it) != "{") throw JobError("The solver directory is empty");
  next(it);
  do {
    const std::string name = parameterName(*it);
    if (name == "length-scale") {
      const std::string v = stringValue(*it);
      static const std::pair<const char*, Double> units[] = {{"METER", 1.0},        {"DECIMETER", 1.0e-1},
                                                             {"CENTIMETER", 1.0e-2}, {"MILLIMETER", 1.0e-3},
                                                             {"MICROMETER", 1.0e-6}, {"NANOMETER", 1.0e-9},
                                                             {"ANGSTROM", 1.0e-10}};
      mesh_.lengthScale_ = doubleValue(*it);
      for (auto& u : units)
        if (v == u.first) mesh_.lengthScale_ = u.second;
    } else if (name == "mesh-lengths") {
      mesh_.meshLength_ = vec3(vectorDoubleValue(*it));
    } else if (name == "mesh-resolution") {
      mesh_.meshResolution_ = vec3(vectorDoubleValue(*it));
    } else if (name == "mesh-center") {
      mesh_.meshCenter_ = vec3(vectorDoubleValue(*it));
    } else if (name == "total-time") {
      mesh_.totalTime_ = doubleValue(*it);
    } else if (name == "total-distance") {
      mesh_.totalDist_ = doubleValue(*it);
    } else if (name == "bunch-time-step") {
      bunch_.timeStep_ = doubleValue(*it);
    } else if (name == "mesh-truncation-order") {
      mesh_.truncationOrder_ = intValue(*it);
      if (mesh_.truncationOrder_ != 1 && mesh_.truncationOrder_ != 2)
        throw JobError("Mesh truncation order can not be different from one or two.");
    } else if (name == "time-scale") {
      const std::string v = stringValue(*it);
      static const std::pair<const char*, Double> units[] = {{"SECOND", 1.0},          {"MILLISECOND", 1.0e-3},
                                                             {"MICROSECOND", 1.0e-6},  {"NANOSECOND", 1.0e-9},
                                                             {"PICOSECOND", 1.0e-12},  {"FEMTOSECOND", 1.0e-15},
                                                             {"ATTOSECOND", 1.0e-18}};
      mesh_.timeScale_ = doubleValue(*it);
      for (auto& u : units)
        if (v == u.first) mesh_.timeScale_ = u.second;
    } else if (name == "space-charge") {
      mesh_.spaceCharge_ = boolValue(*it);
    } else if (name == "optimize-bunch-position") {
      mesh_.optimizePosition_ = boolValue(*it);
    } else if (name == "initial-time-back-shift") {
      mesh_.timeShift_ = doubleValue(*it);
      if (mesh_.timeShift_ < 0.0) throw JobError("The shift back in time should always be positive.");
    } else if (name == "solver") {
      const std::string v = stringValue(*it);
      if (v == "FD") mesh_.solver_ = FD;
      else if (v == "NSFD") mesh_.solver_ = NSFD;
      else throw JobError("The solver type is not among the accepted solvers.");
    } else if (name == "lorentz-factor") {
      mesh_.gamma_ = doubleValue(*it);
    } else {
      throw JobError(name + " is not defined in solver group.");
    }
    next(it);
  } while (*it != "}");
}

void ParseDarius::readBunch(It& it) {
  if (*next(it) != "{") throw JobError("The bunch directory is empty");
  next(it);
  do {
    if (*it == "bunch-initialization") {
      if (*next(it) != "{") throw JobError("The bunch-initialization directory is empty");
      next(it);
      BunchInitialize b;
      do {
        const std::string name = parameterName(*it);
        if (name == "type") b.bunchType_ = stringValue(*it);
        else if (name == "distribution") b.distribution_ = stringValue(*it);
        else if (name == "generator") {
          b.generator_ = stringValue(*it);
          if (b.generator_ != "halton" && b.generator_ != "random") throw JobError("The inserted generator is not accepted !!!");
        } else if (name == "charge") b.cloudCharge_ = doubleValue(*it);
        else if (name == "number-of-particles") b.numberOfParticles_ = intValue(*it);
        else if (name == "gamma") b.initialGamma_ = doubleValue(*it);
        else if (name == "direction") b.initialDirection_ = vec3(vectorDoubleValue(*it));
        else if (name == "position") b.position_.push_back(vec3(vectorDoubleValue(*it)));
        else if (name == "numbers") {
          std::vector<unsigned int> n = vectorIntValue(*it);
          if (n.size() != 3) throw JobError("A field-vector with other than three parameters is not accepted.");
          for (int c = 0; c < 3; ++c) b.numbers_[c] = n[c];
        } else if (name == "lattice-constants") b.latticeConstants_ = vec3(vectorDoubleValue(*it));
        else if (name == "sigma-position") b.sigmaPosition_ = vec3(vectorDoubleValue(*it));
        else if (name == "sigma-momentum") b.sigmaGammaBeta_ = vec3(vectorDoubleValue(*it));
        else if (name == "transverse-truncation") b.tranTrun_ = doubleValue(*it);
        else if (name == "longitudinal-truncation") b.longTrun_ = doubleValue(*it);
        else if (name == "file-name") b.fileName_ = stringValue(*it);
        else if (name == "bunching-factor") b.bF_ = doubleValue(*it);
        else if (name == "bunching-factor-phase") b.bFP_ = doubleValue(*it);
        else if (name == "shot-noise") b.shotNoise_ = boolValue(*it);
        else throw JobError(name + " is not defined in bunch-initialization group.");
        next(it);
      } while (*it != "}");
      bunch_.bunchInit_.push_back(b);
    } else if (*it == "bunch-sampling" || *it == "bunch-visualization" || *it == "bunch-profile") {
      const std::string group = *it;
      if (*next(it) != "{") throw JobError("The " + group + " directory is empty");
      next(it);
      do {
        const std::string name = parameterName(*it);
        if (group == "bunch-sampling") {
          if (name == "sample") bunch_.sampling_ = boolValue(*it);
          else if (name == "directory") bunch_.directory_ = stringValue(*it);
          else if (name == "base-name") bunch_.basename_ = stringValue(*it);
          else if (name == "rhythm") bunch_.rhythm_ = doubleValue(*it);
          else throw JobError(name + " is not defined in the bunch-sampling group.");
        } else if (group == "bunch-visualization") {
          if (name == "sample") bunch_.bunchVTK_ = boolValue(*it);
          else if (name == "directory") bunch_.bunchVTKDirectory_ = stringValue(*it);
          else if (name == "base-name") bunch_.bunchVTKBasename_ = stringValue(*it);
          else if (name == "rhythm") bunch_.bunchVTKRhythm_ = doubleValue(*it);
          else throw JobError(name + " is not defined in the bunch-visualization group.");
        } else {
          if (name == "sample") bunch_.bunchProfile_ = boolValue(*it);
          else if (name == "directory") bunch_.bunchProfileDirectory_ = stringValue(*it);
          else if (name == "base-name") bunch_.bunchProfileBasename_ = stringValue(*it);
          else if (name == "time") bunch_.bunchProfileTime_.push_back(doubleValue(*it));
          else if (name == "rhythm") bunch_.bunchProfileRhythm_ = doubleValue(*it);
          else throw JobError(name + " is not defined in the bunch-profile group.");
        }
        next(it);
      } while (*it != "}");
    } else {
      throw JobError(parameterName(*it) + " is not defined in the bunch-profile group.");
    }
    next(it);
  } while (*it != "}");
}

namespace {

// Keys shared by field-initialization, optical-undulator and electromagnetic-wave.
struct BeamKeys {
  std::string type, signalType;
  std::vector<Double> position = std::vector<Double>(3, 0.0), direction = std::vector<Double>(3, 0.0),
                      polarization = std::vector<Double>(3, 0.0), sigmaInvG = std::vector<Double>(2, 0.0);
  Double a0 = 0.0, offset = 0.0, pulseLength = 0.0, wavelength = 0.0, cep = 0.0;
  unsigned int nR = 2;
  std::vector<Double> radius = std::vector<Double>(2, 0.0);
  std::vector<int> order = std::vector<int>(2, 0);
  bool take(const std::string& name, const std::string& line, const char* typeKey) {
    if (name == typeKey) type = stringValue(line);
    else if (name == "position") position = vectorDoubleValue(line);
    else if (name == "direction") direction = vectorDoubleValue(line);
    else if (name == "polarization") polarization = vectorDoubleValue(line);
    else if (name == "strength-parameter") a0 = doubleValue(line);
    else if (name == "radius-parallel") radius[0] = doubleValue(line);
    else if (name == "radius-perpendicular") radius[1] = doubleValue(line);
    else if (name == "order-parallel") order[0] = (int)doubleValue(line);
    else if (name == "order-perpendicular") order[1] = (int)doubleValue(line);
    else if (name == "signal-type") signalType = stringValue(line);
    else if (name == "offset") offset = doubleValue(line);
    else if (name == "pulse-length") pulseLength = doubleValue(line);
    else if (name == "wavelength") wavelength = doubleValue(line);
    else if (name == "rising-cycles") nR = intValue(line);
    else if (name == "CEP") cep = doubleValue(line);
    else if (name == "sigma-inverse-gaussian") sigmaInvG = vectorDoubleValue(line);
    else return false;
    return true;
  }
  Signal signal() const {
    Signal s;
    s.initialize(signalType, offset, pulseLength, wavelength, cep, nR, sigmaInvG);
    return s;
  }
};

}  // namespace

namespace {

SamplingType samplingTypeOf(const std::string& t) {  // Seed::samplingType, src/classes.cpp:871-876
  if (t == "at-point") return ATPOINT;
  if (t == "over-line") return OVERLINE;
  throw JobError(t + " is an unknown sampling type.");
}

FieldType fieldTypeOf(const std::string& t) {  // Seed::fieldType, src/classes.cpp:896-908
  static const char* names[] = {"Ex", "Ey", "Ez", "Bx", "By", "Bz", "Ax", "Ay", "Az", "F"};
  for (int i = 0; i < 10; ++i)
    if (t == names[i]) return FieldType(i);
  throw JobError(t + " is an unknown sampling field.");
}

}  // namespace

void ParseDarius::readField(It& it) {
  if (*next(it) != "{") throw JobError("The seed directory is empty");
  next(it);
  do {
    if (*it == "field-initialization") {
      if (*next(it) != "{") throw JobError("The seed-initialization directory is empty");
      next(it);
      BeamKeys k;
      do {
        const std::string name = parameterName(*it);
        if (!k.take(name, *it, "type")) throw JobError(name + " is not defined in seed-initialization group.");
        next(it);
      } while (*it != "}");
      seed_.initialize(k.type, k.position, k.direction, k.polarization, k.a0, k.radius, k.order, k.signal());
    } else if (*it == "field-sampling") {  // src/datainput.cpp:330-367
      if (*next(it) != "{") throw JobError("The seed-sampling directory is empty");
      next(it);
      do {
        const std::string name = parameterName(*it);
        if (name == "sample") seed_.sampling_ = boolValue(*it);
        else if (name == "type") seed_.samplingType_ = samplingTypeOf(stringValue(*it));
        else if (name == "field") seed_.samplingField_.push_back(fieldTypeOf(stringValue(*it)));
        else if (name == "directory") seed_.samplingDirectory_ = stringValue(*it);
        else if (name == "base-name") seed_.samplingBasename_ = stringValue(*it);
        else if (name == "rhythm") seed_.samplingRhythm_ = doubleValue(*it);
        else if (name == "position") seed_.samplingPosition_.push_back(vec3(vectorDoubleValue(*it)));
        else if (name == "line-begin") seed_.samplingLineBegin_ = vec3(vectorDoubleValue(*it));
        else if (name == "line-end") seed_.samplingLineEnd_ = vec3(vectorDoubleValue(*it));
        else if (name == "number-of-points") seed_.samplingRes_ = intValue(*it);
        else throw JobError(name + " is not defined in seed-sampling group.");
        next(it);
      } while (*it != "}");
    } else if (*it == "field-visualization") {  // src/datainput.cpp:370-397
      if (*next(it) != "{") throw JobError("The seed-visualization directory is empty");
      next(it);
      seed_.vtk_.emplace_back();
      Seed::Vtk& v = seed_.vtk_.back();
      do {
        const std::string name = parameterName(*it);
        if (name == "sample") v.sample_ = boolValue(*it);
        else if (name == "directory") v.directory_ = stringValue(*it);
        else if (name == "type") {
          const std::string t = stringValue(*it);
          if (t == "in-plane") v.type_ = INPLANE;
          else if (t == "all-domain") v.type_ = ALLDOMAIN;
          else throw JobError(t + " is an unknown vtk type.");
        } else if (name == "plane") {
          const std::string t = stringValue(*it);
          if (t == "yz") v.plane_ = XNORMAL;
          else if (t == "xz") v.plane_ = YNORMAL;
          else if (t == "xy") v.plane_ = ZNORMAL;
          else throw JobError(t + " is an unknown vtk plane type.");
        } else if (name == "base-name") v.basename_ = stringValue(*it);
        else if (name == "field") v.field_.push_back(fieldTypeOf(stringValue(*it)));
        else if (name == "rhythm") v.rhythm_ = doubleValue(*it);
        else if (name == "position") v.position_ = vec3(vectorDoubleValue(*it));
        else throw JobError(name + " is not defined in seed-visualization group.");
        next(it);
      } while (*it != "}");
    } else if (*it == "field-profile") {  // src/datainput.cpp:400-418
      if (*next(it) != "{") throw JobError("The seed-profile directory is empty");
      next(it);
      do {
        const std::string name = parameterName(*it);
        if (name == "sample") seed_.profile_ = boolValue(*it);
        else if (name == "directory") seed_.profileDirectory_ = stringValue(*it);
        else if (name == "base-name") seed_.profileBasename_ = stringValue(*it);
        else if (name == "time") seed_.profileTime_.push_back(doubleValue(*it));
        else if (name == "field") seed_.profileField_.push_back(fieldTypeOf(stringValue(*it)));
        else if (name == "rhythm") seed_.profileRhythm_ = doubleValue(*it);
        else throw JobError(name + " is not defined in seed-profile group.");
        next(it);
      } while (*it != "}");
    }
    next(it);
  } while (*it != "}");
}

void ParseDarius::readUndulator(It& it) {
  if (*next(it) != "{") throw JobError("The undulator directory is empty");
  next(it);
  do {
    if (*it == "static-undulator") {
      if (*next(it) != "{") throw JobError("The static undulator directory is empty");
      next(it);
      Undulator u;
      u.type_ = STATIC;
      do {
        const std::string name = parameterName(*it);
        if (name == "undulator-parameter") u.k_ = doubleValue(*it);
        else if (name == "period") u.lu_ = doubleValue(*it);
        else if (name == "polarization-angle") u.theta_ = PI / 180.0 * doubleValue(*it);
        else if (name == "length") u.length_ = intValue(*it);
        else if (name == "distance-to-bunch-head") u.dist_ = doubleValue(*it);
        else if (name == "offset") u.rb_ = doubleValue(*it);
        else throw JobError(name + " is not defined in the static-undulator group.");
        next(it);
      } while (*it != "}");
      undulator_.push_back(u);
    } else if (*it == "static-undulator-array") {
      if (*next(it) != "{") throw JobError("The static undulator directory is empty");
      next(it);
      Undulator u;
      Double k = u.k_, lu = u.lu_, theta = u.theta_, g = 0.0, t = 0.0, d = 0.0;
      unsigned int l = u.length_, N = 1;
      do {
        const std::string name = parameterName(*it);
        if (name == "undulator-parameter") k = doubleValue(*it);
        else if (name == "period") lu = doubleValue(*it);
        else if (name == "polarization-angle") theta = PI / 180.0 * doubleValue(*it);
        else if (name == "length") l = intValue(*it);
        else if (name == "gap") g = doubleValue(*it);
        else if (name == "number") N = intValue(*it);
        else if (name == "tapering-parameter") t = doubleValue(*it);
        else if (name == "distance-to-bunch-head") d = doubleValue(*it);
        else throw JobError(name + " is not defined in the static-undulator-array group.");
        next(it);
      } while (*it != "}");
      for (unsigned int i = 0; i < N; i++) {  // src/datainput.cpp:478-492
        u.type_ = STATIC;
        u.k_ = k + i * t;
        u.lu_ = lu;
        u.theta_ = theta;
        u.length_ = l;
        u.rb_ = i * (l * lu + g);
        u.dist_ = d;
        undulator_.push_back(u);
      }
    } else if (*it == "optical-undulator") {
      if (*next(it) != "{") throw JobError("The optical undulator directory is empty");
      next(it);
      Undulator u;
      u.type_ = OPTICAL;
      BeamKeys k;
      do {
        const std::string name = parameterName(*it);
        if (name == "distance-to-bunch-head") u.dist_ = doubleValue(*it);
        else if (!k.take(name, *it, "beam-type")) throw JobError(name + " is not defined in the optical-undulator group.");
        next(it);
      } while (*it != "}");
      // the signal, then Undulator::initialize (src/datainput.cpp:545-550, src/classes.cpp:1007-1102)
      const Signal signal = k.signal();
      u.seedType_ = seedTypeOf(k.type);
      u.position_ = vec3(k.position);
      u.direction_ = vec3(k.direction);
      u.polarization_ = vec3(k.polarization);
      normalise_beam(u.direction_, u.polarization_);
      u.a0_ = k.a0;
      u.radius_ = k.radius;
      u.lu_ = k.wavelength;
      if (gaussian_type(u.seedType_) && u.radius_[0] * u.radius_[1] == 0.0)
        throw JobError("One of the radii of the Gaussian beam is set to zero.");
      u.signal_ = signal;
      if (u.seedType_ == SUPERGAUSSIANBEAM || u.seedType_ == STANDINGSUPERGAUSSIANBEAM)
        super_gaussian_setup(u.order_, k.order, u.radius_, u.a0_);
      undulator_.push_back(u);
    }
    next(it);
  } while (*it != "}");
}

void ParseDarius::readExtField(It& it) {
  if (*next(it) != "{") throw JobError("The EXTERNAL-FIELD directory is empty");
  next(it);
  do {
    if (*it == "electromagnetic-wave") {
      if (*next(it) != "{") throw JobError("The electromagnetic-field directory is empty");
      next(it);
      BeamKeys k;
      k.wavelength = 1.0;  // 1 / Signal::f0_ of a default signal (src/datainput.cpp:584-588)
      do {
        const std::string name = parameterName(*it);
        if (!k.take(name, *it, "beam-type"))
          throw JobError(name + " is not defined in the electromagnetic external field group.");
        next(it);
      } while (*it != "}");
      // the signal, then ExtField::initialize (src/datainput.cpp:615-620, src/classes.cpp:1162-1253)
      const Signal signal = k.signal();
      ExtField e;
      e.seedType_ = seedTypeOf(k.type);
      e.position_ = vec3(k.position);
      e.direction_ = vec3(k.direction);
      e.polarization_ = vec3(k.polarization);
      normalise_external(e.direction_, e.polarization_);
      e.a0_ = k.a0;
      e.radius_ = k.radius;
      if (gaussian_type(e.seedType_) && e.radius_[0] * e.radius_[1] == 0.0)
        throw JobError("One of the radii of the Gaussian beam is set to zero.");
      e.signal_ = signal;
      if (e.seedType_ == SUPERGAUSSIANBEAM || e.seedType_ == STANDINGSUPERGAUSSIANBEAM)
        super_gaussian_setup(e.order_, k.order, e.radius_, e.a0_);
      extField_.push_back(e);
    }
    next(it);
  } while (*it != "}");
}

void ParseDarius::readFEL(It& it) {
  if (*next(it) != "{") throw JobError("The FEL-OUTPUT directory is empty");
  next(it);
  do {
    if (*it == "radiation-power" || *it == "radiation-energy") {
      // radiation-energy keys land in radiationPower_ as well (src/datainput.cpp:706-717)
      const bool energy = *it == "radiation-energy";
      if (*next(it) != "{") throw JobError(std::string("The ") + (energy ? "radiation-energy" : "radiation-power") +
                                           " directory is empty");
      next(it);
      FreeElectronLaser F;
      auto& r = F.radiationPower_;
      do {
        const std::string name = parameterName(*it);
        if (name == "sample") r.sampling_ = boolValue(*it);
        else if (name == (energy ? "distance-from-bunch" : "plane-position")) r.z_.push_back(doubleValue(*it));
        else if (name == "directory") r.directory_ = stringValue(*it);
        else if (name == "base-name") r.basename_ = stringValue(*it);
        else if (name == "line-begin") r.lineBegin_ = doubleValue(*it);
        else if (name == "line-end") r.lineEnd_ = doubleValue(*it);
        else if (name == (energy ? "resolution" : "number-of-points")) r.res_ = energy ? (unsigned)doubleValue(*it) : intValue(*it);
        else if (name == (energy ? "normalized-wavelength" : "normalized-frequency")) r.lambda_.push_back(doubleValue(*it));
        else if (name == (energy ? "minimum-normalized-wavelength" : "minimum-normalized-frequency")) r.lambdaMin_ = doubleValue(*it);
        else if (name == (energy ? "maximum-normalized-wavelength" : "maximum-normalized-frequency")) r.lambdaMax_ = doubleValue(*it);
        else if (name == (energy ? "normalized-wavelength-resolution" : "number-of-frequency-points"))
          r.lambdaRes_ = energy ? (unsigned)doubleValue(*it) : intValue(*it);
        else if (name == "type") {
          const std::string v = stringValue(*it);
          if (v == "at-point") r.samplingType_ = ATPOINT;
          else if (v == "over-line") r.samplingType_ = OVERLINE;
          else throw JobError(v + " is an unknown sampling type.");
        } else throw JobError(name + " is not defined in radiation-power group.");
        next(it);
      } while (*it != "}");
      FEL_.push_back(F);
    } else if (*it == "power-visualization") {  // src/datainput.cpp:674-692
      if (*next(it) != "{") throw JobError("The power-visualization directory is empty");
      next(it);
      FreeElectronLaser F;
      auto& v = F.vtkPower_;
      do {
        const std::string name = parameterName(*it);
        if (name == "sample") v.sampling_ = boolValue(*it);
        else if (name == "directory") v.directory_ = stringValue(*it);
        else if (name == "base-name") v.basename_ = stringValue(*it);
        else if (name == "plane-position") v.z_ = doubleValue(*it);
        else if (name == "rhythm") v.rhythm_ = doubleValue(*it);
        else if (name == "normalized-frequency") v.lambda_ = doubleValue(*it);
        else throw JobError(name + " is not defined in power-visualization group.");
        next(it);
      } while (*it != "}");
      FEL_.push_back(F);
    } else if (*it == "bunch-profile-lab-frame") {  // src/datainput.cpp:726-745
      if (*next(it) != "{") throw JobError("The bunch-profile-lab-frame directory is empty");
      next(it);
      FreeElectronLaser F;
      auto& sp = F.screenProfile_;
      do {
        const std::string name = parameterName(*it);
        if (name == "sample") sp.sampling_ = boolValue(*it);
        else if (name == "directory") sp.directory_ = stringValue(*it);
        else if (name == "base-name") sp.basename_ = stringValue(*it);
        else if (name == "position") sp.pos_.push_back(doubleValue(*it));
        else if (name == "rhythm") sp.rhythm_ = doubleValue(*it);
        else throw JobError(name + " is not defined in the bunch-profile-lab-frame group.");
        next(it);
      } while (*it != "}");
      FEL_.push_back(F);
    }
    next(it);
  } while (*it != "}");
}

/* ---------------------------------------------------------------- bunch generators ------ */

void Bunch::initializeManual(BunchInitialize b, std::list<Charge>& out, int ia) {
  Charge c;  // src/classes.cpp:85-98 (one domain: always inside)
  c.q = b.cloudCharge_;
  c.rnp = b.position_[ia];
  for (int k = 0; k < 3; ++k) c.gb[k] = b.initialGamma_ * b.betaVector_[k];
  out.push_back(c);
}

void Bunch::initializeEllipsoid(BunchInitialize b, std::list<Charge>& out, int ia) {
  // src/classes.cpp:104-298 on one rank (rank 0, size 1)
  if (b.numberOfParticles_ % 4 != 0) {
    unsigned int n = b.numberOfParticles_ % 4;
    b.numberOfParticles_ += 4 - n;
    std::cout << "Warning: The number of particles in the bunch is not a multiple of four. It is corrected to "
              << b.numberOfParticles_ << std::endl;
  }
  const unsigned int Np = b.numberOfParticles_, Np0 = (unsigned int)out.size();
  unsigned int i;
  Charge charge;
  charge.q = b.cloudCharge_ / Np;
  Vec3 gb;
  for (int k = 0; k < 3; ++k) gb[k] = b.initialGamma_ * b.betaVector_[k];
  Vec3 r, t;
  Double t0, zmin = 1e100, Ne, bF = 0.0, bFi;
  unsigned int bmi;
  const unsigned int ng = (b.lambda_ == 0.0) ? 1 : 4;
  if (b.bF_ > 2.0 || b.bF_ < 0.0) throw JobError("The bunching factor can not be larger than one or a negative value !!!");
  std::vector<Double> randomNumbers;
  if (b.generator_ == "random") {  // not reproducible by construction, like the reference
    std::srand((unsigned)std::time(nullptr));
    randomNumbers.resize(Np / ng * 20, 0.0);
    for (unsigned int ri = 0; ri < Np / ng * 20; ri++) randomNumbers[ri] = ((double)std::rand()) / RAND_MAX;
  }
  auto generate = [&](unsigned int n, unsigned int m) -> Double {
    if (b.generator_ == "random") return randomNumbers[n * 2 * Np / ng + m];
    return halton(n, m);
  };
  auto insertCharge = [&](Charge q) {
    for (unsigned int ii = 0; ii < ng; ii++) {
      if (b.shotNoise_) {
        bmi = int((charge.rnp[2] - zmin) / b.lambda_);
        bFi = bF * std::sqrt(-2.0 * std::log(generate(8, bmi)));
        q.rnp[2] = charge.rnp[2] - b.lambda_ / 4 * ii;
        q.rnp[2] -= b.lambda_ / PI * bFi * std::sin(2.0 * PI / b.lambda_ * q.rnp[2] + 2.0 * PI * generate(9, bmi));
      } else if (b.lambda_ != 0.0) {
        q.rnp[2] = charge.rnp[2] - b.lambda_ / 4 * ii;
        q.rnp[2] -= b.lambda_ / PI * b.bF_ * std::sin(2.0 * PI / b.lambda_ * q.rnp[2] + b.bFP_ * PI / 180.0);
      }
      out.push_back(q);
    }
  };
  const unsigned int nTail = unsigned(Np / ng * (1.0 + 2.0 * b.lambda_ * std::sqrt(2.0 * PI) / (2.0 * b.sigmaPosition_[2])));
  if (b.shotNoise_) {
    for (i = 0; i < Np / ng; i++) {
      if (b.distribution_ == "uniform")
        zmin = std::min((2.0 * generate(2, i + Np0) - 1.0) * b.sigmaPosition_[2], zmin);
      else if (b.distribution_ == "gaussian")
        zmin = std::min(b.sigmaPosition_[2] * std::sqrt(-2.0 * std::log(generate(2, i + Np0))) *
                            std::sin(2.0 * PI * generate(3, i + Np0)), zmin);
      else
        throw JobError("The longitudinal type is not correctly given to the code !!!");
    }
    if (b.distribution_ == "uniform")
      for (; i < nTail; i++) {
        t0 = 2.0 * b.lambda_ * std::sqrt(-2.0 * std::log(generate(2, i + Np0))) * std::sin(2.0 * PI * generate(3, i + Np0));
        t0 += (t0 < 0.0) ? (-b.sigmaPosition_[2]) : (b.sigmaPosition_[2]);
        zmin = std::min(t0, zmin);
      }
    zmin = zmin + b.position_[ia][2];
    Ne = b.cloudCharge_ * b.lambda_ / (2.0 * b.sigmaPosition_[2]);
    bF = (b.bF_ == 0.0) ? 1.0 / std::sqrt(Ne) : b.bF_;
  }
  auto draw = [&](bool tail) {
    r[0] = b.sigmaPosition_[0] * std::sqrt(-2.0 * std::log(generate(0, i + Np0))) * std::cos(2.0 * PI * generate(1, i + Np0));
    r[1] = b.sigmaPosition_[1] * std::sqrt(-2.0 * std::log(generate(0, i + Np0))) * std::sin(2.0 * PI * generate(1, i + Np0));
    if (tail) {
      r[2] = 2.0 * b.lambda_ * std::sqrt(-2.0 * std::log(generate(2, i + Np0))) * std::sin(2.0 * PI * generate(3, i + Np0));
      r[2] += (r[2] < 0.0) ? (-b.sigmaPosition_[2]) : (b.sigmaPosition_[2]);
    } else if (b.distribution_ == "uniform") {
      r[2] = (2.0 * generate(2, i + Np0) - 1.0) * b.sigmaPosition_[2];
    } else if (b.distribution_ == "gaussian") {
      r[2] = b.sigmaPosition_[2] * std::sqrt(-2.0 * std::log(generate(2, i + Np0))) * std::sin(2.0 * PI * generate(3, i + Np0));
    } else {
      throw JobError("The longitudinal type is not correctly given to the code !!!");
    }
    t[0] = b.sigmaGammaBeta_[0] * std::sqrt(-2.0 * std::log(generate(4, i + Np0))) * std::cos(2.0 * PI * generate(5, i + Np0));
    t[1] = b.sigmaGammaBeta_[1] * std::sqrt(-2.0 * std::log(generate(4, i + Np0))) * std::sin(2.0 * PI * generate(5, i + Np0));
    t[2] = b.sigmaGammaBeta_[2] * std::sqrt(-2.0 * std::log(generate(6, i + Np0))) * std::cos(2.0 * PI * generate(7, i + Np0));
    if (std::fabs(r[0]) < b.tranTrun_ && std::fabs(r[1]) < b.tranTrun_ && std::fabs(r[2]) < b.longTrun_) {
      charge.rnp = b.position_[ia];
      for (int k = 0; k < 3; ++k) charge.rnp[k] += r[k];
      charge.gb = gb;
      for (int k = 0; k < 3; ++k) charge.gb[k] += t[k];
      insertCharge(charge);
    }
  };
  for (i = 0; i < Np / ng; i += 1) draw(false);
  if (b.distribution_ == "uniform")
    for (; i < nTail; i += 1) draw(true);
}

void Bunch::initialize3DCrystal(BunchInitialize b, std::list<Charge>& out, int ia) {
  const unsigned int cells = b.numbers_[0] * b.numbers_[1] * b.numbers_[2];
  if (cells == 0 || b.numberOfParticles_ % cells != 0)
    throw JobError("The number of the particles and their lattice numbers do not match !!!");
  Vec3 gb;
  for (int k = 0; k < 3; ++k) gb[k] = b.initialGamma_ * b.betaVector_[k];
  const unsigned int np = b.numberOfParticles_ / cells;
  out.clear();  // src/classes.cpp:318
  Charge c;
  for (unsigned int i = 0; i < b.numbers_[0]; i++)
    for (unsigned int j = 0; j < b.numbers_[1]; j++)
      for (unsigned int k = 0; k < b.numbers_[2]; k++)
        for (unsigned int l = 0; l < np; l++) {
          c.q = b.cloudCharge_ / b.numberOfParticles_;
          c.rnp[0] = b.position_[ia][0] + (i + 1.0 - 0.5 * b.numbers_[0]) * b.latticeConstants_[0];
          c.rnp[1] = b.position_[ia][1] + (j + 1.0 - 0.5 * b.numbers_[1]) * b.latticeConstants_[1];
          c.rnp[2] = b.position_[ia][2] + (k + 1.0 - 0.5 * b.numbers_[2]) * b.latticeConstants_[2];
          c.rnp[0] += 0.5 * b.sigmaPosition_[0] * std::sqrt(-2.0 * std::log(halton(0, i))) * std::sin(2.0 * PI * halton(1, i));
          c.rnp[1] += 0.5 * b.sigmaPosition_[1] * std::sqrt(-2.0 * std::log(halton(2, i))) * std::sin(2.0 * PI * halton(3, i));
          c.rnp[2] += 0.5 * b.sigmaPosition_[2] * std::sqrt(-2.0 * std::log(halton(4, i))) * std::sin(2.0 * PI * halton(5, i));
          c.gb = gb;
          c.gb[0] += b.sigmaGammaBeta_[0] * std::sqrt(-2.0 * std::log(halton(6, i))) * std::sin(2.0 * PI * halton(7, i));
          c.gb[1] += b.sigmaGammaBeta_[1] * std::sqrt(-2.0 * std::log(halton(8, i))) * std::sin(2.0 * PI * halton(9, i));
          c.gb[2] += b.sigmaGammaBeta_[2] * std::sqrt(-2.0 * std::log(halton(10, i))) * std::sin(2.0 * PI * halton(11, i));
          out.push_back(c);
        }
}

void Bunch::initializeFile(BunchInitialize b, std::list<Charge>& out, int ia) {
  // src/classes.cpp:363-417: six columns x y z gbx gby gbz per particle; the loop tests
  // ifstream::good(), so trailing whitespace reads one extra copy of the last particle and
  // then fails the count check, as in the reference.
  out.clear();
  std::ifstream f(b.fileName_.c_str());
  Charge c;
  bool flag = false;
  c.q = b.cloudCharge_ / b.numberOfParticles_;
  while (f.good()) {
    f >> c.rnp[0] >> c.rnp[1] >> c.rnp[2] >> c.gb[0] >> c.gb[1] >> c.gb[2];
    for (int k = 0; k < 3; ++k) c.rnp[k] += b.position_[ia][k];
    out.push_back(c);
    if (b.tranTrun_ > 0.0 && (std::fabs(c.rnp[0]) > b.tranTrun_ || std::fabs(c.rnp[1]) > b.tranTrun_)) flag = true;
  }
  if (flag)
    std::cout << "Warning: Some particle coordinates are out of the transverse truncation length for the bunch. "
                 "The results may be inaccurate !!!" << std::endl;
  if (b.numberOfParticles_ != out.size())
    throw JobError("The number of the particles and the file size do not match !!! The file contains " +
                   std::to_string(out.size()) + " particles.");
}

}  // namespace mithra_host
