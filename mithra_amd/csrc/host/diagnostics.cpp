// Diagnostics writers of the drop-in Solver: the reference's field sampling, field and bunch
// visualization (VTK XML), field and bunch profiles, bunch moments and lab-frame screens, with
// its file names, layouts and number formats.  The values come from the device: the few nodes a
// sampler interpolates between (mithra_dev_sample_nodes), the bunch's charge-weighted sums
// (mithra_dev_bunch_sums), the screen crossings (mithra_dev_screen_crossings), and whole arrays
// only for the writers that print whole arrays.
//
// Set-up: Solver::initializeSeedSampling / initializeSeedVTK / initializeSeedProfile
// (src/solver.cpp:848-1047), the file part of Solver::initializeBunchUpdate (:1061-1126) and
// Solver::initializeScreenProfile (:2145-2199).  Writers: FdTd::fieldSample (src/fdtd.cpp:851-950,
// FdTdSC src/fdtdSC.cpp:1147-1262), FdTd::fieldVisualizeAllDomain / InPlane{X,Y,Z}Normal
// (src/fdtd.cpp:956-1545), FdTd::fieldProfile (:1546-1595), Solver::bunchSample /
// bunchVisualize / bunchProfile (src/solver.cpp:1582-1792), Solver::screenProfile (:2205-2257),
// Solver::initializePowerVisualize / powerVisualize (src/radiation.cpp:238-450).
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <iostream>
#include <numeric>

#include "solver.h"

namespace mithra_host {

namespace {

bool isabsolute(const std::string& f) { return f.compare(0, 1, "/") == 0; }  // src/stdinclude.h:65-68

void splitFilename(const std::string& s, std::string& path, std::string& file) {  // src/stdinclude.cpp:18-23
  size_t found = s.find_last_of('/');
  path = s.substr(0, found + 1);
  file = s.substr(found + 1);
}

void makeDirectory(const std::string& filename) {  // createDirectory, src/stdinclude.cpp:26-36
  std::string path, file;
  splitFilename(filename, path, file);
  struct stat st;
  if (path.empty() || stat(path.c_str(), &st) == 0) return;
  if (mkdir(path.c_str(), S_IRWXU | S_IRWXG | S_IROTH | S_IXOTH) == -1 && errno != EEXIST)
    throw JobError("Could not create the directory " + path + ". Probably the given address does not exist.");
}

std::unique_ptr<std::ofstream> openFile(const std::string& name) {
  std::unique_ptr<std::ofstream> f(new std::ofstream(name.c_str(), std::ios::trunc));
  if (!f->good()) throw JobError("Could not open the output file " + name + ".");
  return f;
}

// One rank: the rank suffix of every per-rank file name is 0 and the slab is the whole mesh.
const char* const kRank = "0";

void vtsHeader(std::ostream& f, const std::string& extent) {
  f << "<?xml version=\"1.0\"?>" << std::endl;
  f << "<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
       "compressor=\"vtkZLibDataCompressor\">"
    << std::endl;
  f << extent;
  f << "<Points>" << std::endl;
  f << "<DataArray type = \"Float64\" NumberOfComponents=\"3\" format=\"ascii\">" << std::endl;
}

void vtsMiddle(std::ostream& f, size_t components) {
  f << "</DataArray>" << std::endl;
  f << "</Points>" << std::endl;
  f << "<CellData>" << std::endl;
  f << "</CellData>" << std::endl;
  f << "<PointData Vectors = \"field\">" << std::endl;
  f << "<DataArray type=\"Float64\" Name=\"field\" NumberOfComponents=\"" << components << "\" format=\"ascii\">"
    << std::endl;
}

void vtsFooter(std::ostream& f) {
  f << "</DataArray>" << std::endl;
  f << "</PointData>" << std::endl;
  f << "</Piece>" << std::endl;
  f << "</StructuredGrid>" << std::endl;
  f << "</VTKFile>" << std::endl;
}

void pvts(const std::string& name, const std::string& whole, const std::string& piece, size_t components) {
  auto f = openFile(name);
  *f << "<?xml version=\"1.0\"?>" << std::endl;
  *f << "<VTKFile type=\"PStructuredGrid\" version=\"0.1\" >" << std::endl;
  *f << whole;
  *f << "<PPoints>" << std::endl;
  *f << "<DataArray type = \"Float64\" NumberOfComponents=\"3\" format=\"ascii\" />" << std::endl;
  *f << "</PPoints>" << std::endl;
  *f << "<PPointData>" << std::endl;
  *f << "<DataArray type=\"Float64\" NumberOfComponents=\"" << components << "\" Name=\"field\" format=\"ascii\" />"
     << std::endl;
  *f << "</PPointData>" << std::endl;
  *f << piece;
  *f << "</PStructuredGrid>" << std::endl;
  *f << "</VTKFile>" << std::endl;
}

void writeRow(std::ostream& f, const std::vector<Double>& v, size_t L, size_t n) {
  f << v[L * n];
  for (size_t l = 1; l < L; l++) f << " " << v[L * n + l];
  f << std::endl;
}

}  // namespace

/* ------------------------------------------------------------------------------ set-up ---- */

void Solver::initializeSeedSampling() {
  if (seed_.samplingRhythm_ == 0)
    throw JobError("The sampling rhythm of the field is zero although sampling is activated !!!");
  seed_.samplingRhythm_ /= gamma_;
  for (auto& p : seed_.samplingPosition_) p[2] *= gamma_;
  seed_.samplingLineBegin_[2] *= gamma_;
  seed_.samplingLineEnd_[2] *= gamma_;
  if (seed_.samplingType_ == OVERLINE) {
    Vec3 l = seed_.samplingLineEnd_;
    for (int c = 0; c < 3; ++c) l[c] -= seed_.samplingLineBegin_[c];
    const Double n = seed_.samplingRes_;
    for (int c = 0; c < 3; ++c) l[c] /= n;
    for (unsigned i = 0; i < n; i++) {
      Vec3 position;
      for (int c = 0; c < 3; ++c) position[c] = seed_.samplingLineBegin_[c] + i * l[c];
      seed_.samplingPosition_.push_back(position);
    }
  }
  // Points outside the mesh (one cell in from every face) are dropped.  The reference compares
  // with UpdateBunch::dx..dz before Solver::initializeBunchUpdate assigns them (src/solver.cpp:898-900
  // against :1055-1057); the margin here is the mesh step they are assigned.
  const Double dx = mesh_.meshResolution_[0], dy = mesh_.meshResolution_[1], dz = mesh_.meshResolution_[2];
  std::vector<Vec3> kept;
  for (auto& p : seed_.samplingPosition_) {
    if (p[0] < xmax_ - dx && p[0] > xmin_ + dx && p[1] < ymax_ - dy && p[1] > ymin_ + dy && p[2] < zmax_ - dz &&
        p[2] > zmin_ + dz) {
      if (p[2] < zp_[1] && p[2] >= zp_[0]) kept.push_back(p);
    } else {
      std::cout << "The sampling point does not reside in the grid. No data is saved." << std::endl;
    }
  }
  seed_.samplingPosition_ = kept;
  if (!kept.empty()) {
    std::string base = isabsolute(seed_.samplingBasename_) ? "" : seed_.samplingDirectory_;
    base += seed_.samplingBasename_ + "-" + kRank + ".txt";
    makeDirectory(base);
    sfFile_ = openFile(base);
  }
  sfCe_ = mesh_.lengthScale_ / std::pow(mesh_.timeScale_, 2);
  sfCb_ = 1.0 / (mesh_.lengthScale_ * mesh_.timeScale_);
  sfCa_ = 1.0 / mesh_.timeScale_;
  sfCf_ = std::pow(mesh_.lengthScale_ / mesh_.timeScale_, 2);
}

void Solver::initializeSeedVTK() {
  vfName_.assign(seed_.vtk_.size(), std::string());
  for (unsigned i = 0; i < seed_.vtk_.size();) {
    Seed::Vtk& v = seed_.vtk_[i];
    if (!v.sample_) {
      ++i;
      continue;
    }
    if (v.rhythm_ == 0)
      throw JobError("The visualization rhythm of the field is zero although visualization is activated !!!");
    v.rhythm_ /= gamma_;
    v.position_[2] *= gamma_;
    if (!isabsolute(v.basename_)) v.basename_ = v.directory_ + v.basename_;
    std::string path;
    splitFilename(v.basename_, path, vfName_[i]);
    makeDirectory(v.basename_);
    bool outside = false;
    if (v.type_ == INPLANE) {
      if (v.plane_ == XNORMAL)
        outside = v.position_[0] > xmax_ - mesh_.meshResolution_[0] || v.position_[0] < xmin_ + mesh_.meshResolution_[0];
      else if (v.plane_ == YNORMAL)
        outside = v.position_[1] > ymax_ - mesh_.meshResolution_[1] || v.position_[1] < ymin_ + mesh_.meshResolution_[1];
      else
        outside = v.position_[2] > zmax_ - mesh_.meshResolution_[2] || v.position_[2] < zmin_ + mesh_.meshResolution_[2];
    }
    if (outside) {
      // The reference erases the block and then advances its loop index past the block that
      // moved into this slot, leaving that one without its set-up (src/solver.cpp:980-1009); here
      // the next block is set up like any other.
      std::cout << "The plane does not reside in the grid. No data is saved." << std::endl;
      seed_.vtk_.erase(seed_.vtk_.begin() + i);
      vfName_.erase(vfName_.begin() + i);
      continue;
    }
    ++i;
  }
}

void Solver::initializeSeedProfile() {
  if (seed_.profileRhythm_ == 0 && seed_.profileTime_.empty())
    throw JobError("The profiling rhythm of the field is zero and no time is set although profiling of the field "
                   "is activated !!!");
  seed_.profileRhythm_ /= gamma_;
  for (auto& t : seed_.profileTime_) t /= gamma_;
  if (!isabsolute(seed_.profileBasename_)) seed_.profileBasename_ = seed_.profileDirectory_ + seed_.profileBasename_;
  makeDirectory(seed_.profileBasename_);
}

void Solver::initializeBunchDiagnostics() {
  if (bunch_.sampling_) {
    std::string base = isabsolute(bunch_.basename_) ? "" : bunch_.directory_;
    base += bunch_.basename_ + ".txt";
    makeDirectory(base);
    sbFile_ = openFile(base);
    if (bunch_.rhythm_ == 0)
      throw JobError("The sampling rhythm of the bunch is zero although sampling is activated !!!");
  }
  if (bunch_.bunchVTK_) {
    if (!isabsolute(bunch_.bunchVTKBasename_)) bunch_.bunchVTKBasename_ = bunch_.bunchVTKDirectory_ + bunch_.bunchVTKBasename_;
    makeDirectory(bunch_.bunchVTKBasename_);
    if (bunch_.bunchVTKRhythm_ == 0)
      throw JobError("The visualization rhythm of the bunch is zero although visualization is activated !!!");
  }
  if (bunch_.bunchProfile_) {
    if (!isabsolute(bunch_.bunchProfileBasename_))
      bunch_.bunchProfileBasename_ = bunch_.bunchProfileDirectory_ + bunch_.bunchProfileBasename_;
    makeDirectory(bunch_.bunchProfileBasename_);
    if (bunch_.bunchProfileRhythm_ == 0 && bunch_.bunchProfileTime_.empty())
      throw JobError("The profiling rhythm of the bunch is zero and no time is set although profiling of the bunch "
                     "is activated !!!");
  }
}

void Solver::initializeScreenProfile() {
  scrFiles_.clear();
  scrFiles_.resize(FEL_.size());
  for (unsigned jf = 0; jf < FEL_.size(); jf++) {
    auto& S = FEL_[jf].screenProfile_;
    if (!S.sampling_) continue;
    if (!isabsolute(S.basename_)) S.basename_ = S.directory_ + S.basename_;
    if (S.rhythm_ > 0.0) {
      // One screen every `rhythm` up to the end of the undulator.  The reference reads that end
      // from undulator_.end(), one past the last element (src/solver.cpp:2160); the last
      // undulator is used here.
      if (undulator_.empty()) throw JobError("Screens at a rhythm need an undulator to end at.");
      const Undulator& u = undulator_.back();
      Double z = 0.0;
      while (z < (u.rb_ + u.length_ * u.lu_)) {
        S.pos_.push_back(z);
        z += S.rhythm_;
      }
    }
    makeDirectory(S.basename_);
    std::sort(S.pos_.begin(), S.pos_.end());
    for (unsigned i = 0; i < S.pos_.size(); i++) {
      std::cout << "Screen " << i << " is at distance " << S.pos_[i] << " from the undulator beginning." << std::endl;
      scrFiles_[jf].push_back(openFile(S.basename_ + "-p" + kRank + "-screen" + std::to_string(i) + ".txt"));
      scrFiles_[jf].back()->setf(std::ios::scientific);
      scrFiles_[jf].back()->precision(15);
      scrFiles_[jf].back()->width(40);
    }
    if (S.pos_.empty())
      throw JobError("No position is set for the screen although the screen sampling is activated !!!");
  }
}

void Solver::initializePowerVisualize() {
  vizFEL_ = -1;
  for (unsigned jf = 0; jf < FEL_.size(); jf++) {
    auto& V = FEL_[jf].vtkPower_;
    if (!V.sampling_) continue;
    if (V.rhythm_ == 0)
      throw JobError("The power visualization rhythm of the field is zero although power visualization is "
                     "activated !!!");
    V.rhythm_ /= gamma_;
    V.z_ *= gamma_;
    if (!isabsolute(V.basename_)) V.basename_ = V.directory_ + V.basename_;
    makeDirectory(V.basename_);
    if (!(V.z_ < zp_[1] && V.z_ >= zp_[0])) continue;  // Nz = 0: the plane is not in the mesh
    if (vizFEL_ >= 0) throw JobError("More than one power-visualization block is not supported by this build.");
    if (undulator_.empty()) throw JobError("Power visualization needs an undulator for its wavelength.");
    vizFEL_ = (int)jf;
    const Double dt = undulator_[0].lu_ / V.lambda_ / (gamma_ * c0_);
    vizNf_ = unsigned(3.0 * dt / mesh_.timeStep_);
    vizW_ = 2 * PI / dt;
    vizPc_ = 2.0 * mesh_.meshResolution_[0] * mesh_.meshResolution_[1] / (m0_ * vizNf_ * vizNf_) *
             std::pow(mesh_.lengthScale_, 2) / std::pow(mesh_.timeScale_, 3);
    if (vizNf_ == 0) throw JobError("The power visualization needs at least one time sample per period.");
  }
}

void Solver::powerVisualizeSetup() {
  if (vizFEL_ < 0) return;
  const auto& V = FEL_[vizFEL_].vtkPower_;
  Double c;
  mithra_power_map_params p;
  p.dzr = std::modf((V.z_ - zmin_) / mesh_.meshResolution_[2], &c);
  p.k = (int)c - k0_;
  p.Nf = vizNf_;
  std::vector<double> cs(vizNf_), sn(vizNf_);
  for (unsigned j = 0; j < vizNf_; j++) {  // src/radiation.cpp:311-318
    cs[j] = std::cos(vizW_ * j * mesh_.timeStep_);
    sn[j] = std::sin(vizW_ * j * mesh_.timeStep_);
  }
  p.cos_wt = cs.data();
  p.sin_wt = sn.data();
  p.pc = vizPc_;
  p.gamma = gamma_;
  p.beta = beta_;
  p.c0 = c0_;
  check(mithra_dev_power_map_setup(dev_, &p), "power visualization set-up");
}

void Solver::powerVisualize() {
  if (vizFEL_ < 0) return;
  const auto& V = FEL_[vizFEL_].vtkPower_;
  check(mithra_dev_power_map_sample(dev_, nTime_), "power visualization sample");
  if (!(std::fmod(time_, V.rhythm_) < mesh_.timeStep_)) return;
  std::vector<double> pl((size_t)N1N0_);
  check(mithra_dev_power_map_fetch(dev_, pl.data()), "power visualization");
  Double c;
  const Double dzr = std::modf((V.z_ - zmin_) / mesh_.meshResolution_[2], &c);
  const int k = (int)c - k0_;
  auto f = openFile(V.basename_ + "-" + std::to_string(nTime_) + ".vts");
  f->setf(std::ios::scientific);
  f->precision(4);
  const std::string box = "0 " + std::to_string(N0_ - 1) + " 0 " + std::to_string(N1_ - 1) + " 0 0";
  vtsHeader(*f, "<StructuredGrid WholeExtent=\"" + box + "\">\n<Piece Extent=\"" + box + "\">\n");
  for (int j = 0; j < N1_; j++)
    for (int i = 0; i < N0_; i++) {
      const unsigned int m = k * N1_ * N0_ + i * N1_ + j;
      const Vec3 r1 = rc(m), r2 = rc(m + N1N0_);
      *f << r1[0] * (1.0 - dzr) + r2[0] * dzr << " " << r1[1] << " " << r1[2] << std::endl;
    }
  *f << "</DataArray>" << std::endl;
  *f << "</Points>" << std::endl;
  *f << "<CellData>" << std::endl;
  *f << "</CellData>" << std::endl;
  *f << "<PointData Vectors = \"power\">" << std::endl;
  *f << "<DataArray type=\"Float64\" Name=\"power\" NumberOfComponents=\"" << 1 << "\" format=\"ascii\">" << std::endl;
  for (int j = 0; j < N1_; j++)
    for (int i = 0; i < N0_; i++) *f << pl[(size_t)i * N1_ + j] << std::endl;
  vtsFooter(*f);
  f->close();
}

/* ----------------------------------------------------------------------- rhythm tests ---- */

// Whether a field writer fires at the current time_ (the tests of Solver::solve,
// src/solver.cpp:1329-1352); decided before the step's field update so that the update
// evaluates E/B on every node that step.
bool Solver::fieldWritersDue() const {
  const Double dt = mesh_.timeStep_;
  if (time_ <= 0.0) {
    // the profile's time list is the only field test that does not require time_ > 0
    if (!seed_.profile_) return false;
  }
  if (seed_.sampling_ && std::fmod(time_, seed_.samplingRhythm_) < dt && time_ > 0.0 && !seed_.samplingPosition_.empty())
    return true;
  for (auto& v : seed_.vtk_)
    if (v.sample_ && std::fmod(time_, v.rhythm_) < dt && time_ > 0.0) return true;
  if (seed_.profile_) {
    for (Double t : seed_.profileTime_)
      if (time_ - t < dt && time_ > t) return true;
    if (std::fmod(time_, seed_.profileRhythm_) < dt && time_ > 0.0 && seed_.profileRhythm_ != 0) return true;
  }
  return false;
}

void Solver::fieldWriters() {
  const Double dt = mesh_.timeStep_;
  fullValid_ = false;
  if (seed_.sampling_ && std::fmod(time_, seed_.samplingRhythm_) < dt && time_ > 0.0) fieldSample();
  for (unsigned i = 0; i < seed_.vtk_.size(); i++)
    if (seed_.vtk_[i].sample_ && std::fmod(time_, seed_.vtk_[i].rhythm_) < dt && time_ > 0.0) {
      if (seed_.vtk_[i].type_ == ALLDOMAIN) fieldVisualizeAllDomain(i);
      else if (seed_.vtk_[i].type_ == INPLANE) fieldVisualizeInPlane(i);
    }
  if (seed_.profile_) {
    for (Double t : seed_.profileTime_)
      if (time_ - t < dt && time_ > t) fieldProfile();
    if (std::fmod(time_, seed_.profileRhythm_) < dt && time_ > 0.0 && seed_.profileRhythm_ != 0) fieldProfile();
  }
}

void Solver::bunchWriters() {
  const Double dt = mesh_.timeStep_, t = time_ + mesh_.timeShift_;
  if (bunch_.sampling_ && std::fmod(t, bunch_.rhythm_) < dt && t > 0.0) bunchSample();
  if (bunch_.bunchVTK_ && std::fmod(t, bunch_.bunchVTKRhythm_) < dt && t > 0.0) bunchVisualize();
  if (bunch_.bunchProfile_) {
    for (Double bt : bunch_.bunchProfileTime_)
      if (time_ - bt < dt && time_ > bt) bunchProfile();
    if (std::fmod(t, bunch_.bunchProfileRhythm_) < dt && t > 0.0 && bunch_.bunchProfileRhythm_ != 0.0) bunchProfile();
  }
}

bool Solver::diagnosticsActive() const {
  if (seed_.sampling_ || seed_.profile_ || bunch_.sampling_ || bunch_.bunchVTK_ || bunch_.bunchProfile_) return true;
  for (auto& v : seed_.vtk_)
    if (v.sample_) return true;
  for (auto& f : FEL_)
    if (f.screenProfile_.sampling_) return true;
  return vizFEL_ >= 0;
}

/* ---------------------------------------------------------------------- field writers ---- */

// The whole E/B and A^n (and phi^n) arrays, downloaded once per step for the writers that
// print every node.
void Solver::fetchFull() {
  if (fullValid_) return;
  const size_t n = (size_t)N1N0_ * np_;
  fullE_.resize(3 * n);
  fullB_.resize(3 * n);
  fullA_.resize(3 * n);
  check(mithra_dev_get_fields(dev_, fullE_.data(), fullB_.data()), "field download");
  if (mesh_.spaceCharge_) fullF_.resize(n);
  check(mithra_dev_get_potentials(dev_, 0, fullA_.data(), mesh_.spaceCharge_ ? fullF_.data() : nullptr),
        "potential download");
  fullValid_ = true;
}

Vec3 Solver::rc(long m) const {  // Solver::rc, src/solver.cpp:2302-2315
  unsigned int i, j, k;
  k = m / N1N0_;
  i = (m % N1N0_) / N1_;
  j = m - (N1N0_ * k + N1_ * i);
  Vec3 v;
  v[0] = xmin_ + i * mesh_.meshResolution_[0];
  v[1] = ymin_ + j * mesh_.meshResolution_[1];
  v[2] = zmin_ + (k + k0_) * mesh_.meshResolution_[2];
  return v;
}

// The value of field `f` at node m from the full arrays (0 for F without space charge).
Double Solver::nodeValue(FieldType f, long m) const {
  switch (f) {
    case Ex: return fullE_[3 * m];
    case Ey: return fullE_[3 * m + 1];
    case Ez: return fullE_[3 * m + 2];
    case Bx: return fullB_[3 * m];
    case By: return fullB_[3 * m + 1];
    case Bz: return fullB_[3 * m + 2];
    case Ax: return fullA_[3 * m];
    case Ay: return fullA_[3 * m + 1];
    case Az: return fullA_[3 * m + 2];
    case F: return fullF_.empty() ? 0.0 : fullF_[m];
  }
  return 0.0;
}

void Solver::fieldSample() {
  const size_t N = seed_.samplingPosition_.size();
  if (N == 0) return;
  std::ofstream& f = *sfFile_;
  f.setf(std::ios::scientific);
  f.precision(4);
  f << time_ * gamma_ << "\t";
  // the eight corners of every point, in the reference's order
  std::vector<long> nodes(8 * N);
  std::vector<Double> w(3 * N);
  for (size_t n = 0; n < N; ++n) {
    const Vec3& p = seed_.samplingPosition_[n];
    Double c1;
    const Double dxr = std::modf((p[0] - xmin_) / mesh_.meshResolution_[0], &c1);
    const int i = (int)c1;
    const Double dyr = std::modf((p[1] - ymin_) / mesh_.meshResolution_[1], &c1);
    const int j = (int)c1;
    const Double dzr = std::modf((p[2] - zmin_) / mesh_.meshResolution_[2], &c1);
    const int k = (int)c1;
    const long m = (k - k0_) * N1N0_ + i * N1_ + j;
    const long corner[8] = {m, m + N1_, m + 1, m + N1_ + 1, m + N1N0_, m + N1N0_ + N1_, m + N1N0_ + 1,
                            m + N1N0_ + N1_ + 1};
    std::copy(corner, corner + 8, nodes.begin() + 8 * n);
    w[3 * n] = dxr;
    w[3 * n + 1] = dyr;
    w[3 * n + 2] = dzr;
  }
  std::vector<float> e(3 * 8 * N), b(3 * 8 * N);
  std::vector<double> a(3 * 8 * N), phi(8 * N);
  check(mithra_dev_sample_nodes(dev_, nodes.data(), (long)nodes.size(), e.data(), b.data(), a.data(), phi.data()),
        "field sampling");
  const Double g1 = std::sqrt(std::pow(gamma_, 2) - 1);
  for (size_t n = 0; n < N; ++n) {
    const Double dxr = w[3 * n], dyr = w[3 * n + 1], dzr = w[3 * n + 2];
    const Double wt[8] = {(1.0 - dxr) * (1.0 - dyr) * (1.0 - dzr), dxr * (1.0 - dyr) * (1.0 - dzr),
                          (1.0 - dxr) * dyr * (1.0 - dzr),         dxr * dyr * (1.0 - dzr),
                          (1.0 - dxr) * (1.0 - dyr) * dzr,         dxr * (1.0 - dyr) * dzr,
                          (1.0 - dxr) * dyr * dzr,                 dxr * dyr * dzr};
    Double et[3], bt[3], at[3], ft = 0.0;
    for (int c = 0; c < 3; ++c) {  // FieldVector::mv then seven pmv (src/fieldvector.h:62-76)
      const size_t q0 = 8 * n;
      et[c] = wt[0] * e[3 * q0 + c];
      bt[c] = wt[0] * b[3 * q0 + c];
      at[c] = wt[0] * a[3 * q0 + c];
      for (int q = 1; q < 8; ++q) {
        et[c] += wt[q] * e[3 * (q0 + q) + c];
        bt[c] += wt[q] * b[3 * (q0 + q) + c];
        at[c] += wt[q] * a[3 * (q0 + q) + c];
      }
    }
    if (mesh_.spaceCharge_) {
      ft = wt[0] * phi[8 * n];
      for (int q = 1; q < 8; ++q) ft += wt[q] * phi[8 * n + q];
    }
    const Vec3& p = seed_.samplingPosition_[n];
    f << p[0] << "\t" << p[1] << "\t" << p[2] << "\t";
    for (FieldType s : seed_.samplingField_) {
      switch (s) {
        case Ex: f << (gamma_ * et[0] + c0_ * g1 * bt[1]) * sfCe_ << "\t"; break;
        case Ey: f << (gamma_ * et[1] - c0_ * g1 * bt[0]) * sfCe_ << "\t"; break;
        case Ez: f << et[2] * sfCe_ << "\t"; break;
        case Bx: f << (gamma_ * bt[0] - g1 / c0_ * et[1]) * sfCb_ << "\t"; break;
        case By: f << (gamma_ * bt[1] + g1 / c0_ * et[0]) * sfCb_ << "\t"; break;
        case Bz: f << bt[2] * sfCb_ << "\t"; break;
        case Ax: f << at[0] * sfCa_ << "\t"; break;
        case Ay: f << at[1] * sfCa_ << "\t"; break;
        case Az: f << at[2] * sfCa_ << "\t"; break;
        case F:  // FdTdSC only (src/fdtdSC.cpp:1254-1255); FdTd writes nothing for it
          if (mesh_.spaceCharge_) f << ft * sfCf_ << "\t";
          break;
      }
    }
  }
  f << std::endl;
}


void Solver::fieldVisualizeAllDomain(unsigned ivtk) {
  const Seed::Vtk& V = seed_.vtk_[ivtk];
  const size_t L = V.field_.size();
  const std::string nt = std::to_string(nTime_);
  fetchFull();
  std::vector<Double> v((size_t)N1N0_ * np_ * L, 0.0);
  for (int k = 0; k < np_; k++)
    for (int j = 1; j < N1_ - 1; j++)
      for (int i = 1; i < N0_ - 1; i++) {
        const long m = k * N1_ * N0_ + i * N1_ + j;
        for (size_t l = 0; l < L; l++) v[L * m + l] = nodeValue(V.field_[l], m);
      }
  auto f = openFile(V.basename_ + "-p" + kRank + "-" + nt + ".vts");
  f->setf(std::ios::scientific);
  f->precision(4);
  const std::string kz = std::to_string(k0_) + " " + std::to_string(k0_ + np_ - 2 + 1);
  const std::string box = "0 " + std::to_string(N0_ - 1) + " 0 " + std::to_string(N1_ - 1) + " ";
  vtsHeader(*f, "<StructuredGrid WholeExtent=\"" + box + kz + "\">\n<Piece Extent=\"" + box + kz + "\">\n");
  for (int k = 0; k < np_; k++)
    for (int j = 0; j < N1_; j++)
      for (int i = 0; i < N0_; i++) {
        const Vec3 r = rc((long)k * N1_ * N0_ + i * N1_ + j);
        *f << r[0] << " " << r[1] << " " << r[2] << std::endl;
      }
  vtsMiddle(*f, L);
  for (int k = 0; k < np_; k++)
    for (int j = 0; j < N1_; j++)
      for (int i = 0; i < N0_; i++) writeRow(*f, v, L, (size_t)k * N1_ * N0_ + i * N1_ + j);
  vtsFooter(*f);
  f->close();
  pvts(V.basename_ + "-" + nt + ".pvts",
       "<PStructuredGrid WholeExtent=\"0 " + std::to_string(N0_ - 1) + " 0 " + std::to_string(N1_ - 1) + " 0 " +
           std::to_string(N2_ - 1) + "\" GhostLevel = \"0\" >\n",
       "<Piece Extent=\"0 " + std::to_string(N0_ - 1) + " 0 " + std::to_string(N1_ - 1) + " 0 " +
           std::to_string(N2_ - 2 + 1) + "\" Source=\"" + vfName_[ivtk] + "-p" + kRank + "-" + nt + ".vts\" />\n",
       L);
}

void Solver::fieldVisualizeInPlane(unsigned ivtk) {
  const Seed::Vtk& V = seed_.vtk_[ivtk];
  if (V.plane_ == ZNORMAL && !(V.position_[2] < zp_[1] && V.position_[2] >= zp_[0])) return;
  const size_t L = V.field_.size();
  const std::string nt = std::to_string(nTime_);
  auto f = openFile(V.basename_ + "-p" + kRank + "-" + nt + ".vts");
  f->setf(std::ios::scientific);
  f->precision(4);
  const int axis = V.plane_ == XNORMAL ? 0 : V.plane_ == YNORMAL ? 1 : 2;
  const Double lo = axis == 0 ? xmin_ : axis == 1 ? ymin_ : zmin_;
  Double c;
  const Double fr = std::modf((V.position_[axis] - lo) / mesh_.meshResolution_[axis], &c);
  const long step = axis == 0 ? N1_ : axis == 1 ? 1 : N1N0_;
  // the plane's nodes (lower side of the cell) and their value index n, as the reference loops
  std::vector<long> nodes;
  std::vector<size_t> slot;
  size_t nv = 0;
  if (axis == 0) {
    const int i = (int)c;
    nv = (size_t)N1_ * np_;
    for (int k = 0; k < np_; k++)
      for (int j = 1; j < N1_ - 1; j++) {
        nodes.push_back((long)k * N1_ * N0_ + i * N1_ + j);
        slot.push_back((size_t)k * N1_ + j);
      }
  } else if (axis == 1) {
    const int j = (int)c;
    nv = (size_t)N0_ * np_;
    for (int k = 0; k < np_; k++)
      for (int i = 1; i < N0_ - 1; i++) {
        nodes.push_back((long)k * N1_ * N0_ + i * N1_ + j);
        slot.push_back((size_t)k * N0_ + i);
      }
  } else {
    const int k = (int)c - k0_;
    nv = (size_t)N1N0_;
    for (int j = 1; j < N1_ - 1; j++)
      for (int i = 1; i < N0_ - 1; i++) {
        nodes.push_back((long)k * N1_ * N0_ + i * N1_ + j);
        slot.push_back((size_t)i * N1_ + j);
      }
  }
  // each node and its neighbour across the plane
  const size_t P = nodes.size();
  std::vector<long> both(2 * P);
  for (size_t q = 0; q < P; ++q) {
    both[2 * q] = nodes[q];
    both[2 * q + 1] = nodes[q] + step;
  }
  std::vector<float> e(6 * P), b(6 * P);
  std::vector<double> a(6 * P), phi(2 * P);
  check(mithra_dev_sample_nodes(dev_, both.data(), (long)both.size(), e.data(), b.data(), a.data(), phi.data()),
        "plane sampling");
  std::vector<Double> v(nv * L, 0.0);
  for (size_t q = 0; q < P; ++q)
    for (size_t l = 0; l < L; l++) {
      const size_t lo_q = 2 * q, hi_q = 2 * q + 1;
      Double val = 0.0;
      switch (V.field_[l]) {  // en_[m][c] (1 - r) + en_[m + step][c] r, float operands widened
        case Ex: val = e[3 * lo_q] * (1.0 - fr) + e[3 * hi_q] * fr; break;
        case Ey: val = e[3 * lo_q + 1] * (1.0 - fr) + e[3 * hi_q + 1] * fr; break;
        case Ez: val = e[3 * lo_q + 2] * (1.0 - fr) + e[3 * hi_q + 2] * fr; break;
        case Bx: val = b[3 * lo_q] * (1.0 - fr) + b[3 * hi_q] * fr; break;
        case By: val = b[3 * lo_q + 1] * (1.0 - fr) + b[3 * hi_q + 1] * fr; break;
        case Bz: val = b[3 * lo_q + 2] * (1.0 - fr) + b[3 * hi_q + 2] * fr; break;
        case Ax: val = a[3 * lo_q] * (1.0 - fr) + a[3 * hi_q] * fr; break;
        case Ay: val = a[3 * lo_q + 1] * (1.0 - fr) + a[3 * hi_q + 1] * fr; break;
        case Az: val = a[3 * lo_q + 2] * (1.0 - fr) + a[3 * hi_q + 2] * fr; break;
        case F:
          if (!mesh_.spaceCharge_) continue;  // FdTd leaves the value at zero
          val = phi[lo_q] * (1.0 - fr) + phi[hi_q] * fr;
          break;
      }
      v[L * slot[q] + l] = val;
    }
  const std::string kz = std::to_string(k0_) + " " + std::to_string(k0_ + np_ - 2 + 1);
  const std::string wz = " 0 " + std::to_string(N2_ - 1);
  const std::string wk = std::to_string(0) + " " + std::to_string(N2_ - 2 + 1);
  const std::string src = vfName_[ivtk] + "-p" + kRank + "-" + nt + ".vts";
  if (axis == 0) {
    vtsHeader(*f, "<StructuredGrid WholeExtent=\"0  0  0 " + std::to_string(N1_ - 1) + " " + kz +
                      "\">\n<Piece Extent=\" 0 0 0 " + std::to_string(N1_ - 1) + " " + kz + "\">\n");
    for (int k = 0; k < np_; k++)
      for (int j = 0; j < N1_; j++) {
        const long m = (long)k * N1_ * N0_ + (long)c * N1_ + j;
        const Vec3 r1 = rc(m), r2 = rc(m + N1_);
        *f << r1[0] * (1.0 - fr) + r2[0] * fr << " " << r1[1] << " " << r1[2] << std::endl;
      }
    vtsMiddle(*f, L);
    for (int k = 0; k < np_; k++)
      for (int j = 0; j < N1_; j++) writeRow(*f, v, L, (size_t)k * N1_ + j);
    vtsFooter(*f);
    f->close();
    pvts(V.basename_ + "-" + nt + ".pvts",
         "<PStructuredGrid WholeExtent=\" 0 0 0 " + std::to_string(N1_ - 1) + wz + "\" GhostLevel = \"0\" >\n",
         "<Piece Extent=\"0 0 0 " + std::to_string(N1_ - 1) + " " + wk + "\" Source=\"" + src + "\" />\n", L);
  } else if (axis == 1) {
    vtsHeader(*f, "<StructuredGrid WholeExtent=\"0 " + std::to_string(N0_ - 1) + " 0 0 " + kz +
                      "\">\n<Piece Extent=\"0 " + std::to_string(N0_ - 1) + " 0 0 " + kz + "\">\n");
    for (int k = 0; k < np_; k++)
      for (int i = 0; i < N0_; i++) {
        const long m = (long)k * N1_ * N0_ + i * N1_ + (long)c;
        const Vec3 r1 = rc(m), r2 = rc(m + 1);
        // x is weighted with the y fraction, as the reference writes it (src/fdtd.cpp:1349)
        *f << r1[0] * (1.0 - fr) + r2[0] * fr << " " << r1[1] << " " << r1[2] << std::endl;
      }
    vtsMiddle(*f, L);
    for (int k = 0; k < np_; k++)
      for (int i = 0; i < N0_; i++) writeRow(*f, v, L, (size_t)k * N0_ + i);
    vtsFooter(*f);
    f->close();
    pvts(V.basename_ + "-" + nt + ".pvts",
         "<PStructuredGrid WholeExtent=\"0 " + std::to_string(N0_ - 1) + " 0 0" + wz + "\" GhostLevel = \"0\" >\n",
         "<Piece Extent=\"0 " + std::to_string(N0_ - 1) + " 0 0 " + wk + "\" Source=\"" + src + "\" />\n", L);
  } else {
    const int k = (int)c - k0_;
    vtsHeader(*f, "<StructuredGrid WholeExtent=\"0 " + std::to_string(N0_ - 1) + " 0 " + std::to_string(N1_ - 1) +
                      " 0 0\">\n<Piece Extent=\"0 " + std::to_string(N0_ - 1) + " 0 " + std::to_string(N1_ - 1) +
                      " 0 0\">\n");
    for (int j = 0; j < N1_; j++)
      for (int i = 0; i < N0_; i++) {
        const long m = (long)k * N1_ * N0_ + i * N1_ + j;
        const Vec3 r1 = rc(m), r2 = rc(m + N1N0_);
        // x is weighted with the z fraction and z is the lower plane's, as the reference writes it
        *f << r1[0] * (1.0 - fr) + r2[0] * fr << " " << r1[1] << " " << r1[2] << std::endl;
      }
    vtsMiddle(*f, L);
    for (int j = 0; j < N1_; j++)
      for (int i = 0; i < N0_; i++) writeRow(*f, v, L, (size_t)i * N1_ + j);
    vtsFooter(*f);
    f->close();
  }
}

void Solver::fieldProfile() {
  fetchFull();
  auto f = openFile(seed_.profileBasename_ + "-p" + kRank + "-" + std::to_string(nTime_) + ".txt");
  f->setf(std::ios::scientific);
  f->precision(4);
  for (int i = 0; i < N0_; i++)
    for (int j = 0; j < N1_; j++)
      for (int k = 0; k < np_; k++) {
        const long m = (long)k * N1_ * N0_ + i * N1_ + j;
        const Vec3 r = rc(m);
        *f << r[0] << "\t" << r[1] << "\t" << r[2] << "\t";
        for (FieldType s : seed_.profileField_) {
          if (s == F && !mesh_.spaceCharge_) continue;  // FdTd prints nothing for F
          *f << nodeValue(s, m) << "\t";
        }
        *f << std::endl;
      }
  f->close();
}

/* ---------------------------------------------------------------------- bunch writers ---- */

void Solver::bunchSample() {
  double s[13];
  check(mithra_dev_bunch_sums(dev_, s), "bunch sums");
  const Double qT = s[0];
  Double rT[3], r2T[3], gbT[3], gb2T[3];
  for (int l = 0; l < 3; ++l) {
    rT[l] = s[1 + l] / qT;
    r2T[l] = s[4 + l] / qT;
    gbT[l] = s[7 + l] / qT;
    gb2T[l] = s[10 + l] / qT;
  }
  std::ofstream& f = *sbFile_;
  f.setf(std::ios::scientific);
  f.precision(4);
  f << timeBunch_ << "\t";
  f << rT[0] << "\t" << rT[1] << "\t" << rT[2] << "\t";
  f << gbT[0] << "\t" << gbT[1] << "\t" << gbT[2] << "\t";
  f << std::sqrt(r2T[0] - rT[0] * rT[0]) << "\t";
  f << std::sqrt(r2T[1] - rT[1] * rT[1]) << "\t";
  f << std::sqrt(r2T[2] - rT[2] * rT[2]) << "\t";
  f << std::sqrt(gb2T[0] - gbT[0] * gbT[0]) << "\t";
  f << std::sqrt(gb2T[1] - gbT[1] * gbT[1]) << "\t";
  f << std::sqrt(gb2T[2] - gbT[2] * gbT[2]) << std::endl;
}

// Every particle record in load order (the reference's list order).
const std::vector<double>& Solver::fetchParticles() {
  const long n = mithra_dev_particle_count(dev_);
  partCache_.resize((size_t)11 * n);
  check(mithra_dev_get_particles(dev_, partCache_.data(), n), "particle download");
  return partCache_;
}

void Solver::bunchVisualize() {
  const std::vector<double>& q = fetchParticles();
  const size_t N = q.size() / 11;
  const std::string nt = std::to_string(nTimeBunch_);
  auto f = openFile(bunch_.bunchVTKBasename_ + "-p" + kRank + "-" + nt + ".vtu");
  f->setf(std::ios::scientific);
  f->precision(4);
  *f << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">" << std::endl;
  *f << "<UnstructuredGrid>" << std::endl;
  *f << "<Piece NumberOfPoints=\"" << N + 1 << "\" NumberOfCells=\"" << 1 << "\">" << std::endl;
  *f << "<Points>" << std::endl;
  *f << "<DataArray type = \"Float64\" NumberOfComponents=\"3\" format=\"ascii\">" << std::endl;
  for (size_t p = 0; p < N; ++p) *f << q[11 * p + 1] << " " << q[11 * p + 2] << " " << q[11 * p + 3] << std::endl;
  *f << xmin_ << " " << ymin_ << " " << zmin_ << std::endl;
  *f << "</DataArray>" << std::endl;
  *f << "</Points>" << std::endl;
  *f << "<Cells>" << std::endl;
  *f << "<DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">" << std::endl;
  for (unsigned i = 0; i < N + 1; ++i) *f << i << " ";
  *f << std::endl;
  *f << "</DataArray>" << std::endl;
  *f << "<DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">" << std::endl;
  *f << N + 1 << std::endl;
  *f << "</DataArray>" << std::endl;
  *f << "<DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">" << std::endl;
  *f << 2 << std::endl;
  *f << "</DataArray>" << std::endl;
  *f << "</Cells>" << std::endl;
  *f << "<PointData Vectors = \"charge\">" << std::endl;
  *f << "<DataArray type=\"Float64\" Name=\"charge\" NumberOfComponents=\"3\" format=\"ascii\">" << std::endl;
  for (size_t p = 0; p < N; ++p) {
    const double* c = &q[11 * p];
    const Double gb0 = c[7], gb1 = c[8], gb2 = c[9];
    const Double gamma = std::sqrt(1.0 + (gb0 * gb0 + gb1 * gb1 + gb2 * gb2));
    const Double beta = gb2 / gamma;
    *f << c[0] << " " << gamma * gamma_ * (1.0 + beta_ * beta) << " " << gamma * gamma_ * (1.0 + beta_ * beta) * 0.512
       << std::endl;
  }
  *f << 0.0 << " " << 0.0 << " " << 0.0 << std::endl;
  *f << 0.0 << " " << 0.0 << " " << 0.0 << std::endl;
  *f << "</DataArray>" << std::endl;
  *f << "</PointData>" << std::endl;
  *f << "</Piece>" << std::endl;
  *f << "</UnstructuredGrid>" << std::endl;
  *f << "</VTKFile>" << std::endl;
  f->close();

  std::string path, name;
  splitFilename(bunch_.bunchVTKBasename_, path, name);
  auto pf = openFile(bunch_.bunchVTKBasename_ + "-" + nt + ".pvtu");
  *pf << "<VTKFile type=\"PUnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">" << std::endl;
  *pf << "<PUnstructuredGrid> GhostLevel = \"0\"" << std::endl;
  *pf << "<PPoints>" << std::endl;
  *pf << "<PDataArray type = \"Float64\" NumberOfComponents=\"3\" format=\"ascii\" />" << std::endl;
  *pf << "</PPoints>" << std::endl;
  *pf << "<PPointData>" << std::endl;
  *pf << "<PDataArray type=\"Float64\" Name=\"charge\" NumberOfComponents=\"3\" format=\"ascii\" />" << std::endl;
  *pf << "</PPointData>" << std::endl;
  *pf << "<Piece  Source=\"" << name + "-p" + kRank + "-" + nt + ".vtu" << "\"/>" << std::endl;
  *pf << "</PUnstructuredGrid>" << std::endl;
  *pf << "</VTKFile>" << std::endl;
  pf->close();
}

void Solver::bunchProfile() {
  const std::vector<double>& q = fetchParticles();
  const size_t N = q.size() / 11;
  auto f = openFile(bunch_.bunchProfileBasename_ + "-p" + kRank + "-" + std::to_string(nTime_) + ".txt");
  f->setf(std::ios::scientific);
  f->precision(15);
  f->width(40);
  *f << time_ * gamma_ << std::endl;
  for (size_t p = 0; p < N; ++p) {
    const double* c = &q[11 * p];
    *f << c[0] << "\t" << c[1] << "\t" << c[2] << "\t" << c[3] << "\t" << c[7] << "\t" << c[8] << "\t" << c[9]
       << std::endl;
  }
  f->close();
}

void Solver::screenProfile() {
  for (unsigned jf = 0; jf < FEL_.size(); jf++) {
    const auto& S = FEL_[jf].screenProfile_;
    if (!S.sampling_ || S.pos_.empty()) continue;
    mithra_screen_params sp;
    sp.gamma = gamma_;
    sp.beta = beta_;
    sp.c0 = c0_;
    sp.time_bunch = timeBunch_;
    sp.time_step = mesh_.timeStep_;
    sp.dt_shift = dt_;
    long cap = (long)scrRec_.size() / 6, count = 0;
    for (;;) {
      scrScreen_.resize(std::max(cap, 1L));
      scrIndex_.resize(std::max(cap, 1L));
      scrRec_.resize(6 * std::max(cap, 1L));
      check(mithra_dev_screen_crossings(dev_, S.pos_.data(), (int)S.pos_.size(), &sp, cap, &count, scrScreen_.data(),
                                        scrIndex_.data(), scrRec_.data()),
            "screen crossings");
      if (count <= cap) break;
      cap = count;  // a second pass with room for every crossing (the call changes no state)
    }
    if (count == 0) continue;
    // the reference writes screen by screen, each in list order
    std::vector<long> order(count);
    std::iota(order.begin(), order.end(), 0L);
    std::sort(order.begin(), order.end(), [&](long a, long b) {
      return scrScreen_[a] != scrScreen_[b] ? scrScreen_[a] < scrScreen_[b] : scrIndex_[a] < scrIndex_[b];
    });
    for (long o : order) {
      std::ofstream& f = *scrFiles_[jf][scrScreen_[o]];
      const double* r = &scrRec_[6 * o];
      f << r[0] << "\t" << r[1] << "\t" << r[2] << "\t" << r[3] << "\t" << r[4] << "\t" << r[5] << std::endl;
    }
  }
}

}  // namespace mithra_host
