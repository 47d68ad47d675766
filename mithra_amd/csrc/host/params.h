// Parsed job parameters and the bunch generators of the MITHRA drop-in host.
//
// The structs keep the reference's class names and member meanings (reference src/classes.h,
// src/database.h:18-80) so that the job grammar maps one to one; the set-up arithmetic follows
// the reference in source order (the library is built with -O3 -ffp-contract=off), which the
// tests check bit for bit against the reference's own dumps.
#pragma once

#include <cmath>
#include <list>
#include <stdexcept>
#include <string>
#include <vector>

namespace mithra_host {

typedef double Double;

// Constants of the reference (src/stdinclude.h:43-52); PI is truncated exactly as there.
const Double PI = 3.1415926535;
const Double EPSILON_ZERO = 8.85418782e-12;
const Double MU_ZERO = 4.0 * PI * 1.0e-7;
const Double C0 = 1.0 / std::sqrt(EPSILON_ZERO * MU_ZERO);
const Double EC = 1.602e-19;
const Double EM = 9.109e-31;

enum SignalType { NEUMANN, GAUSSIAN, SECANT, FLATTOP, INVGAUSSIAN };
enum SeedType {
  PLANEWAVE, PLANEWAVETRUNCATED, GAUSSIANBEAM, SUPERGAUSSIANBEAM,
  STANDINGPLANEWAVE, STANDINGPLANEWAVETRUNCATED, STANDINGGAUSSIANBEAM, STANDINGSUPERGAUSSIANBEAM
};
enum SamplingType { ATPOINT, OVERLINE, INPLANE, ALLDOMAIN };
enum FieldType { Ex, Ey, Ez, Bx, By, Bz, Ax, Ay, Az, F };  // src/classes.cpp:896-908
enum PlaneType { XNORMAL, YNORMAL, ZNORMAL };              // src/classes.cpp:887-893
enum UndulatorType { STATIC, OPTICAL };
enum SolverType { FD, NSFD };

// Raised for every job error the reference reports with printmessage + exit(1); the command
// line front end prints it and exits with status 1, library callers get a status code.
struct JobError : std::runtime_error {
  explicit JobError(const std::string& m) : std::runtime_error(m) {}
};

struct Vec3 {
  Double a[3] = {0.0, 0.0, 0.0};
  Vec3() = default;
  Vec3(Double x, Double y, Double z) { a[0] = x; a[1] = y; a[2] = z; }
  Double& operator[](int i) { return a[i]; }
  const Double& operator[](int i) const { return a[i]; }
  Double norm2() const { return a[0] * a[0] + a[1] * a[1] + a[2] * a[2]; }
  Double norm() const { return std::sqrt(norm2()); }
};
Vec3 vec3(const std::vector<Double>& v);  // JobError unless 3 values

// The reference's Charge record (src/stdinclude.h:130-144).
struct Charge {
  Double q = 0.0;
  Vec3 rnp, rnm, gb;
  Double e = 0.0;
};

Double halton(unsigned int i, unsigned int j);  // src/stdinclude.cpp:45-73
Double pmod(Double a, Double b);                // src/stdinclude.cpp:88-93

struct Mesh {
  Double lengthScale_ = 0.0, timeScale_ = 0.0, timeStep_ = 0.0, totalTime_ = 0.0, totalDist_ = 0.0;
  Vec3 meshCenter_, meshLength_, meshResolution_;
  unsigned int truncationOrder_ = 0;
  bool spaceCharge_ = false, optimizePosition_ = false;
  Double timeShift_ = 0.0;
  SolverType solver_ = NSFD;
  Double gamma_ = -1.0;  // Mesh::initialize, src/classes.cpp:38-46
};

struct BunchInitialize {  // src/database.h:18-80
  std::string bunchType_, distribution_, generator_, fileName_;
  unsigned int numberOfParticles_ = 0;
  Double cloudCharge_ = 0.0, initialGamma_ = 0.0, initialBeta_ = 0.0;
  Vec3 initialDirection_;
  std::vector<Vec3> position_;
  unsigned int numbers_[3] = {0, 0, 0};
  Vec3 latticeConstants_, sigmaPosition_, sigmaGammaBeta_;
  Double tranTrun_ = 0.0, longTrun_ = 0.0, lambda_ = 0.0, bF_ = 0.0, bFP_ = 0.0;
  bool shotNoise_ = false;
  Vec3 betaVector_;
};

struct Bunch {
  std::vector<BunchInitialize> bunchInit_;
  std::string directory_ = "./", basename_;
  bool sampling_ = false;
  Double timeStep_ = 0.0, rhythm_ = 0.0;
  bool bunchVTK_ = false;
  std::string bunchVTKDirectory_ = "./", bunchVTKBasename_;
  Double bunchVTKRhythm_ = 0.0;
  bool bunchProfile_ = false;
  std::string bunchProfileDirectory_ = "./", bunchProfileBasename_;
  std::vector<Double> bunchProfileTime_;
  Double bunchProfileRhythm_ = 0.0;
  Double zu_ = 0.0, beta_ = 0.0;

  // Generators of the reference (src/classes.cpp:85-417), appending in the reference's order.
  void initializeManual(BunchInitialize b, std::list<Charge>& out, int ia);
  void initializeEllipsoid(BunchInitialize b, std::list<Charge>& out, int ia);
  void initialize3DCrystal(BunchInitialize b, std::list<Charge>& out, int ia);
  void initializeFile(BunchInitialize b, std::list<Charge>& out, int ia);
};

struct Signal {  // src/classes.cpp:471-575
  SignalType signalType_ = GAUSSIAN;
  Double t0_ = 0.0, s_ = 0.0, f0_ = 1.0, cep_ = 0.0;
  unsigned int nR_ = 1;
  std::vector<Double> sigmaInvG_ = std::vector<Double>(2, 0.0);
  void initialize(const std::string& type, Double l0, Double s, Double l, Double cep, unsigned int nR,
                  const std::vector<Double>& sigmaInvG);
  Double self(Double t, Double phase) const;
};

struct Seed {
  SeedType seedType_ = PLANEWAVE;
  Double c0_ = 0.0;
  Vec3 position_, direction_, polarization_;
  Double amplitude_ = 0.0, a0_ = 0.0;
  std::vector<Double> radius_ = std::vector<Double>(2, 0.0);
  Signal signal_;
  std::vector<int> order_ = std::vector<int>(2, 0);
  Double l_ = 0.0;
  std::vector<Double> zR_ = std::vector<Double>(2, 0.0);
  Double beta_ = 0.0, gamma_ = 1.0, dt_ = 0.0;
  // field-sampling, field-visualization and field-profile groups (src/classes.cpp:623-641,
  // :858-868; read by src/datainput.cpp:330-427)
  bool sampling_ = false;
  SamplingType samplingType_ = ATPOINT;
  std::vector<FieldType> samplingField_;
  std::string samplingDirectory_, samplingBasename_;
  Double samplingRhythm_ = 0.0;
  std::vector<Vec3> samplingPosition_;
  Vec3 samplingLineBegin_, samplingLineEnd_;
  Double samplingRes_ = 0.0;
  struct Vtk {
    bool sample_ = false;
    std::vector<FieldType> field_;
    std::string directory_, basename_;
    Double rhythm_ = 0.0;
    SamplingType type_ = ALLDOMAIN;
    PlaneType plane_ = ZNORMAL;
    Vec3 position_;
  };
  std::vector<Vtk> vtk_;
  bool profile_ = false;
  std::vector<FieldType> profileField_;
  std::string profileDirectory_, profileBasename_;
  std::vector<Double> profileTime_;
  Double profileRhythm_ = 0.0;
  void initialize(const std::string& type, const std::vector<Double>& position, const std::vector<Double>& direction,
                  const std::vector<Double>& polarization, Double a0, const std::vector<Double>& radius,
                  const std::vector<int>& order, const Signal& signal);
  // Seed::fields (src/classes.cpp:740-855): `a` is accumulated into for the super-gaussian type.
  void fields(const Vec3& r, Double time, Vec3& a) const;
};

struct Undulator {
  Double k_ = 0.0, lu_ = 0.0, rb_ = 0.0;
  unsigned int length_ = 0;
  Double dist_ = 0.0, beta_ = 0.0, gamma_ = 0.0, dt_ = 0.0, theta_ = 0.0;
  UndulatorType type_ = STATIC;
  SeedType seedType_ = PLANEWAVE;
  Double c0_ = 0.0, amplitude_ = 0.0, a0_ = 0.0, l_ = 0.0;
  Vec3 position_, direction_, polarization_;
  std::vector<Double> radius_ = std::vector<Double>(2, 0.0);
  Signal signal_;
  std::vector<int> order_ = std::vector<int>(2, 0);
  std::vector<Double> zR_ = std::vector<Double>(2, 0.0);
};

struct ExtField {  // src/classes.cpp:1145-1253
  SeedType seedType_ = PLANEWAVE;
  Double c0_ = 0.0, amplitude_ = 0.0, a0_ = 0.0, l_ = 0.0;
  Vec3 position_, direction_, polarization_;
  std::vector<Double> radius_ = std::vector<Double>(2, 0.0);
  Signal signal_;
  std::vector<int> order_ = std::vector<int>(2, 0);
  std::vector<Double> zR_ = std::vector<Double>(2, 0.0);
};

struct FreeElectronLaser {
  struct RadiationSampling {
    std::vector<Double> z_;
    bool sampling_ = false;
    std::string directory_, basename_;
    Double lineBegin_ = 0.0, lineEnd_ = 0.0;
    unsigned int res_ = 0;
    SamplingType samplingType_ = ATPOINT;
    std::vector<Double> lambda_;
    Double lambdaMin_ = 0.0, lambdaMax_ = 0.0;
    unsigned int lambdaRes_ = 0;
  };
  RadiationSampling radiationPower_;
  struct RadiationVisualization {  // power-visualization (src/datainput.cpp:674-692)
    bool sampling_ = false;
    std::string directory_, basename_;
    Double z_ = 0.0, lambda_ = 0.0, rhythm_ = 0.0;
  };
  RadiationVisualization vtkPower_;
  struct ScreenProfile {  // bunch-profile-lab-frame (src/datainput.cpp:726-745, src/classes.cpp:1324-1331)
    bool sampling_ = false;
    std::string directory_ = "./", basename_;
    std::vector<Double> pos_;
    Double rhythm_ = 0.0;
  };
  ScreenProfile screenProfile_;
};

// Job-file reader and grammar (reference src/readdata.cpp, src/datainput.cpp).
std::list<std::string> read_file(const std::string& filename);
void cleanJobFile(std::list<std::string>& jobFile);

class ParseDarius {
 public:
  ParseDarius(std::list<std::string>& jobFile, Mesh& mesh, Bunch& bunch, Seed& seed, std::vector<Undulator>& undulator,
              std::vector<ExtField>& extField, std::vector<FreeElectronLaser>& FEL)
      : jobFile_(jobFile), mesh_(mesh), bunch_(bunch), seed_(seed), undulator_(undulator), extField_(extField),
        FEL_(FEL) {}
  void setJobParameters();

 private:
  typedef std::list<std::string>::iterator It;
  void readMesh(It& it);
  void readBunch(It& it);
  void readField(It& it);
  void readUndulator(It& it);
  void readExtField(It& it);
  void readFEL(It& it);
  It next(It& it);
  std::list<std::string>& jobFile_;
  Mesh& mesh_;
  Bunch& bunch_;
  Seed& seed_;
  std::vector<Undulator>& undulator_;
  std::vector<ExtField>& extField_;
  std::vector<FreeElectronLaser>& FEL_;
};

}  // namespace mithra_host
