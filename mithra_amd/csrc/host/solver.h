// The drop-in Solver: the reference's set-up and main loop, with the time march on the GPU.
//
// Keeps the reference's class interface (MITHRA::Solver with the FdTd / FdTdSC plugins chosen
// by mesh.spaceCharge_, reference src/solver.h:20-345, src/mithra.cpp:86-90) and output
// layout; the field/particle work of each step goes through the device C-ABI of
// include/mithra_gpu.h instead of the reference's CPU loops.
#pragma once

#include <fstream>
#include <memory>
#include <string>
#include <vector>

#include "mithra_gpu.h"
#include "params.h"

namespace mithra_host {

class Solver {
 public:
  Solver(Mesh& mesh, Bunch& bunch, Seed& seed, std::vector<Undulator>& undulator, std::vector<ExtField>& extField,
         std::vector<FreeElectronLaser>& FEL);
  virtual ~Solver();

  // Set-up, in the order of Solver::initialize (src/solver.cpp:547-595).
  void setSimulationParameters();  // src/solver.cpp:68-214
  void initializeBunch();          // src/solver.cpp:1130-1178
  void lorentzBoostMesh();         // src/solver.cpp:220-257
  void initializeMesh();           // src/solver.cpp:601-689
  void lorentzBoostBunch();        // src/solver.cpp:263-423
  void computeFileGamma(BunchInitialize& b);  // src/solver.cpp:509-540
  void initializeField();          // src/solver.cpp:695-842 (coefficients + seeded A^0, A^{-1})
  void initializeBunchUpdate();    // src/solver.cpp:1050-1059
  void initializePowerSample();    // src/radiation.cpp:18-121
  void shiftBackInTime();          // src/solver.cpp:1184-1206
  void initialize();

  // Diagnostics writers (diagnostics.cpp): set-up in Solver::initialize's order, writers called
  // from solve() where the reference's main loop calls them (src/solver.cpp:1248-1370).
  void initializeSeedSampling();   // src/solver.cpp:848-932
  void initializeSeedVTK();        // src/solver.cpp:938-1016
  void initializeSeedProfile();    // src/solver.cpp:1022-1044
  void initializeBunchDiagnostics();  // src/solver.cpp:1061-1126
  void initializeScreenProfile();  // src/solver.cpp:2145-2199
  void fieldSample();              // src/fdtd.cpp:851-950, src/fdtdSC.cpp:1147-1262
  void fieldVisualizeAllDomain(unsigned ivtk);  // src/fdtd.cpp:956-1106
  void fieldVisualizeInPlane(unsigned ivtk);    // src/fdtd.cpp:1112-1545
  void fieldProfile();             // src/fdtd.cpp:1546-1595, src/fdtdSC.cpp:1857-1910
  void bunchSample();              // src/solver.cpp:1582-1641
  void bunchVisualize();           // src/solver.cpp:1647-1757
  void bunchProfile();             // src/solver.cpp:1763-1792
  void screenProfile();            // src/solver.cpp:2205-2257
  void initializePowerVisualize(); // src/radiation.cpp:238-318
  void powerVisualize();           // src/radiation.cpp:324-450
  Vec3 rc(long m) const;           // src/solver.cpp:2302-2315

  // Device hand-over of the state built by initialize().
  void upload(int precision);
  // Solver::solve main loop (src/solver.cpp:1212-1414); max_steps < 0 runs to totalTime_.
  long solve(long max_steps = -1, int precision = MITHRA_PRECISION_FP64, bool quiet = false);

  // C-ABI views of the set-up
  mithra_grid grid() const;
  mithra_coeffs coeffs() const;
  mithra_seed seed() const;
  mithra_bunch_params bunchParams() const;
  std::vector<mithra_undulator> undulators() const;
  int nSubSteps() const;   // iterations of the sub-step loop (src/solver.cpp:1316)

  Mesh& mesh_;
  Bunch& bunch_;
  Seed& seed_;
  std::vector<Undulator>& undulator_;
  std::vector<ExtField>& extField_;
  std::vector<FreeElectronLaser>& FEL_;

  // state after initialize()
  std::vector<Double> an_, anm1_, fn_, fnm1_;   // AoS3 / scalar host copies of A^n, A^{n-1}, phi
  std::list<Charge> chargeVectorn_;
  int N0_ = 0, N1_ = 0, N2_ = 0, np_ = 0, k0_ = 0;
  long N1N0_ = 0;
  Double zp_[2] = {0.0, 0.0};
  Double xmin_ = 0, xmax_ = 0, ymin_ = 0, ymax_ = 0, zmin_ = 0, zmax_ = 0;
  Double timep1_ = 0, time_ = 0, timem1_ = 0;
  unsigned int nTime_ = 0;
  Double timeBunch_ = 0;
  unsigned int nTimeBunch_ = 0;
  Double nUpdateBunch_ = 0;
  Double gamma_ = 0, beta_ = 0, dt_ = 0;
  Double c0_, m0_, e0_;
  // update coefficients (UpdateField, UpdateCurrent, UpdateBunch)
  Double a_[6], bB_[5], cB_[5], dB_[5], eE_[5], fE_[5], gE_[5], hC_[17], alpha_ = 0, betaNS_ = 0;
  Double dv_ = 0, rc_ = 0, r1_ = 0, r2_ = 0, dtb_ = 0;
  // radiation power (one radiation-power block, SampleRadiationPower)
  int powerFEL_ = -1;
  unsigned pN_ = 0, pNl_ = 0, pNf_ = 0;
  Double pc_ = 0;
  std::vector<Double> pw_;
  std::vector<std::unique_ptr<std::ofstream>> pfile_;
  // external bunch for type = other (the reference leaves it to an external program)
  std::vector<Charge> external_;
  // Device generation: the bunch lives in a mithra_bunch on the GPU from its generation on, and
  // the particle loops of the set-up run there (ellipsoid initializations and `other` bunches
  // handed over on the device only).  dbunch_ is owned unless it was handed over.
  bool deviceBunch_ = false;
  mithra_bunch* dbunch_ = nullptr;
  bool ownsDbunch_ = false;
  long particleCount() const;
  void fill(mithra_dev* dev);
  void particles(double* charge11, long n) const;
  void setBeams(mithra_dev* dev);  // optical undulator and external-field beams of the push
  bool initialized_ = false;
  mithra_dev* dev_ = nullptr;

 protected:
  void check(int status, const char* what) const;
  void deviceBunch();
  void deviceBoost();
  void writePowerRows(const std::vector<double>& rows, const std::vector<double>& tb, unsigned count);
  // diagnostics state
  bool diagnosticsActive() const;
  bool fieldWritersDue() const;
  void fieldWriters();
  void bunchWriters();
  void fetchFull();
  Double nodeValue(FieldType f, long m) const;
  const std::vector<double>& fetchParticles();
  std::unique_ptr<std::ofstream> sfFile_, sbFile_;
  Double sfCe_ = 0, sfCb_ = 0, sfCa_ = 0, sfCf_ = 0;  // SampleField unit factors
  std::vector<std::string> vfName_;                   // file part of each visualization's base name
  std::vector<std::vector<std::unique_ptr<std::ofstream>>> scrFiles_;
  bool fullValid_ = false;
  std::vector<float> fullE_, fullB_;
  std::vector<double> fullA_, fullF_, partCache_, scrRec_;
  std::vector<int> scrScreen_;
  std::vector<long> scrIndex_;
  void powerVisualizeSetup();
  int vizFEL_ = -1;  // the FEL block whose power map is written, or -1
  unsigned vizNf_ = 0;
  Double vizW_ = 0, vizPc_ = 0;
};

// The two plugins of the reference, selected by the space-charge flag; on the GPU both run the
// same kernels, with the scalar potential enabled for FdTdSC.
class FdTd : public Solver {
 public:
  using Solver::Solver;
};
class FdTdSC : public Solver {
 public:
  using Solver::Solver;
};

}  // namespace mithra_host
