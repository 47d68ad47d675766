/*
 * mithra_gpu.h -- C-ABI of the MI355X FDTD/PIC time-march library (libmithra_gpu.so).
 *
 * This is the drop-in boundary under MITHRA's solver plugins.  In the reference the hot
 * path sits behind the abstract MITHRA::Solver (reference src/solver.h:20-345): the
 * FdTd / FdTdSC plugins override currentReset / currentUpdate / fieldUpdate /
 * fieldShift / fieldEvaluate (src/solver.h:139-160), while bunchUpdate and powerSample
 * are non-virtual Solver members called from Solver::solve() (src/solver.cpp:1300-1399).
 * Every entry point below replaces one of those calls, cited on its declaration.  The
 * host keeps the reference's job-file parser and set-up (the C++ drop-in in
 * mithra_amd/csrc/host), computes every coefficient the reference's way and hands them
 * over in the plain structs below; the library owns all device memory and one HIP
 * stream per handle.
 *
 * Conventions
 *  - All functions return 0 on success, a negative MITHRA_E* code on a validation error
 *    and MITHRA_E_HIP - hipError_t when the HIP runtime fails; mithra_strerror() gives
 *    the text.  The host wrapper turns a non-zero status into printmessage + exit(1),
 *    which is the reference's error behaviour (e.g. src/datainput.cpp:41).
 *  - Node index m = N1*N0*k + N1*i + j (y fastest, then x, then z), as the reference
 *    (src/solver.cpp:674, src/solver.cpp:2301-2315).  Host-side vector fields are the
 *    reference's FieldVector layout: 3 consecutive values per node ("AoS3").
 *  - Particles cross the boundary as the reference's Charge record, 11 doubles each:
 *    q, rnp[3], rnm[3], gb[3], e   (src/stdinclude.h:130-144).
 *  - A handle is not thread safe: one host thread per handle.  Work is ordered on the
 *    handle's stream; only the get/fetch calls and mithra_dev_synchronize wait for it.
 */
#ifndef MITHRA_GPU_H_
#define MITHRA_GPU_H_

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------- */
#define MITHRA_OK 0
#define MITHRA_E_ARG (-1)         /* invalid argument or shape                           */
#define MITHRA_E_UNSUPPORTED (-2) /* a job feature this library does not implement       */
#define MITHRA_E_STATE (-3)       /* call out of order (e.g. power sample before set-up) */
#define MITHRA_E_NOMEM (-4)       /* device allocation failed                            */
#define MITHRA_E_HIP (-1000)      /* MITHRA_E_HIP - hipError_t                           */

/* ---- enumerations: values equal the reference's enums (src/stdinclude.h:17-41) ------- */
enum { MITHRA_SOLVER_FD = 0, MITHRA_SOLVER_NSFD = 1 };
enum {
  MITHRA_SIGNAL_NEUMANN = 0, MITHRA_SIGNAL_GAUSSIAN = 1, MITHRA_SIGNAL_SECANT = 2,
  MITHRA_SIGNAL_FLATTOP = 3, MITHRA_SIGNAL_INVGAUSSIAN = 4
};
enum {
  MITHRA_SEED_PLANEWAVE = 0, MITHRA_SEED_PLANEWAVE_TRUNCATED = 1, MITHRA_SEED_GAUSSIANBEAM = 2,
  MITHRA_SEED_SUPERGAUSSIANBEAM = 3, MITHRA_SEED_STANDING_PLANEWAVE = 4,
  MITHRA_SEED_STANDING_PLANEWAVE_TRUNCATED = 5, MITHRA_SEED_STANDING_GAUSSIANBEAM = 6,
  MITHRA_SEED_STANDING_SUPERGAUSSIANBEAM = 7
  /* the standing-wave types 4..7 are beam types of optical undulators and external fields
     (src/beam.cc:280-496); the reference's Seed::fields does not handle them
     (src/classes.cpp:758-851) and a seed of those types is rejected with MITHRA_E_UNSUPPORTED */
};
enum { MITHRA_UNDULATOR_STATIC = 0, MITHRA_UNDULATOR_OPTICAL = 1 };
enum { MITHRA_PRECISION_FP64 = 0, MITHRA_PRECISION_FP32_BULK = 1 };

/* Mesh of one device domain after the Lorentz boost (Solver::initializeMesh,
 * src/solver.cpp:601-689; on one domain np_ = N2_ and k0_ = 0). */
typedef struct mithra_grid {
  int N0, N1, N2;                 /* nodes along x, y, z                                 */
  int solver;                     /* MITHRA_SOLVER_*  (Mesh::solver_)                   */
  int truncation_order;           /* 1 or 2 (Mesh::truncationOrder_)                     */
  int space_charge;               /* 0: FdTd, 1: FdTdSC (src/mithra.cpp:87-90)           */
  double xmin, xmax, ymin, ymax, zmin, zmax;
  double zp0, zp1;                /* owned z range, Solver::zp_                          */
  double dx, dy, dz;              /* boosted mesh resolution                             */
  double dt;                      /* field time step (Mesh::timeStep_)                   */
  double Lz;                      /* Mesh::meshLength_[2], period of the pmod ownership  */
} mithra_grid;

/* Update coefficients exactly as UpdateField carries them (src/database.h:178-213),
 * computed by the host in Solver::initializeField order (src/solver.cpp:695-824). */
typedef struct mithra_coeffs {
  double a[6];                    /* NSFD/FD interior coefficients, a[4] J, a[5] rho     */
  double bB[5], cB[5], dB[5];     /* Mur faces normal to x, y, z                        */
  double eE[5], fE[5], gE[5];     /* Mur edges along z, x, y                             */
  double hC[17];                  /* Mur corners                                         */
  double alpha;                   /* AdvanceField::alpha_                                */
  double beta;                    /* AdvanceField::beta_ (= beta / alpha)                */
} mithra_coeffs;

/* Seed after Solver::setSimulationParameters scaling (src/solver.cpp:74-120) and the boost
 * parameters copied into it (src/solver.cpp:211-212, :330). */
typedef struct mithra_seed {
  int seed_type;                  /* MITHRA_SEED_*                                       */
  int signal_type;                /* MITHRA_SIGNAL_*                                     */
  int order[2];                   /* super-gaussian orders                               */
  unsigned nR;                    /* flat-top rising cycles                              */
  double amplitude;               /* Seed::amplitude_ ; |amplitude| <= 1e-50 means no seed*/
  double position[3], direction[3], polarization[3];
  double radius[2];
  double t0, s, f0, cep;          /* Signal members after scaling                        */
  double sigma_invg[2];
  double gamma, beta, dt_shift, c0;
} mithra_seed;

/* One undulator module (class Undulator, src/classes.h) after sorting and shifting.  An
 * optical undulator's laser beam is the next unused entry of the optical beams given to
 * mithra_dev_set_beams, in table order. */
typedef struct mithra_undulator {
  int type;                       /* MITHRA_UNDULATOR_STATIC or MITHRA_UNDULATOR_OPTICAL  */
  unsigned length;                /* number of periods                                   */
  double k, lu, rb, dist, theta;
} mithra_undulator;

/* An analytic laser beam the bunch push adds in the lab frame: the beam of an optical undulator
 * (class Undulator) or of an EXTERNAL-FIELD electromagnetic wave (class ExtField), after
 * Undulator::initialize / ExtField::initialize (src/classes.cpp:1007-1102, :1162-1253) and the
 * scaling of Solver::setSimulationParameters (src/solver.cpp:88-123). */
typedef struct mithra_beam {
  int beam_type;                  /* MITHRA_SEED_* (all eight types)                     */
  int signal_type;                /* MITHRA_SIGNAL_*                                     */
  int order[2];                   /* super-gaussian orders                               */
  unsigned nR;                    /* flat-top rising cycles                              */
  double amplitude;               /* field amplitude, amplitude_                          */
  double position[3], direction[3], polarization[3];
  double radius[2], zR[2], l;     /* radius_, Rayleigh lengths zR_, wavelength l_          */
  double t0, s, f0, cep;          /* Signal members after scaling                        */
  double sigma_invg[2];
} mithra_beam;

/* Constants of one bunch sub-step (Solver::bunchUpdate, src/solver.cpp:1424-1549, with the
 * UpdateBunch values of Solver::initializeBunchUpdate, src/solver.cpp:1050-1059). */
typedef struct mithra_bunch_params {
  double gamma, beta, c0;         /* frame Lorentz factor, velocity, speed of light       */
  double dt_shift;                /* Solver::dt_                                         */
  double r1, r2, dtb;             /* UpdateBunch::r1, r2, dtb                             */
} mithra_bunch_params;

/* Radiation-power sampling (Solver::initializePowerSample, src/radiation.cpp:18-121).
 * Only the sampling planes inside [zp0, zp1) are handed over, in the reference's order. */
typedef struct mithra_power_params {
  unsigned N;                     /* number of sampling planes                           */
  unsigned Nl;                    /* number of normalised frequencies                    */
  unsigned Nf;                    /* ring length (time samples per DFT)                   */
  double pc;                      /* SampleRadiationPower::pc                            */
  double gamma, beta, c0;
  const double* z;                /* N boosted plane positions                            */
  const double* w;                /* Nl angular frequencies                               */
} mithra_power_params;

typedef struct mithra_dev mithra_dev;

const char* mithra_strerror(int status);

/* Allocate the device domain (A^{n+1}, A^n, A^{n-1}, phi levels when space_charge, E/B).
 * Replaces Solver::initializeMesh's allocations (src/solver.cpp:644-658) and
 * FdTd::FdTd / FdTdSC::FdTdSC (src/fdtd.cpp:10-17, src/fdtdSC.cpp:10-17).  All fields
 * start at zero. precision: MITHRA_PRECISION_*. */
int mithra_dev_create(const mithra_grid* grid, const mithra_coeffs* coeffs, int precision,
                      mithra_dev** out);
int mithra_dev_destroy(mithra_dev* dev);
int mithra_dev_synchronize(mithra_dev* dev);

/* Potentials.  level 0 = A^n (an_), 1 = A^{n-1} (anm1_), 2 = A^{n+1} (anp1_, which also
 * holds J between currentUpdate and fieldUpdate, src/fdtd.cpp:25-27).  a is AoS3 with
 * N0*N1*N2 nodes; phi (N0*N1*N2 values) is used only when space_charge and may be NULL.
 * The set call replaces the seed initialisation of Solver::initializeField
 * (src/solver.cpp:828-839) as a copy from the host. */
int mithra_dev_set_potentials(mithra_dev* dev, int level, const double* a, const double* phi);
int mithra_dev_get_potentials(mithra_dev* dev, int level, double* a, double* phi);

/* E and B (float AoS3) as last evaluated: Solver::en_, bn_ (src/solver.h:244-245). */
int mithra_dev_get_fields(mithra_dev* dev, float* e, float* b);

/* Particle list (Solver::chargeVectorn_, src/solver.h:271) as n Charge records. */
int mithra_dev_set_particles(mithra_dev* dev, const double* charge11, long n);
int mithra_dev_get_particles(mithra_dev* dev, double* charge11, long n);
long mithra_dev_particle_count(const mithra_dev* dev);

/* The analytic seed on the device at n listed points, as the TF/SF injection evaluates it.
 * signal_evaluate: out[q] = Signal::self(t[q], phase[q]) (src/classes.cpp:534-575) for the
 * seed's signal.  seed_evaluate: a[3q..3q+2] = Seed::fields(r[3q..3q+2], time[q])
 * (src/classes.cpp:740-855), the boosted-frame vector potential at boosted-frame point r; each
 * point on its own, so the super-gaussian type, whose reference form adds into the caller's
 * previous output, returns MITHRA_E_UNSUPPORTED (as the field update does).  Host arrays in
 * and out; the calls wait for the device. */
int mithra_signal_evaluate(const mithra_seed* seed, const double* t, const double* phase, long n, double* out);
int mithra_seed_evaluate(const mithra_seed* seed, const double* r, const double* time, long n, double* a);

/* One field step: NSFD/FD interior update of A (and phi), TF/SF seed injection, Mur faces,
 * edges and corners, then E/B on every node a particle or sampler may read.
 * Replaces FdTd::fieldUpdate (src/fdtd.cpp:231-800) / FdTdSC::fieldUpdate
 * (src/fdtdSC.cpp:260-1088) together with the lazy FdTd::fieldEvaluate
 * (src/fdtd.cpp:818-845, src/fdtdSC.cpp:1110-1145).  time = Solver::time_. */
int mithra_dev_field_update(mithra_dev* dev, double time, const mithra_seed* seed);

/* E/B from the potentials as they stand, without advancing them: what the reference's lazy
 * FdTd::fieldEvaluate (src/fdtd.cpp:818-845) returns while fieldUpdate is not running, i.e.
 * during the particle-only march before time zero (src/solver.cpp:1232-1291). */
int mithra_dev_field_evaluate(mithra_dev* dev);

/* rnm = rnp for every particle (src/solver.cpp:1311-1312). */
int mithra_dev_bunch_begin(mithra_dev* dev);

/* One bunch sub-step: undulator field, gather of E/B, Boris push and drift.
 * Replaces Solver::bunchUpdate (src/solver.cpp:1424-1576) with Solver::undulatorField
 * (src/solver.cpp:1798-1880) and Solver::staticUndulator (src/beam.cc:14-76).
 * time_bunch = Solver::timeBunch_ at this sub-step. */
int mithra_dev_bunch_update(mithra_dev* dev, const mithra_bunch_params* p,
                            const mithra_undulator* und, int n_und, double time_bunch);

/* rnm = rnp followed by n_sub bunch sub-steps in one launch: the loop
 * `for (t = 0; t < nUpdateBunch_; t += 1) { bunchUpdate(); timeBunch_ += dt_b; }` of
 * Solver::solve() (src/solver.cpp:1311-1321).  Sub-step s runs at time_bunch accumulated s
 * times with bunch_time_step, in the reference's order.  The E/B grid is constant over a
 * field step, so each particle is loaded once and advanced n_sub times in registers. */
int mithra_dev_bunch_advance(mithra_dev* dev, const mithra_bunch_params* p, const mithra_undulator* und,
                             int n_und, double time_bunch, double bunch_time_step, int n_sub);

/* The laser beams of the push: optical[q] is the beam of the q-th optical undulator of the
 * undulator table, external[q] the q-th EXTERNAL-FIELD wave.  Each sub-step adds, per particle,
 * every undulator's field in table order (static: Solver::staticUndulator; optical: the beam's
 * lab-frame E and B transformed into the bunch frame, Solver::undulatorField,
 * src/solver.cpp:1825-1877) and then every external wave's (Solver::externalField,
 * src/solver.cpp:1886-1947), before the grid's E/B, with the beam fields of src/beam.cc:80-496.
 * Replaces the previous beams; n = 0 removes them.  Two uses of state the reference leaves
 * undefined are given their evident meaning: a standing-wave optical undulator takes its
 * backward-wave time from its own beam (the reference reads the value the previous external
 * wave left in UpdateBunchParallel::tlm, src/solver.cpp:1836-1838), and a standing super-gaussian
 * beam's phase uses its wavelength l_ (the reference reads the never-set
 * UpdateBunchParallel::l, src/beam.cc:466). */
int mithra_dev_set_beams(mithra_dev* dev, const mithra_beam* optical, int n_optical, const mithra_beam* external,
                         int n_external);

/* Number of particles that were in the z-range but outside the transverse domain during
 * the bunch updates since the last call (UpdateBunch::nt, src/solver.cpp:1500-1505,
 * :1572-1575), and the number that crossed the periodic z boundary (which the reference
 * duplicates on one rank, src/solver.cpp:1544-1568).  Both counters are reset. */
int mithra_dev_bunch_counters(mithra_dev* dev, long* transverse_out, long* z_crossings);

/* Radiation power. setup allocates the ring of lab-frame fields (SampleRadiationPower::fdt)
 * and the DFT phases; sample computes P for every plane and frequency at step n_time and
 * stores them in history row `row` (row < rows given to setup).  fetch copies `count` rows
 * starting at `first` to the host as [row][plane][frequency].  Replaces
 * Solver::initializePowerSample / Solver::powerSample (src/radiation.cpp:18-232). */
int mithra_dev_power_setup(mithra_dev* dev, const mithra_power_params* p, unsigned rows);
int mithra_dev_power_sample(mithra_dev* dev, unsigned n_time, unsigned row);
int mithra_dev_power_fetch(mithra_dev* dev, unsigned first, unsigned count, double* out);

/* ---- diagnostics: what the host's file writers print, gathered on the device ------------- */

/* E/B as last evaluated, A^n and phi^n at n listed nodes (m = N1 N0 k + N1 i + j): the values
 * FdTd::fieldSample (src/fdtd.cpp:851-950) and the in-plane visualizations
 * (src/fdtd.cpp:1128-1545) read from en_, bn_, an_ and fn_ (FdTdSC, src/fdtdSC.cpp:1147-1262)
 * to interpolate.  e, b, a: 3 per node; phi: 1 per node (0 without space charge); any may be
 * NULL.  A node outside the mesh is MITHRA_E_ARG.  Waits for the device. */
int mithra_dev_sample_nodes(mithra_dev* dev, const long* nodes, long n, float* e, float* b, double* a,
                            double* phi);

/* The charge-weighted sums of Solver::bunchSample (src/solver.cpp:1582-1610) over every
 * particle: out[0] = sum q, out[1..3] = sum q r, out[4..6] = sum r^2 q, out[7..9] = sum q gb,
 * out[10..12] = sum gb^2 q.  Summed on a fixed grid in a fixed order (the same particle state
 * gives the same sums every time); the reference sums in list order, so the two agree to
 * rounding.  Waits for the device. */
int mithra_dev_bunch_sums(mithra_dev* dev, double* out13);

/* Lab-frame screens (Solver::screenProfile, src/solver.cpp:2205-2257): the particles whose lab
 * position passed screen s (lz[s], lengths) during the field step that just ended, with the
 * record the reference writes for each: x, y and lab time interpolated to the screen, gb_x,
 * gb_y and the lab-frame gb_z.  time_bunch = Solver::timeBunch_ after the step's sub-steps,
 * time_step = the field step, dt_shift = Solver::dt_.  count receives the number of crossings;
 * the first min(count, cap) are written (screen index, particle load index, 6 doubles each) in
 * no particular order.  Waits for the device. */
typedef struct mithra_screen_params {
  double gamma, beta, c0, time_bunch, time_step, dt_shift;
} mithra_screen_params;
int mithra_dev_screen_crossings(mithra_dev* dev, const double* lz, int n_screens, const mithra_screen_params* p,
                                long cap, long* count, int* screen, long* index, double* rec6);

/* Power map of one plane (Solver::initializePowerVisualize / powerVisualize,
 * src/radiation.cpp:238-450).  setup: the plane's cell k and fraction dzr (modf of its boosted
 * position over dz), the ring length Nf, the DFT phases cos(w j dt), sin(w j dt) for j < Nf, the
 * power constant pc and the frame; the ring starts at zero, and the lazy E/B evaluation covers
 * the plane from then on.  sample: the lab-frame Ex, Ey, Bx, By of the plane into ring row
 * n_time % Nf (every field step, after the bunch push).  fetch: the power of every node of the
 * plane, out[N1 i + j] (zero on the boundary nodes), from the ring as it stands. */
typedef struct mithra_power_map_params {
  int k;
  double dzr;
  unsigned Nf;
  const double* cos_wt;
  const double* sin_wt;
  double pc, gamma, beta, c0;
} mithra_power_map_params;
int mithra_dev_power_map_setup(mithra_dev* dev, const mithra_power_map_params* p);
int mithra_dev_power_map_sample(mithra_dev* dev, unsigned n_time);
int mithra_dev_power_map_fetch(mithra_dev* dev, double* out);

/* Rotate A^{n-1} <- A^n <- A^{n+1} (and phi): FdTd::fieldShift (src/fdtd.cpp:806-812). */
int mithra_dev_field_shift(mithra_dev* dev);

/* Zero J (held in A^{n+1}) and rho: FdTd::currentReset (src/fdtd.cpp:23-32),
 * FdTdSC::currentReset (src/fdtdSC.cpp:23-36). */
int mithra_dev_current_reset(mithra_dev* dev);

/* Charge-conserving zigzag deposition of J (and rho at rnp) from rnm -> rnp:
 * FdTd::currentUpdate (src/fdtd.cpp:38-185), FdTdSC::currentUpdate (src/fdtdSC.cpp:38-210).
 * On one domain FdTd::currentCommunicate (src/fdtd.cpp:191-225) does nothing else. */
int mithra_dev_current_update(mithra_dev* dev);

/* One whole iteration of the Solver::solve() main loop (src/solver.cpp:1303-1399) on the
 * device, without the host-side diagnostics: field_update, bunch_advance, power_sample (skipped
 * when power_row < 0), field_shift, current_reset, current_update. */
int mithra_dev_step(mithra_dev* dev, double time, const mithra_seed* seed, const mithra_bunch_params* p,
                    const mithra_undulator* und, int n_und, double time_bunch, double bunch_time_step,
                    int n_sub, unsigned n_time, int power_row);

/* lazy = 1: each field update evaluates E/B only on the z-planes the bunch can read during the
 * coming push (its cell extent after the previous push, widened by the one cell a particle can
 * move per field step) and on the power-sampling planes -- the reference's rule of evaluating
 * E/B only where particles and samplers need it (FdTd::fieldEvaluate, src/fdtd.cpp:818-845,
 * called per touched node; flags pic_, src/solver.h:241).  Elsewhere the E/B array keeps older
 * values, as the reference's en_/bn_ do.  lazy = 0 (default) evaluates every node, for callers
 * that read the whole field with mithra_dev_get_fields.  Until a push has measured the bunch
 * (after mithra_dev_set_particles) every node is evaluated either way. */
int mithra_dev_set_lazy_fields(mithra_dev* dev, int lazy);

/* Deposits between two reorders of the particle arrays into the cell-row order of the
 * deposition (default 16; 0 never reorders).  The deposition sorts its own index lists every
 * step; reordering the arrays only makes those lists walk memory in sequence.  No per-particle
 * result changes, and mithra_dev_get_particles always returns the records in the order they
 * were set.  No counterpart in the reference (its particles live in a std::list,
 * src/solver.h:271). */
int mithra_dev_set_reorder_interval(mithra_dev* dev, int interval);

/* Fixed-order current deposition (on = 1; default 0).  The default deposit sums each node's
 * contributions in LDS with atomic adds, so FP64 runs agree to rounding but not bit for bit.
 * In fixed order every node sums the weights of the halves in its eight cells in an order set
 * by the particles alone (cell planes K-1, K, rows I-1, I, columns J-1, J; in each cell the
 * rnp halves, then the rnm halves, in particle-array order), so a run repeated from the same
 * state with the same reorder interval repeats every bit.  It costs a cell-keyed sort, the
 * weights of every half stored (12-20 doubles) and a pass over every node; pipelined steps run
 * as plain steps meanwhile.  MITHRA_E_STATE mid-way through a pipelined run, MITHRA_E_UNSUPPORTED
 * for a mesh of 2^32 cells or more.  A batch refuses a
 * fixed-order domain as an instance.  Replaces the reference's particle-order sums
 * (FdTd::currentUpdate, src/fdtd.cpp:38-185) by another fixed order; no counterpart switch in
 * the reference, whose sums are always in list order. */
int mithra_dev_set_deterministic(mithra_dev* dev, int on);

/* Tile shapes of the two marched field kernels (interior update, E/B evaluation): variant v
 * has `rows` x-rows of 64 y-nodes per workgroup and marches `planes` z-planes; variants 7-9
 * stage the interior update's A^n planes in LDS.  When no TF/SF seed is injected, the interior
 * update of every shape also updates the Mur y faces, which the boundary pass then skips.
 * Every shape computes identical values; only the speed differs. */
#define MITHRA_TILE_VARIANTS 10
int mithra_tile_shape(int variant, int* rows, int* planes);
int mithra_dev_set_tiles(mithra_dev* dev, int interior, int eval);
/* Autotuner: times `reps` launches of every tile shape of both kernels (the interior update
 * together with the Mur boundaries, the E/B evaluation alone) on this domain's mesh and keeps
 * the fastest of each.  The domain's state is left as it was (the interior update
 * runs on a saved copy of the J it consumes).  chosen[2] (optional) receives the two variant
 * indices, ms[2 * MITHRA_TILE_VARIANTS] (optional) the mean time of each shape.  No
 * counterpart in the reference (a CPU code with no launch shapes). */
int mithra_dev_autotune(mithra_dev* dev, int reps, int* chosen, float* ms);

/* ---- bunches generated and boosted on the device ------------------------------------------
 * A mithra_bunch holds Charge records on the device.  The generators append lab-frame records;
 * the boost calls are the particle loops of Solver::lorentzBoostBunch (src/solver.cpp:263-423)
 * and Solver::shiftBackInTime (src/solver.cpp:1185-1206), the host keeping the scalar formulas
 * between them; mithra_dev_take_bunch then makes the records a domain's particles (device to
 * device).  The drop-in host uses them when a job asks for device generation. */
typedef struct mithra_bunch mithra_bunch;
enum { MITHRA_DISTRIBUTION_UNIFORM = 0, MITHRA_DISTRIBUTION_GAUSSIAN = 1 };

/* One `ellipsoid` bunch-initialization at one position (BunchInitialize, src/database.h, after
 * Solver::setSimulationParameters scaled it): Bunch::initializeEllipsoid (src/classes.cpp:104-298)
 * with the halton generator.  first_index = the records this initialization already produced at
 * earlier positions (the reference's Np0); lambda = BunchInitialize::lambda_ (0 without an
 * undulator: one record per draw instead of four beamlets). */
typedef struct mithra_ellipsoid {
  long n;                         /* number-of-particles (rounded up to a multiple of four)   */
  unsigned first_index;
  int distribution;               /* MITHRA_DISTRIBUTION_*                                    */
  int shot_noise;
  double charge;                  /* cloudCharge_ (electrons)                                 */
  double gamma, beta_vector[3];   /* initialGamma_, betaVector_                               */
  double position[3], sigma_position[3], sigma_gb[3];
  double tran_trunc, long_trunc;  /* tranTrun_, longTrun_                                     */
  double lambda, bunching, bunching_phase;  /* lambda_, bF_, bFP_ (degrees)                   */
} mithra_ellipsoid;

/* A quiet-start flat-top slice for `type = other` bunches (no counterpart in the reference,
 * which leaves such bunches to an external program, src/solver.cpp:1164-1165): n records (a
 * multiple of four) of charge/n electrons; n/4 base particles uniform in z over [z_center -
 * length/2, z_center + length/2], Gaussian in x, y (sigma_r) and gb_x, gb_y (sigma_gb_perp),
 * gb_z = sqrt(gamma^2 - 1)(1 + energy_spread N(0,1)); four beamlets lambda/4 apart per base
 * particle, and per lambda bucket one shot-noise modulation of mean-square amplitude
 * 2 / electrons_per_bucket with a uniform phase.  Every record is a function of (seed, index). */
typedef struct mithra_flattop {
  long n;
  double charge, gamma, lambda, length, z_center, sigma_r, sigma_gb_perp, energy_spread;
  double electrons_per_bucket;
} mithra_flattop;

int mithra_bunch_create(mithra_bunch** out);
int mithra_bunch_destroy(mithra_bunch* bunch);
long mithra_bunch_count(const mithra_bunch* bunch);
/* the records as n Charge records (load order) on the host */
int mithra_bunch_get(mithra_bunch* bunch, double* charge11, long n);
/* appends; *added (optional) = records appended */
int mithra_bunch_ellipsoid(mithra_bunch* bunch, const mithra_ellipsoid* spec, long* added);
int mithra_bunch_flattop(mithra_bunch* bunch, const mithra_flattop* spec, unsigned long long seed);
/* first loop of lorentzBoostBunch: z *= gamma, gb_z = gamma g (beta_z - beta); *zmax = max z */
int mithra_bunch_boost(mithra_bunch* bunch, double gamma, double beta, double* zmax);
/* second loop: r += gb / g (z - zu) beta */
int mithra_bunch_drift(mithra_bunch* bunch, double beta, double zu);
/* sums of z and beta_z, and min z (per-workgroup partials added in a fixed order) */
int mithra_bunch_moments(mithra_bunch* bunch, double* z_sum, double* bz_sum, double* z_min);
/* z -= shift (the optimised bunch position, src/solver.cpp:369-377) */
int mithra_bunch_shift(mithra_bunch* bunch, double shift);
/* r -= c0 time_shift / g gb (Solver::shiftBackInTime) */
int mithra_bunch_time_shift(mithra_bunch* bunch, double c0, double time_shift);
/* the bunch's records become the domain's particles (mithra_dev_set_particles from device
 * memory); the bunch keeps its copy */
int mithra_dev_take_bunch(mithra_dev* dev, const mithra_bunch* bunch);

/* ---- batched parameter scans ------------------------------------------------------------
 * A batch marches many independent domains of uneven size together: each step is one fixed
 * sequence of launches whose persistent workgroups walk tables of (instance, tile) entries, and
 * it replays as one captured HIP graph.  Each instance is an ordinary mithra_dev, created and
 * filled through the calls above (potentials, particles, power set-up) and read back through
 * them afterwards (power fetch, particles, potentials); the batch does not own the handles,
 * which must outlive it and must not be used while it runs.  Supported: with or without space
 * charge, TF/SF seeds of every type but the super-gaussian (mithra_batch_seed), static
 * undulators, rows of the mesh that fit the row-tiled deposit.  Each instance computes exactly
 * its single-domain march (mithra_dev_step with a fused push and power row = step): only the
 * deposit's summation order differs.  No counterpart in the reference (one job per process
 * run). */
typedef struct mithra_batch mithra_batch;
int mithra_batch_create(mithra_dev* const* devs, int n, mithra_batch** out);
int mithra_batch_destroy(mithra_batch* batch);
/* Describes instance i's march: start time and bunch time, the bunch time step and sub-steps
 * per field step, the number of field steps, the push constants and undulators; and its
 * bunching-factor samples b = |sum q exp(i k z)| / |sum q| taken after the push of every step
 * that is a multiple of `bunching_every`, into slots 0 .. bunching_slots-1. */
int mithra_batch_instance(mithra_batch* batch, int i, double time, double time_bunch, double bunch_time_step,
                          int n_sub, long n_steps, const mithra_bunch_params* p, const mithra_undulator* und,
                          int n_und, double bunching_k, long bunching_every, int bunching_slots);
/* Instance i's analytic seed, injected through the TF/SF planes every step at the instance's
 * own time (FdTd::fieldUpdate, src/fdtd.cpp:307-373), as mithra_dev_field_update does; after
 * mithra_batch_instance, before the first run.  The super-gaussian type is
 * MITHRA_E_UNSUPPORTED here (its running vector is a single-domain path). */
int mithra_batch_seed(mithra_batch* batch, int i, const mithra_seed* seed);
/* Tile shape of the field kernels (a variant of mithra_tile_shape) and whether steps replay as
 * a captured graph (1) or are launched one by one (0); before the first run. */
int mithra_batch_set_tiles(mithra_batch* batch, int variant, int use_graph);
/* Times the field kernels (interior, Mur boundaries, E/B) of the whole batch for each tile
 * shape and keeps the fastest (the
 * domains' state is left as it was); before the first run.  ms[MITHRA_TILE_VARIANTS]
 * (optional): per shape. */
int mithra_batch_autotune(mithra_batch* batch, int reps, int* chosen, float* ms);
/* Advances every unfinished instance by up to `steps` field steps (instances stop at their own
 * n_steps); *done receives the steps taken.  Returns with each mithra_dev updated. */
int mithra_batch_run(mithra_batch* batch, long steps, long* done);
int mithra_batch_steps(const mithra_batch* batch, int i, long* step);
int mithra_batch_bunching(mithra_batch* batch, int i, double* out);

/* ---- learned surrogate: an ensemble of multilayer perceptrons --------------------------
 * `members` networks of `layers` fully connected layers of widths dims[0] -> dims[1] -> ... ->
 * dims[layers], SiLU after every layer but the last.  Weights live in FP32 (the Adam master)
 * and BF16 (what the products read); every product is a BF16 x BF16 -> FP32 MFMA GEMM.  The
 * parameters of one member are laid out W_0 (dims[0] x dims[1], row-major), b_0, W_1, b_1, ...
 * Inputs and outputs cross the boundary as FP32 host arrays, n rows each.  No counterpart in
 * the reference (a CPU FDTD/PIC code). */
typedef struct mithra_mlp mithra_mlp;
int mithra_mlp_create(int members, int layers, const int* dims, unsigned long long seed, mithra_mlp** out);
int mithra_mlp_destroy(mithra_mlp* mlp);
long mithra_mlp_param_count(const mithra_mlp* mlp);  /* per member */
int mithra_mlp_set_params(mithra_mlp* mlp, int member, const float* params);  /* member -1: all */
int mithra_mlp_get_params(mithra_mlp* mlp, int member, float* params);
/* y: members x n x dims[layers] */
int mithra_mlp_forward(mithra_mlp* mlp, const float* x, long n, float* y);
/* mean squared error over n x dims[layers] targets and its gradient for one member */
int mithra_mlp_gradients(mithra_mlp* mlp, int member, const float* x, const float* t, long n, float* grads,
                         float* loss);
/* one Adam step (0.9, 0.999, 1e-8) of every member on one mini-batch; loss[members] before it */
int mithra_mlp_train_step(mithra_mlp* mlp, const float* x, const float* t, long n, float lr, float* loss);
/* the training set kept on the device: n rows of x (n x dims[0], stored as BF16) and targets t
 * (n x dims[layers]); replaces any earlier set */
int mithra_mlp_set_data(mithra_mlp* mlp, const float* x, const float* t, long n);
/* `epochs` epochs of Adam on the resident set: each epoch a fresh permutation (from `seed`) cut
 * into n / batch mini-batches, every member stepping on each; learning rate
 * max(lr_max (1 + cos(pi epoch / epochs)) / 2, lr_min).  An epoch is one captured HIP graph
 * replayed without the host.  loss[epochs] (optional): each epoch's mean loss over members and
 * mini-batches.  MITHRA_E_STATE without a data set or with batch > n. */
int mithra_mlp_train_epochs(mithra_mlp* mlp, int epochs, int batch, float lr_max, float lr_min,
                            unsigned long long seed, float* loss);
/* one product of the surrogate's GEMM on host data (BF16 as 16-bit patterns): C (M x N, FP32,
 * row-major) = op(A) op(B) with op(A)(m,k) = A[m*sam + k*sak], op(B)(k,n) = B[k*sbk + n*sbn];
 * tile 0 (chosen by size), 1 (64 x 64 workgroups) or 2 (128 x 128); reps timed launches, mean
 * time in *ms.  MITHRA_E_ARG when op(A) or op(B) reaches past a_len or b_len elements. */
int mithra_mlp_gemm(const unsigned short* A, long sam, long sak, long a_len, const unsigned short* B, long sbk,
                    long sbn, long b_len, float* C, int M, int N, int K, int tile, int reps, float* ms);

/* Bunching factor b = |sum q exp(i k z)| / |sum q| over all particles at their current boosted-
 * frame z.  With k = 2 pi / lambda', lambda' = lambda_u / (gamma lambda_norm) the radiation
 * wavelength in the frame (the one the power sampler uses, src/radiation.cpp:91-92), this is
 * the FEL bunching of the beam; the phase k z - omega t is Lorentz invariant, so b needs no
 * lab-frame transform.  No counterpart in the reference (its radiation diagnostics only). */
int mithra_dev_bunching(mithra_dev* dev, double k, double* b_out);

/* mithra_dev_bunching without any device allocation: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  Its per-workgroup partial sums (3 x 512 doubles, the same
 * fixed grid as mithra_dev_bunching, so every particle count the library accepts fits) live in a
 * static device array that each GPU holds until the process ends; b equals mithra_dev_bunching's
 * bit for bit.  Every call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (several host threads, or handles on several streams of one GPU) race. */
int mithra_dev_bunching_noalloc(mithra_dev* dev, double k, double* b_out);

/* mithra_dev_step with HIP events recorded on the handle's stream between its phases; ms[i]
 * receives the duration of phase i: 0 interior update, 1 TF/SF, 2 Mur boundaries, 3 E/B
 * evaluation, 4 bunch sub-steps, 5 power sample, 6 current reset, 7 current deposition.
 * Waits for the step to finish.  No counterpart in the reference: the measurement hook. */
#define MITHRA_STEP_PHASES 8
int mithra_dev_step_timed(mithra_dev* dev, double time, const mithra_seed* seed, const mithra_bunch_params* p,
                          const mithra_undulator* und, int n_und, double time_bunch, double bunch_time_step,
                          int n_sub, unsigned n_time, int power_row, float* ms);

/* mithra_dev_step pipelined across calls.  The interior update sums its terms left to right
 * with the source a4 J last (src/database.cpp:44-134), so A^{n+2} = X + a4 J^{n+1} splits
 * exactly, and X reads the same two levels (A^{n+1}, A^n) as the E/B evaluation of step n.
 * more = 1: after the seed injection and Mur boundaries of A^{n+1}, one kernel evaluates E/B and
 * writes X into the free level; the deposition then writes X + a4 J at the interior nodes
 * instead of J, so the next call starts from the marched A^{n+2} and skips its interior update.
 * Per node and step this saves one read each of A^n, A^{n+1} and J (72 B in FP64) for one read
 * of X (24 B).  more = 0 ends the run: that step evaluates E/B alone and deposits J, leaving
 * exactly the state mithra_dev_step leaves.  While a run is pending the potentials are not in
 * the reference's between-step state (J is not held), so mithra_dev_get/set_potentials,
 * field_update, field_evaluate, field_shift, current_reset, current_update, autotune and
 * batch_add return MITHRA_E_STATE; particles, E/B, power samples and counters stay readable.
 * Under space charge phi is a fourth component: its X takes the same stencil, the fused E/B
 * adds the gradient of phi^n, and the deposit writes X + a5 rho for it (src/fdtdSC.cpp).
 * Pipelining needs FP64, particles, the row-tiled deposit and no fixed-order deposit;
 * otherwise more = 1 runs as more = 0 (and nothing is pending).  Its results are bit-identical to
 * mithra_dev_step's (the deposit's summation order aside).  ms: NULL, or MITHRA_STEP_PHASES durations as
 * mithra_dev_step_timed gives them (phase 0 is zero after a pipelined step; phase 3 is the
 * fused kernel when more = 1).  No counterpart in the reference: its loop body reordered. */
int mithra_dev_step_pipelined(mithra_dev* dev, double time, const mithra_seed* seed, const mithra_bunch_params* p,
                              const mithra_undulator* und, int n_und, double time_bunch, double bunch_time_step,
                              int n_sub, unsigned n_time, int power_row, int more, float* ms);
/* 1 while a pipelined run is pending (the last step had more = 1 and was pipelined), else 0. */
int mithra_dev_pipeline_pending(const mithra_dev* dev);
/* Tile shapes of the pipelined step's fused kernel: `rows` x-rows of 64 y-nodes, `planes`
 * z-planes per workgroup; every shape computes identical values.  autotune_fused times each
 * (reps launches after one warm-up, ms[v] the mean) on the handle's current levels, keeps the
 * fastest and restores the level and the E/B array it writes.  No counterpart in the reference. */
#define MITHRA_FUSED_TILE_VARIANTS 4
int mithra_fused_tile_shape(int variant, int* rows, int* planes);
int mithra_dev_set_fused_tile(mithra_dev* dev, int variant);
int mithra_dev_autotune_fused(mithra_dev* dev, int reps, int* chosen, float* ms);

/* Device-to-device copy rate of a plain streaming kernel over two buffers of `bytes` each
 * (read + write bytes per second, GB/s), the measured HBM ceiling the roofline fractions are
 * also reported against.  No counterpart in the reference. */
int mithra_copy_bandwidth(unsigned long bytes, int reps, double* gbps);
/* The same with elem_bytes (8, 16 or 32) bytes moved per lane and load. */
int mithra_copy_bandwidth_elem(unsigned long bytes, int reps, int elem_bytes, double* gbps);
/* Rate (GB/s, all bytes moved) of the interior stencil's stream pattern without its arithmetic:
 * three buffers of `bytes` read and one written per launch.  The measured ceiling of the
 * stencil's access pattern.  No counterpart in the reference. */
int mithra_stream_bandwidth(unsigned long bytes, int reps, double* gbps);

#ifdef __cplusplus
}
#endif

#endif /* MITHRA_GPU_H_ */
