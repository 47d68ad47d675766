// Shared declarations of the MI355X (gfx950) FDTD/PIC library.
//
// Device layout of one domain (one GPU holds the whole mesh):
//   potentials  SoA per component, three time levels rotated by index:  A[level][c][m],
//               phi[level][m]; m = N1*N0*k + N1*i + j (y fastest, then x, then z), the
//               reference's node order, so every z-march reads contiguous y rows.
//   E/B         two float4 per node: {Ex, Ey, Ez, Bx}, {By, Bz, 0, 0}; the trilinear gather
//               reads a node's six values with two 16-byte loads.
//   particles   SoA doubles: x y z | xm ym zm | gx gy gz | q | e  (the Charge record split by
//               member so that every kernel streams only what it uses).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../../include/mithra_gpu.h"

namespace mithra {

// Pointers into device memory (address space 1).  Kernel bodies receive their buffers through
// structures (field sets, batch views); casting them back to global pointers lets the compiler
// use global_load/store instead of flat accesses, which also count against the LDS wait counter.
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
// Four floats as a builtin vector (float4's member operators cannot act on address space 1).
typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));

// E/B of node m: six floats {Ex, Ey | Ez, Bx | By, Bz}, 24 bytes, read and written as three
// 8-byte pairs at eb[3 m], eb[3 m + 1], eb[3 m + 2] (f2 units).
constexpr int kEBPairs = 3;
template <typename T>
__device__ inline gptr<T> G(T* p) {
  return (gptr<T>)p;
}

// Constants of the reference (src/stdinclude.h:43-50); PI is truncated on purpose.
constexpr double kPI = 3.1415926535;
constexpr double kEC = 1.602e-19;
constexpr double kEM = 9.109e-31;

enum Particle { PX, PY, PZ, PXM, PYM, PZM, PGX, PGY, PGZ, PQ, PE, PCOUNT };

struct FieldSet {           // one time level of every potential component
  void* c[4];               // A_x, A_y, A_z, phi (phi only with space charge)
};

// Launch constants of the kernels, shared by the single-domain launchers and the batch views.
struct MurCoeffs {
  double bB[5], cB[5], dB[5], eE[5], fE[5], gE[5], hC[17];
};

// Mur y faces (j = 0 and j = N1-1) that the interior kernels (every tile shape) update together
// with the interior when no TF/SF write comes between the two: bit 0 the j = 0 face, bit 1 the
// j = N1-1 face, the latter only when node N1-2 shares its 64-node tile row.
__host__ __device__ inline int yface_mask(int, int N1) {
  if (N1 < 3) return 0;
  return 1 | ((N1 - 1) % 64 != 0 ? 2 : 0);
}

// Face nodes the Mur face kernel updates: both x faces, the y faces not in `ymask`, both z faces.
__host__ __device__ inline long mur_face_count(int N0, int N1, int N2, int ymask) {
  const long ny = (long)(N0 - 2) * (N2 - 2);
  return 2L * (long)(N1 - 2) * (N2 - 2) + ((ymask & 1) ? 0 : ny) + ((ymask & 2) ? 0 : ny) +
         2L * (long)(N0 - 2) * (N1 - 2);
}

// Streaming accesses: values read or written once per launch, marked nontemporal so they do
// not displace reused lines (the stencil's A^n planes) from the caches.
template <typename T>
__device__ inline T ld_stream(gptr<const T> p, long i) {
  return __builtin_nontemporal_load(p + i);
}
template <typename T>
__device__ inline void st_stream(gptr<T> p, long i, T v) {
  __builtin_nontemporal_store(v, p + i);
}

// Component c of a FieldSet by selects: indexing the pointer array with a runtime value would
// place the struct in scratch memory.
__host__ __device__ inline void* component(const FieldSet& f, int c) {
  return c == 0 ? f.c[0] : c == 1 ? f.c[1] : c == 2 ? f.c[2] : f.c[3];
}

struct MurYFace {  // coefficients of the y faces (MurCoeffs::cB) and the faces to update
  double B[5];
  int mask;
};

struct EvalConst {  // -dt, twice the steps, and their correctly rounded reciprocals (div_rn)
  double mdt, dx2, dy2, dz2, rmdt, rdx2, rdy2, rdz2;
};

// Device copy of one static undulator with its per-undulator constants, computed once at upload
// in the push's operation order (Solver::undulatorField, src/solver.cpp:1798-1824): b0, the
// wavenumber ku, the polarisation cos/sin and the length L in lengths.
struct UndDev {
  double k, lu, rb, dist, theta;
  unsigned length;
  int type;                 // MITHRA_UNDULATOR_*
  int beam;                 // optical: index of its laser beam in Domain::beams
  double b0, ku, ct, st, L;
};

// A laser beam of an optical undulator or an external field (mithra_beam) with its transverse
// axis yv = direction x polarization, computed once at upload as the reference computes it for
// every particle (cross(direction_, polarization_), src/beam.cc:96, :121).
struct BeamDev {
  int beam_type, signal_type;
  int order[2];
  unsigned nR;
  double amplitude;
  double position[3], direction[3], polarization[3], yv[3];
  double radius[2], zR[2], l;
  double t0, s, f0, cep;
  double sigma_invg[2];
};

// Signal::self (src/classes.cpp:534-575) of a seed or a beam: the carrier at phase `phase` times
// the pulse envelope of the signal type.
template <typename S>
__device__ double signal(const S& s, double t, double phase) {
  const double t0 = s.t0, ss = s.s, f0 = s.f0, cep = s.cep;
  if (fabs(t - t0) > 10.0 * ss) return 0.0;
  const double c = cos(2 * kPI * f0 * (t - t0) + cep + phase);
  switch (s.signal_type) {
    case MITHRA_SIGNAL_NEUMANN:
      return -cos(2 * kPI * f0 * (t - t0) + cep + phase) * 2.7724 * (t - t0) / (ss * ss) *
             exp(-1.3863 * (t - t0) * (t - t0) / (ss * ss));
    case MITHRA_SIGNAL_GAUSSIAN:
      return c * exp(-1.3863 * pow((t - t0) / ss, 2));
    case MITHRA_SIGNAL_SECANT:
      return c / cosh((t - t0) / ss);
    case MITHRA_SIGNAL_FLATTOP:
      if (t - t0 <= -ss / 2.0) return c * exp(-pow((t - t0 + ss / 2.0) * f0 / s.nR, 2));
      if (t - t0 <= ss / 2.0) return c;
      return c * exp(-pow((t - t0 - ss / 2.0) * f0 / s.nR, 2));
    case MITHRA_SIGNAL_INVGAUSSIAN: {
      double env = pow((1.0 + pow((t - t0) / s.sigma_invg[0], 2)) * (1.0 + pow((t - t0) / s.sigma_invg[1], 2)), 0.25);
      if (t - t0 <= -ss / 2.0) return c * env * exp(-pow((t - t0 + ss / 2.0) * f0 / s.nR, 2));
      if (t - t0 <= ss / 2.0) return c * env;
      return c * env * exp(-pow((t - t0 - ss / 2.0) * f0 / s.nR, 2));
    }
  }
  return 0.0;
}

struct PushConst {
  // mesh
  double xmin, xmax, ymin, ymax, zmin, zp0, zp1, dx, dy, dz, Lz;
  double ix, iy, iz;  // RN(1/dx), RN(1/dy), RN(1/dz) for div_rn
  int N0, N1, N2;
  // frame and Boris constants
  double gamma, beta, c0, dt_shift, r1, r2, dtb;
  // time of the first sub-step and the bunch time step added after each sub-step
  double tb0, dtb_time;
  int n_sub, n_und, save_rnm;
  double dist0;  // undulator[0].dist_: entry test of the emission flag
};

struct DepositConst {
  double xmin, xmax, ymin, ymax, zmin, zp0, zp1, dx, dy, dz;
  double ix, iy, iz;        // RN(1/dx), RN(1/dy), RN(1/dz) for div_rn
  int N0, N1, N2, sc;
};

struct CellOrder {          // particle halves sorted by the (k, i) row of their cell
  unsigned* key;            // row k*N0 + i, or `rows` for a half that deposits nothing
  unsigned* idx;            // particle index
  unsigned* row;            // row r starts at row[r] in key/idx; rows + 1 entries
  unsigned* cell;           // fixed-order deposit: keys are cells (k N0 + i)(N1 - 1) + j, and
                            // cell c starts at cell[c]; cells + 1 entries (else unused)
};

struct Domain {
  mithra_grid g;
  mithra_coeffs co;
  int precision;            // MITHRA_PRECISION_*
  int ncomp;                // 3, or 4 with space charge
  size_t real_size;         // bytes per potential value
  long N1N0, nodes;
  int device;               // the GPU the domain's memory and stream belong to
  hipStream_t stream;
  void* pot_block;          // one allocation: 3 levels x ncomp x nodes
  FieldSet lvl[3];          // physical buffers
  int i_np1, i_n, i_nm1;    // which physical buffer is A^{n+1}, A^n, A^{n-1}
  float* eb;                // E/B, 6 floats per node (kEBPairs pairs)
  // particles
  long np;                  // particle count
  long cap;                 // allocated capacity
  double* part;             // PCOUNT arrays of cap doubles
  unsigned* pid;            // load-order index of each particle (the arrays are kept in row order)
  // cell-row orders of the particle halves (sort.hip); ord[0]: rnp halves of the next deposit,
  // ord[1]: rnm halves
  CellOrder ord[2];
  unsigned* key_alt;        // radix-sort double buffers (key_alt: the counting order's keys)
  unsigned* idx_alt;
  unsigned* row_cnt;        // per-row counts of the counting order (rows + 1), zero between orders
  unsigned* sort_block;     // one allocation holding all of the above
  void* sort_tmp;
  size_t sort_tmp_bytes;
  long sort_cap;            // capacity the order buffers were sized for
  unsigned sort_rows;
  int deterministic;        // fixed-order deposit (mithra_dev_set_deterministic)
  void* det_block;          // its cell starts (2 orders) and per-half weights (2 halves)
  size_t det_bytes;
  double* det_w[2];         // weights of the listed rnp / rnm halves, weight t of entry s at [t cap + s]
  int rm_order_valid;       // ord[1] lists the current rnm halves
  int order1_is_rnp;        // ord[1] lists the current rnp (true right after a deposit)
  double* part_alt;         // target of a reorder, swapped with part
  unsigned* pid_alt;
  int tile_interior;        // tile shape of the interior update (index into kTileShapes)
  int tile_eval;            // tile shape of the E/B evaluation
  int tile_fused;           // tile shape of the pipelined stencil + E/B kernel (kFusedShapes)
  // pipelined step: the A^{n+1} buffer holds the marched A^{n+1} (interior and source) awaiting
  // its seed injection and boundaries, instead of J; J itself was never stored
  int pipe;
  // lazy E/B evaluation: only the planes the bunch (z cells zrange[0..1], device) and the power
  // sampler (planes pk_lo .. pk_hi + 1) read
  int lazy_fields;
  int* zrange;              // device: bunch z-cell extent after the last push
  void* zpart;              // device: per-workgroup partial extents
  int zrange_valid;         // zrange describes the current particle positions
  int pk_lo, pk_hi;         // power planes' cells (pk_lo > pk_hi: none)
  int zero_pending;         // the A^{n+1} buffer is logically zero (current_reset), not yet written
  int yfaces_pending;       // Mur y faces the last interior launch updated (yface_mask bits)
  double und_dist0;         // undulator[0].dist of the uploaded table (the emission flag's entry test)
  int sort_interval;        // deposits between reorders of the particle arrays (0: never)
  long sort_count;          // deposits since the particles were set
  unsigned long long* counters;   // [0] transverse escapes, [1] z-boundary crossings
  // power sampling
  unsigned pN, pNz, pNl, pNf, prows;
  double pc, pgamma, pbeta, pc0;
  double* pz;               // device: plane positions (only planes inside the domain)
  int* pk;                  // device: plane cell index
  double* pdzr;             // device: plane fractional position
  double* dft;              // device: [Nl][Nf][4] = cos, sin  (ep = cos + i sin, em = cos - i sin)
  double* ring;             // device: [Nf][4][Nz][N1N0]
  double* partial;          // device: per-block partial sums
  double* history;          // device: [rows][N][Nl]
  int pblocks;
  int* plane_of;            // host: for each of N planes, its index among inside planes or -1
  // laser beams of the push (mithra_dev_set_beams): n_opt optical-undulator beams, then n_ext
  // external-field beams
  BeamDev* beams;
  int n_opt, n_ext;
  // super-gaussian TF/SF: per call its beam term, running vector and reset flag (field.hip)
  void* sg_buf;
  long sg_cap;
  // staging of the diagnostics gathers (diagnostics.hip), grown on demand
  void* diag;
  size_t diag_bytes;
  // power map of one plane (power visualization, diagnostics.hip): ring [Nf][4][N1N0] of the
  // lab-frame fields, DFT phases [Nf] (cos, sin), the map; vk < 0: none
  double* vring;
  double* vdft;
  double* vmap;
  unsigned vNf;
  int vk;
  double vdzr, vpc, vgamma, vbeta, vc0;
};

// Planes pb0..pb1 the lazy E/B evaluation adds to the bunch's: the power-sampling planes' cells
// and the power map's cell, each with the plane above (pb1 < 0: none).
inline void lazy_planes(const Domain* d, int* pb0, int* pb1) {
  int lo = d->pk_lo, hi = d->pk_hi;
  if (d->vk >= 0) {
    if (lo > hi) lo = hi = d->vk;
    lo = lo < d->vk ? lo : d->vk;
    hi = hi > d->vk ? hi : d->vk;
  }
  *pb0 = lo;
  *pb1 = lo <= hi ? hi + 1 : -1;
}

// Correctly rounded x / d given r = RN(1 / d): q = RN(x r) is within one ulp of x / d, and one
// residual step with exact FMA residual rounds it correctly (Markstein, IBM J. Res. Dev. 34,
// 1990, theorem 1).  Valid while the operands and quotient are normal numbers, which holds for
// the mesh coordinates and potentials divided here by fixed steps; an exact quotient (zero
// residual) is returned as q, so the sign of a zero quotient is the IEEE one.  Three
// instructions instead of the general division sequence, same result bit for bit.
__device__ inline double div_rn(double x, double d, double r) {
  const double q = x * r;
  const double e = fma(-q, d, x);
  return e == 0.0 ? q : fma(e, r, q);
}

// One instance of a batch as the batched kernels read it (device memory, one per instance).  The
// first block is advanced on the device after every step, so one captured step graph replays
// for the whole run; the rest describes the instance's domain and does not change.
struct BatchView {
  int active;               // step < n_steps
  int i_np1, i_n, i_nm1;    // potential levels, rotated like Domain's
  int parity;               // ord[parity]: rnp order of this step's deposit, ord[1-parity]: rnm order
  long step, n_steps;
  double time, time_bunch;
  // mesh, potentials, E/B
  int N0, N1, N2, ncomp, nsfd, trunc2;
  mithra_grid grid;         // the instance's mesh (TF/SF node positions)
  mithra_seed seed;         // its analytic seed, injected when `seeded`
  int seeded;
  FieldSet lvl[3];
  float* eb;
  int tilesJ, tilesI, chunks;  // interior / E-B tiles of the batch's tile shape
  int yfaces;               // Mur y faces the interior kernels update (yface_mask of the tile shape)
  double a[6], alpha, beta;
  MurCoeffs mur;
  EvalConst ev;
  double dt;
  // particles
  double* part;
  double* part_alt;         // spare arrays of the periodic reorder (swapped in by kb_advance)
  unsigned* pid;
  unsigned* pid_alt;
  int reorder;              // steps between reorders of the particle arrays (0: never)
  long cap, np;
  PushConst push;           // tb0 is taken from time_bunch each step
  const UndDev* und;
  unsigned long long* counters;
  int n_sub;
  double dtb;
  DepositConst dep;
  CellOrder ord[2];
  unsigned rows;            // cell rows (N2-1)*N0 of the deposition
  unsigned* cnt;            // rows+1 counters of the counting sort, zero between sorts
  unsigned* bsum;           // per-1024-row block sums of the scan
  int dTI, dtilesI;         // row-tiled deposit geometry
  // power sampling
  double* ring;
  const double* dft;
  const int* pk;
  const double* pdzr;
  double* partial;
  double* history;
  const int* plane_of;
  unsigned pNz, pN, pNl, pNf, prows;
  int pblocks;
  double pgamma, pbeta, pc0, pc;
  // bunching factor samples: every bevery steps into bout[step / bevery] (bslots of them)
  double bk;
  long bevery;
  double* bout;
  double* bpart;            // per particle-block partial sums, 3 doubles each
  int bslots;
};

inline double* particle_array(Domain* d, int which) { return d->part + (size_t)which * d->cap; }

// Launch helpers implemented in field.hip, particles.hip and power.hip.
// fuse_yfaces: no TF/SF write follows, so the LDS-staged variants may update the Mur y faces
// too; launch_boundaries then leaves them alone.
hipError_t launch_interior(Domain* d, bool fuse_yfaces);
hipError_t launch_tfsf(Domain* d, double time, const mithra_seed* seed);
hipError_t launch_boundaries(Domain* d);
hipError_t launch_eval_eb(Domain* d);
hipError_t launch_eval_eb_range(Domain* d, int klo, int khi);  // planes klo..khi (clipped to 1..N2-2)
hipError_t launch_eval_eb_step(Domain* d);  // all planes, or the lazy subset when enabled and known
// Pipelined step: E/B of step n and X, the next interior update without its source term, into
// the A^{n-1} buffer (no space charge).  Lazy E/B as launch_eval_eb_step.
hipError_t launch_stencil_eb(Domain* d);
constexpr int kFusedVariants = 4;
extern const int kFusedShapes[kFusedVariants][2];
hipError_t launch_zrange(Domain* d);
// Tile shapes {rows, planes} of the marched field kernels, selected by Domain::tile_*.
constexpr int kTileVariants = 10;  // 7..9: LDS-staged interior update
extern const int kTileShapes[kTileVariants][2];
hipError_t launch_fill_zero(Domain* d, void* ptr, size_t bytes);
hipError_t launch_bunch_begin(Domain* d);
hipError_t launch_push(Domain* d, const mithra_bunch_params* p, const UndDev* und, int n_und,
                       double time_bunch, double bunch_time_step, int n_sub, int save_rnm);
// Current deposition into the A^{n+1} buffer: accumulate = 0 replaces its contents with the
// deposit, 1 adds the deposit to them, kDepositSource (row-tiled path only) replaces each
// interior node's X with X + a4 J (the pipelined step) and leaves the boundary nodes alone.
constexpr int kDepositSource = 2;
hipError_t launch_deposit(Domain* d, int accumulate);
unsigned long long fixed_order_cells(const Domain* d);  // cells of the fixed-order deposit, 0: too many
bool deposit_tiled(const Domain* d);
hipError_t launch_bunching(Domain* d, double k, double* partial, int blocks);
hipError_t launch_power(Domain* d, unsigned n_time, unsigned row);
hipError_t launch_particle_ids(Domain* d);
void free_sort(Domain* d);
hipError_t ensure_orders(Domain* d, unsigned rows);
hipError_t sort_order(Domain* d, CellOrder& o, long n, unsigned rows, unsigned span);
hipError_t scan_row_counts(Domain* d, const unsigned* cnt, unsigned* row, unsigned rows);
hipError_t reorder_particles(Domain* d, CellOrder& o);

// Batched forms of the kernels (persistent workgroups walking a table of (instance, local tile)
// entries; an entry of an instance that has finished is skipped).
struct BatchTiles {
  const int2* t;
  int n;
};
hipError_t launch_batch_interior(BatchView* V, BatchTiles T, int variant, int precision, int nsfd, hipStream_t s);
hipError_t launch_batch_mur(BatchView* V, BatchTiles faces, BatchTiles edges, BatchTiles corners, int precision,
                            hipStream_t s);
hipError_t launch_batch_tfsf(BatchView* V, const BatchTiles* axes, int precision, hipStream_t s);  // axes[3]
hipError_t launch_batch_eval(BatchView* V, BatchTiles T, int variant, int precision, hipStream_t s);
hipError_t launch_batch_push(BatchView* V, BatchTiles T, hipStream_t s);
hipError_t launch_batch_bunching(BatchView* V, BatchTiles particles, BatchTiles inst, hipStream_t s);
hipError_t launch_batch_sort(BatchView* V, BatchTiles particles, BatchTiles rowblocks, BatchTiles inst, int use_rm,
                             int slot_sel, hipStream_t s);
hipError_t launch_batch_deposit(BatchView* V, BatchTiles T, int precision, size_t lds, hipStream_t s);
hipError_t launch_batch_reorder(BatchView* V, BatchTiles particles, hipStream_t s);
hipError_t launch_batch_power(BatchView* V, BatchTiles nodes, BatchTiles reduce, hipStream_t s);
int batch_grid(int ntiles, int per_cu);  // persistent grid size for a table of ntiles entries

int deposit_tile_rows(const Domain* d);      // rows per tile of the row-tiled deposit (0: atomic path)
long deposit_blocks(int N2, int tilesI, int KC);  // workgroups of the row-tiled deposit (row tiles x z chunks)
int deposit_chunk();                         // z planes per chunk of the batched deposit
size_t deposit_lds_bytes(int ncomp, int TI, int N1);  // its LDS per workgroup
DepositConst make_deposit_const(const Domain* d);

}  // namespace mithra

// The C-ABI handle of one domain (include/mithra_gpu.h).
struct mithra_dev : mithra::Domain {
  mithra::UndDev* und_dev = nullptr;                // device copy of the undulator table
  double und_c0 = 0.0;                      // the speed of light its b0 were computed with
  int und_cap = 0;
  std::vector<mithra_undulator> und_host;   // what und_dev holds
  hipEvent_t ev[MITHRA_STEP_PHASES + 1] = {};  // phase timers of mithra_dev_step_timed
  bool ev_ready = false;
};

namespace mithra {
int upload_undulators(mithra_dev* d, const mithra_undulator* und, int n, double c0);  // capi.hip
int bunch_records(const mithra_bunch* b, const double** part, long* cap, long* n);   // bunch.hip
int settle(mithra_dev* d);  // writes the zeros current_reset left pending
}  // namespace mithra
