// C-ABI of libmithra_gpu.so (include/mithra_gpu.h): device memory, transfers between the
// reference's host layouts and the device layout, and the launch sequence of each call.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "common.h"

using namespace mithra;


namespace {

// Deposits between two cell sorts of the particles.  In the boosted frame a particle moves
// well under a cell per step, so the order decays slowly; a sort costs a few deposits' worth
// of bandwidth on the production mesh.
constexpr int kDefaultSortInterval = 16;

inline int hip_status(hipError_t e) { return e == hipSuccess ? MITHRA_OK : MITHRA_E_HIP - (int)e; }

#define MITHRA_TRY(expr)                      \
  do {                                        \
    hipError_t _e = (expr);                   \
    if (_e != hipSuccess) return hip_status(_e); \
  } while (0)

// Bunching factor: k_bunching runs on a fixed grid, each workgroup leaving (sum q cos kz,
// sum q sin kz, sum q) over its particles; the host adds the partials in workgroup order.
constexpr int kBunchingBlocks = 512;

double bunching_from_partials(const double* h, int blocks) {
  double c = 0.0, s = 0.0, q = 0.0;
  for (int b = 0; b < blocks; ++b) {
    c += h[3 * b];
    s += h[3 * b + 1];
    q += h[3 * b + 2];
  }
  return q != 0.0 ? std::sqrt(c * c + s * s) / std::fabs(q) : 0.0;
}

// Device memory of mithra_dev_bunching_noalloc, which may not allocate: a file-scope array of
// kBunchingScratch doubles that every GPU the library runs on holds until the process ends.
// Its name is a host-side stand-in; the array's address on each GPU comes from
// hipGetSymbolAddress with that GPU current, once, and only that address is handed to kernels.
// Every call on one GPU shares that GPU's array, so calls that overlap on one GPU race.
constexpr long kBunchingScratch = 3L * kBunchingBlocks;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kBunchingScratch];
std::mutex scratch_mutex;
double* scratch_ptr[kMaxDevices] = {};

int bunching_scratch(int device, double** out) {
  std::lock_guard<std::mutex> lock(scratch_mutex);
  if (!scratch_ptr[device]) {
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e == hipSuccess && cur != device) e = hipSetDevice(device);
    void* p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur >= 0 && cur != device) {
      const hipError_t r = hipSetDevice(cur);
      if (e == hipSuccess) e = r;
    }
    if (e != hipSuccess) return MITHRA_E_HIP - (int)e;
    scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = scratch_ptr[device];
  return MITHRA_OK;
}

bool valid_grid(const mithra_grid* g) {
  return g && g->N0 >= 3 && g->N1 >= 3 && g->N2 >= 3 && (g->solver == MITHRA_SOLVER_FD || g->solver == MITHRA_SOLVER_NSFD) &&
         (g->truncation_order == 1 || g->truncation_order == 2) && g->dx > 0 && g->dy > 0 && g->dz > 0 && g->dt > 0 &&
         (long)g->N0 * g->N1 * g->N2 < (1L << 40);
}

}  // namespace

int mithra::upload_undulators(mithra_dev* d, const mithra_undulator* und, int n, double c0) {
  if (n < 0 || (n > 0 && !und)) return MITHRA_E_ARG;
  int optical = 0;
  for (int u = 0; u < n; ++u) {
    if (und[u].type == MITHRA_UNDULATOR_OPTICAL) ++optical;
    else if (und[u].type != MITHRA_UNDULATOR_STATIC) return MITHRA_E_ARG;
  }
  if (optical > d->n_opt) return MITHRA_E_STATE;  // an optical undulator without its beam
  if ((int)d->und_host.size() == n && d->und_c0 == c0 &&
      (n == 0 || std::memcmp(d->und_host.data(), und, sizeof(*und) * n) == 0))
    return MITHRA_OK;
  if (n > d->und_cap) {
    if (d->und_dev) MITHRA_TRY(hipFree(d->und_dev));
    d->und_dev = nullptr;
    MITHRA_TRY(hipMalloc(&d->und_dev, sizeof(UndDev) * n));
    d->und_cap = n;
  }
  std::vector<UndDev> t(n);
  int beam = 0;
  for (int u = 0; u < n; ++u) {  // the push's expressions, src/solver.cpp:1805-1812
    const mithra_undulator& U = und[u];
    UndDev& v = t[u];
    v.k = U.k; v.lu = U.lu; v.rb = U.rb; v.dist = U.dist; v.theta = U.theta; v.length = U.length;
    v.type = U.type;
    v.beam = U.type == MITHRA_UNDULATOR_OPTICAL ? beam++ : -1;
    v.b0 = (U.lu != 0.0) ? kEM * c0 * 2 * kPI / U.lu * U.k / kEC : 0.0;
    v.ku = (U.lu != 0.0) ? 2 * kPI / U.lu : 0.0;
    v.ct = std::cos(U.theta);
    v.st = std::sin(U.theta);
    v.L = U.length * U.lu;
  }
  if (n) {
    MITHRA_TRY(hipStreamSynchronize(d->stream));
    MITHRA_TRY(hipMemcpy(d->und_dev, t.data(), sizeof(UndDev) * n, hipMemcpyHostToDevice));
  }
  d->und_host.assign(und, und + n);
  d->und_c0 = c0;
  d->und_dist0 = n > 0 ? und[0].dist : 0.0;
  return MITHRA_OK;
}

namespace {

// current_reset only marks the A^{n+1} buffer as zero: the next deposit writes every node of
// it anyway.  Anything else that reads or moves the potentials writes the zeros first.
}  // namespace

int mithra::settle(mithra_dev* d) {
  if (d->pipe) return MITHRA_E_STATE;  // mid-way through a pipelined run: J is not held
  if (!d->zero_pending) return MITHRA_OK;
  d->zero_pending = 0;
  for (int c = 0; c < d->ncomp; ++c)
    MITHRA_TRY(launch_fill_zero(d, d->lvl[d->i_np1].c[c], (size_t)d->nodes * d->real_size));
  return MITHRA_OK;
}

namespace {

// One iteration of Solver::solve()'s loop (src/solver.cpp:1303-1399) on the device: the interior
// update (skipped when the previous pipelined step already marched A^{n+1}), seed injection, Mur
// boundaries, E/B (fused with the next update's stencil when `more` keeps the run pipelined),
// bunch sub-steps, power sample, field shift and deposition (of J, or, pipelined, of X + a4 J).
// ms (optional) receives the phase durations of mithra_dev_step_timed.
int run_step(mithra_dev* d, double time, const mithra_seed* seed, const mithra_bunch_params* p,
             const mithra_undulator* und, int n_und, double time_bunch, double bunch_time_step, int n_sub,
             unsigned n_time, int power_row, int more, float* ms) {
  if (!p || n_sub < 0) return MITHRA_E_ARG;
  const bool seeded = seed && std::fabs(seed->amplitude) > 1.0e-50;
  if (seeded && (seed->seed_type < MITHRA_SEED_PLANEWAVE || seed->seed_type > MITHRA_SEED_STANDING_SUPERGAUSSIANBEAM))
    return MITHRA_E_ARG;
  if (power_row >= 0 && (!d->prows || (unsigned)power_row >= d->prows)) return MITHRA_E_STATE;
  int rc = upload_undulators(d, und, n_und, p->c0);
  if (rc) return rc;
  const bool pipe_in = d->pipe != 0;
  if (!pipe_in && (rc = settle(d))) return rc;
  // FP64 only: in FP32-bulk the fused kernel saves no time overall (the separate Mur y faces and
  // the deposit's X reads cost what it saves), and X would be rounded to FP32 before a4 J
  const bool pipe_out = more && d->precision == MITHRA_PRECISION_FP64 && d->np > 0 &&
                        deposit_tiled(d) && !d->deterministic;
  if (ms && !d->ev_ready) {
    for (auto& e : d->ev) MITHRA_TRY(hipEventCreate(&e));
    d->ev_ready = true;
  }
  hipStream_t s = d->stream;
  auto mark = [&](int i) { return ms ? hipEventRecord(d->ev[i], s) : hipSuccess; };
  const bool tfsf = seeded && d->g.N0 > 4 && d->g.N1 > 4 && d->g.N2 > 4;
  MITHRA_TRY(mark(0));
  if (!pipe_in) MITHRA_TRY(launch_interior(d, !tfsf));
  MITHRA_TRY(mark(1));
  if (tfsf) MITHRA_TRY(launch_tfsf(d, time, seed));
  MITHRA_TRY(mark(2));
  MITHRA_TRY(launch_boundaries(d));
  MITHRA_TRY(mark(3));
  MITHRA_TRY(pipe_out ? launch_stencil_eb(d) : launch_eval_eb_step(d));
  MITHRA_TRY(mark(4));
  MITHRA_TRY(launch_push(d, p, d->und_dev, n_und, time_bunch, bunch_time_step, n_sub, 1));
  MITHRA_TRY(mark(5));
  if (power_row >= 0) MITHRA_TRY(launch_power(d, n_time, (unsigned)power_row));
  MITHRA_TRY(mark(6));
  const int t = d->i_nm1;  // field shift; the free level then holds X when pipelined
  d->i_nm1 = d->i_n;
  d->i_n = d->i_np1;
  d->i_np1 = t;
  d->zero_pending = 0;
  MITHRA_TRY(mark(7));
  const hipError_t e = launch_deposit(d, pipe_out ? kDepositSource : 0);
  if (e == hipErrorOutOfMemory) return MITHRA_E_NOMEM;
  MITHRA_TRY(e);
  d->pipe = pipe_out ? 1 : 0;
  MITHRA_TRY(mark(8));
  if (ms) {
    MITHRA_TRY(hipEventSynchronize(d->ev[8]));
    for (int i = 0; i < MITHRA_STEP_PHASES; ++i) MITHRA_TRY(hipEventElapsedTime(&ms[i], d->ev[i], d->ev[i + 1]));
  }
  return MITHRA_OK;
}

int level_of(const mithra_dev* d, int level) {
  return level == 0 ? d->i_n : level == 1 ? d->i_nm1 : level == 2 ? d->i_np1 : -1;
}

void free_power(mithra_dev* d) {
  hipFree(d->pz); hipFree(d->pk); hipFree(d->pdzr); hipFree(d->dft); hipFree(d->ring);
  hipFree(d->partial); hipFree(d->history);
  d->pz = nullptr; d->pk = nullptr; d->pdzr = nullptr; d->dft = nullptr; d->ring = nullptr;
  d->partial = nullptr; d->history = nullptr;
  delete[] d->plane_of;
  d->plane_of = nullptr;
  d->prows = 0;
}

}  // namespace

extern "C" {

const char* mithra_strerror(int status) {
  switch (status) {
    case MITHRA_OK: return "success";
    case MITHRA_E_ARG: return "invalid argument or shape";
    case MITHRA_E_UNSUPPORTED: return "job feature not supported by the GPU library";
    case MITHRA_E_STATE: return "call out of order";
    case MITHRA_E_NOMEM: return "device allocation failed";
  }
  if (status <= MITHRA_E_HIP) return hipGetErrorString((hipError_t)(MITHRA_E_HIP - status));
  return "unknown status";
}

int mithra_dev_create(const mithra_grid* grid, const mithra_coeffs* coeffs, int precision, mithra_dev** out) {
  if (!out || !coeffs || !valid_grid(grid)) return MITHRA_E_ARG;
  if (precision != MITHRA_PRECISION_FP64 && precision != MITHRA_PRECISION_FP32_BULK) return MITHRA_E_ARG;
  *out = nullptr;
  mithra_dev* d = new (std::nothrow) mithra_dev();
  if (!d) return MITHRA_E_NOMEM;
  std::memset(static_cast<Domain*>(d), 0, sizeof(Domain));
  d->g = *grid;
  d->co = *coeffs;
  d->precision = precision;
  d->ncomp = grid->space_charge ? 4 : 3;
  d->real_size = precision == MITHRA_PRECISION_FP64 ? sizeof(double) : sizeof(float);
  d->N1N0 = (long)grid->N1 * grid->N0;
  d->nodes = d->N1N0 * grid->N2;
  d->i_np1 = 0; d->i_n = 1; d->i_nm1 = 2;
  d->sort_interval = kDefaultSortInterval;
  d->pk_lo = 1;
  d->pk_hi = 0;
  d->vk = -1;
  hipError_t e = hipGetDevice(&d->device);
  if (e != hipSuccess) { delete d; return hip_status(e); }
  e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete d; return hip_status(e); }
  const size_t per = (size_t)d->nodes * d->real_size;
  // 256-byte aligned component arrays inside one allocation
  const size_t stride = (per + 255) / 256 * 256;
  if (hipMalloc(&d->pot_block, stride * 3 * d->ncomp) != hipSuccess ||
      hipMalloc(&d->eb, sizeof(float) * 2 * kEBPairs * (size_t)d->nodes) != hipSuccess ||
      hipMalloc(&d->counters, 2 * sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&d->zrange, 2 * sizeof(int)) != hipSuccess ||
      hipMalloc(&d->zpart, 1024 * 2 * sizeof(int)) != hipSuccess) {
    mithra_dev_destroy(d);
    return MITHRA_E_NOMEM;
  }
  for (int l = 0; l < 3; ++l)
    for (int c = 0; c < 4; ++c)
      d->lvl[l].c[c] = c < d->ncomp ? (char*)d->pot_block + stride * (l * d->ncomp + c) : nullptr;
  int rc = hip_status(launch_fill_zero(d, d->pot_block, stride * 3 * d->ncomp));
  if (!rc) rc = hip_status(launch_fill_zero(d, d->eb, sizeof(float) * 2 * kEBPairs * (size_t)d->nodes));
  if (!rc) rc = hip_status(launch_fill_zero(d, d->counters, 2 * sizeof(unsigned long long)));
  if (!rc) rc = hip_status(hipStreamSynchronize(d->stream));
  if (rc) { mithra_dev_destroy(d); return rc; }
  *out = d;
  return MITHRA_OK;
}

int mithra_dev_destroy(mithra_dev* d) {
  if (!d) return MITHRA_E_ARG;
  if (d->stream) hipStreamSynchronize(d->stream);
  hipFree(d->pot_block);
  hipFree(d->eb);
  hipFree(d->counters);
  hipFree(d->zrange);
  hipFree(d->zpart);
  hipFree(d->part);
  hipFree(d->pid);
  free_sort(d);
  hipFree(d->und_dev);
  hipFree(d->beams);
  hipFree(d->det_block);
  free_power(d);
  hipFree(d->diag);
  hipFree(d->sg_buf);
  hipFree(d->vring);
  hipFree(d->vdft);
  hipFree(d->vmap);
  if (d->ev_ready)
    for (auto& e : d->ev) hipEventDestroy(e);
  if (d->stream) hipStreamDestroy(d->stream);
  delete d;
  return MITHRA_OK;
}

int mithra_dev_set_beams(mithra_dev* d, const mithra_beam* optical, int n_optical, const mithra_beam* external,
                         int n_external) {
  if (!d || n_optical < 0 || n_external < 0 || (n_optical > 0 && !optical) || (n_external > 0 && !external))
    return MITHRA_E_ARG;
  std::vector<BeamDev> t;
  for (int q = 0; q < n_optical + n_external; ++q) {
    const mithra_beam& b = q < n_optical ? optical[q] : external[q - n_optical];
    if (b.beam_type < MITHRA_SEED_PLANEWAVE || b.beam_type > MITHRA_SEED_STANDING_SUPERGAUSSIANBEAM ||
        b.signal_type < MITHRA_SIGNAL_NEUMANN || b.signal_type > MITHRA_SIGNAL_INVGAUSSIAN || b.order[0] < 0 ||
        b.order[1] < 0)
      return MITHRA_E_ARG;
    BeamDev v;
    v.beam_type = b.beam_type; v.signal_type = b.signal_type;
    v.order[0] = b.order[0]; v.order[1] = b.order[1]; v.nR = b.nR;
    v.amplitude = b.amplitude;
    for (int c = 0; c < 3; ++c) {
      v.position[c] = b.position[c]; v.direction[c] = b.direction[c]; v.polarization[c] = b.polarization[c];
    }
    // cross(direction_, polarization_) (src/fieldvector.h:197-203)
    v.yv[0] = b.direction[1] * b.polarization[2] - b.direction[2] * b.polarization[1];
    v.yv[1] = b.direction[2] * b.polarization[0] - b.direction[0] * b.polarization[2];
    v.yv[2] = b.direction[0] * b.polarization[1] - b.direction[1] * b.polarization[0];
    v.radius[0] = b.radius[0]; v.radius[1] = b.radius[1];
    v.zR[0] = b.zR[0]; v.zR[1] = b.zR[1]; v.l = b.l;
    v.t0 = b.t0; v.s = b.s; v.f0 = b.f0; v.cep = b.cep;
    v.sigma_invg[0] = b.sigma_invg[0]; v.sigma_invg[1] = b.sigma_invg[1];
    t.push_back(v);
  }
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  MITHRA_TRY(hipFree(d->beams));
  d->beams = nullptr;
  d->n_opt = d->n_ext = 0;
  d->und_host.clear();  // the undulator table is checked against the new beams at the next push
  if (!t.empty()) {
    if (hipMalloc(&d->beams, sizeof(BeamDev) * t.size()) != hipSuccess) return MITHRA_E_NOMEM;
    MITHRA_TRY(hipMemcpy(d->beams, t.data(), sizeof(BeamDev) * t.size(), hipMemcpyHostToDevice));
  }
  d->n_opt = n_optical;
  d->n_ext = n_external;
  return MITHRA_OK;
}

int mithra_dev_synchronize(mithra_dev* d) {
  if (!d) return MITHRA_E_ARG;
  return hip_status(hipStreamSynchronize(d->stream));
}

int mithra_dev_set_potentials(mithra_dev* d, int level, const double* a, const double* phi) {
  if (!d) return MITHRA_E_ARG;
  if (int rc = settle(d)) return rc;
  const int li = level_of(d, level);
  if (li < 0) return MITHRA_E_ARG;
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  const size_t n = (size_t)d->nodes;
  std::vector<double> dbl;
  std::vector<float> flt;
  auto put = [&](int c, const double* src, size_t stride) -> hipError_t {
    if (d->precision == MITHRA_PRECISION_FP64) {
      dbl.resize(n);
      for (size_t m = 0; m < n; ++m) dbl[m] = src[stride * m];
      return hipMemcpy(d->lvl[li].c[c], dbl.data(), n * sizeof(double), hipMemcpyHostToDevice);
    }
    flt.resize(n);
    for (size_t m = 0; m < n; ++m) flt[m] = (float)src[stride * m];
    return hipMemcpy(d->lvl[li].c[c], flt.data(), n * sizeof(float), hipMemcpyHostToDevice);
  };
  if (a)
    for (int c = 0; c < 3; ++c) MITHRA_TRY(put(c, a + c, 3));
  if (phi && d->ncomp == 4) MITHRA_TRY(put(3, phi, 1));
  return MITHRA_OK;
}

int mithra_dev_get_potentials(mithra_dev* d, int level, double* a, double* phi) {
  if (!d) return MITHRA_E_ARG;
  if (int rc = settle(d)) return rc;
  const int li = level_of(d, level);
  if (li < 0) return MITHRA_E_ARG;
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  const size_t n = (size_t)d->nodes;
  std::vector<double> dbl(n);
  std::vector<float> flt;
  auto get = [&](int c, double* dst, size_t stride) -> hipError_t {
    hipError_t e;
    if (d->precision == MITHRA_PRECISION_FP64) {
      e = hipMemcpy(dbl.data(), d->lvl[li].c[c], n * sizeof(double), hipMemcpyDeviceToHost);
    } else {
      flt.resize(n);
      e = hipMemcpy(flt.data(), d->lvl[li].c[c], n * sizeof(float), hipMemcpyDeviceToHost);
      for (size_t m = 0; m < n; ++m) dbl[m] = flt[m];
    }
    if (e == hipSuccess)
      for (size_t m = 0; m < n; ++m) dst[stride * m] = dbl[m];
    return e;
  };
  if (a)
    for (int c = 0; c < 3; ++c) MITHRA_TRY(get(c, a + c, 3));
  if (phi && d->ncomp == 4) MITHRA_TRY(get(3, phi, 1));
  return MITHRA_OK;
}

int mithra_dev_get_fields(mithra_dev* d, float* e, float* b) {
  if (!d) return MITHRA_E_ARG;
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  std::vector<float> h(2 * kEBPairs * (size_t)d->nodes);  // per node: Ex Ey Ez Bx By Bz
  MITHRA_TRY(hipMemcpy(h.data(), d->eb, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
  for (size_t m = 0; m < (size_t)d->nodes; ++m) {
    if (e) { e[3 * m] = h[6 * m]; e[3 * m + 1] = h[6 * m + 1]; e[3 * m + 2] = h[6 * m + 2]; }
    if (b) { b[3 * m] = h[6 * m + 3]; b[3 * m + 1] = h[6 * m + 4]; b[3 * m + 2] = h[6 * m + 5]; }
  }
  return MITHRA_OK;
}

}  // extern "C"

namespace {

// Room for n particles (keeping nothing: the caller fills every array), and the state of a
// domain whose particles were just replaced.
int prepare_particles(mithra_dev* d, long n) {
  if (n >= (1L << 31)) return MITHRA_E_ARG;
  if (n > d->cap) {
    hipFree(d->part);
    hipFree(d->pid);
    free_sort(d);
    d->part = nullptr;
    d->pid = nullptr;
    d->cap = 0;
    // until the new arrays exist the domain holds no particles: a failed allocation below
    // must leave no launch reading the freed ones
    d->np = 0;
    d->rm_order_valid = 0;
    d->order1_is_rnp = 0;
    d->zrange_valid = 0;
    const long cap = (n + 63) / 64 * 64;
    if (hipMalloc(&d->part, sizeof(double) * PCOUNT * (size_t)cap) != hipSuccess ||
        hipMalloc(&d->pid, sizeof(unsigned) * (size_t)cap) != hipSuccess) {
      hipFree(d->part);
      d->part = nullptr;
      d->pid = nullptr;
      return MITHRA_E_NOMEM;
    }
    d->cap = cap;
  }
  d->np = n;
  d->sort_count = 0;
  d->rm_order_valid = 0;  // no cell order describes the new particles yet
  d->order1_is_rnp = 0;
  d->zrange_valid = 0;    // nor a bunch extent
  return n ? hip_status(launch_particle_ids(d)) : MITHRA_OK;
}

}  // namespace

extern "C" {

int mithra_dev_set_particles(mithra_dev* d, const double* q, long n) {
  if (!d || n < 0 || (n > 0 && !q)) return MITHRA_E_ARG;
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  if (int rc = prepare_particles(d, n)) return rc;
  if (n == 0) return MITHRA_OK;
  // Charge record q, rnp[3], rnm[3], gb[3], e  ->  x y z xm ym zm gx gy gz q e
  static const int from[PCOUNT] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 0, 10};
  std::vector<double> col(n);
  for (int a = 0; a < PCOUNT; ++a) {
    for (long p = 0; p < n; ++p) col[p] = q[11 * p + from[a]];
    MITHRA_TRY(hipMemcpy(particle_array(d, a), col.data(), sizeof(double) * n, hipMemcpyHostToDevice));
  }
  return MITHRA_OK;
}

int mithra_dev_take_bunch(mithra_dev* d, const mithra_bunch* b) {
  const double* src = nullptr;
  long cap = 0, n = 0;
  if (!d || bunch_records(b, &src, &cap, &n)) return MITHRA_E_ARG;
  MITHRA_TRY(hipDeviceSynchronize());  // the bunch's kernels run on the null stream
  if (int rc = prepare_particles(d, n)) return rc;
  for (int a = 0; a < PCOUNT && n > 0; ++a)
    MITHRA_TRY(hipMemcpyAsync(particle_array(d, a), src + (size_t)a * cap, sizeof(double) * n,
                              hipMemcpyDeviceToDevice, d->stream));
  return hip_status(hipStreamSynchronize(d->stream));
}

int mithra_dev_get_particles(mithra_dev* d, double* q, long n) {
  if (!d || n != d->np || (n > 0 && !q)) return MITHRA_E_ARG;
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  static const int from[PCOUNT] = {1, 2, 3, 4, 5, 6, 7, 8, 9, 0, 10};
  if (n == 0) return MITHRA_OK;
  std::vector<unsigned> id(n);
  MITHRA_TRY(hipMemcpy(id.data(), d->pid, sizeof(unsigned) * n, hipMemcpyDeviceToHost));
  std::vector<double> col(n);
  for (int a = 0; a < PCOUNT; ++a) {
    MITHRA_TRY(hipMemcpy(col.data(), particle_array(d, a), sizeof(double) * n, hipMemcpyDeviceToHost));
    for (long p = 0; p < n; ++p) q[11 * (long)id[p] + from[a]] = col[p];  // back to load order
  }
  return MITHRA_OK;
}

long mithra_dev_particle_count(const mithra_dev* d) { return d ? d->np : -1; }

int mithra_dev_field_update(mithra_dev* d, double time, const mithra_seed* seed) {
  if (!d) return MITHRA_E_ARG;
  const bool seeded = seed && std::fabs(seed->amplitude) > 1.0e-50;
  if (seeded && (seed->seed_type < MITHRA_SEED_PLANEWAVE || seed->seed_type > MITHRA_SEED_STANDING_SUPERGAUSSIANBEAM))
    return MITHRA_E_ARG;
  if (int rc = settle(d)) return rc;
  const bool tfsf = seeded && d->g.N0 > 4 && d->g.N1 > 4 && d->g.N2 > 4;
  MITHRA_TRY(launch_interior(d, !tfsf));
  if (tfsf) MITHRA_TRY(launch_tfsf(d, time, seed));
  MITHRA_TRY(launch_boundaries(d));
  MITHRA_TRY(launch_eval_eb_step(d));
  return MITHRA_OK;
}

int mithra_dev_field_evaluate(mithra_dev* d) {
  if (!d) return MITHRA_E_ARG;
  if (int rc = settle(d)) return rc;
  return hip_status(launch_eval_eb(d));
}

int mithra_dev_bunch_begin(mithra_dev* d) {
  if (!d) return MITHRA_E_ARG;
  return hip_status(launch_bunch_begin(d));
}

int mithra_dev_bunch_update(mithra_dev* d, const mithra_bunch_params* p, const mithra_undulator* und, int n_und,
                            double time_bunch) {
  if (!d || !p) return MITHRA_E_ARG;
  int rc = upload_undulators(d, und, n_und, p->c0);
  if (rc) return rc;
  return hip_status(launch_push(d, p, d->und_dev, n_und, time_bunch, 0.0, 1, 0));
}

int mithra_dev_bunch_advance(mithra_dev* d, const mithra_bunch_params* p, const mithra_undulator* und, int n_und,
                             double time_bunch, double bunch_time_step, int n_sub) {
  if (!d || !p || n_sub < 0) return MITHRA_E_ARG;
  int rc = upload_undulators(d, und, n_und, p->c0);
  if (rc) return rc;
  return hip_status(launch_push(d, p, d->und_dev, n_und, time_bunch, bunch_time_step, n_sub, 1));
}

int mithra_dev_bunch_counters(mithra_dev* d, long* transverse_out, long* z_crossings) {
  if (!d) return MITHRA_E_ARG;
  unsigned long long h[2];
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  MITHRA_TRY(hipMemcpy(h, d->counters, sizeof h, hipMemcpyDeviceToHost));
  if (transverse_out) *transverse_out = (long)h[0];
  if (z_crossings) *z_crossings = (long)h[1];
  MITHRA_TRY(launch_fill_zero(d, d->counters, sizeof h));
  return MITHRA_OK;
}

int mithra_dev_power_setup(mithra_dev* d, const mithra_power_params* p, unsigned rows) {
  if (!d || !p || p->Nf == 0 || rows == 0 || (p->N > 0 && !p->z) || (p->Nl > 0 && !p->w)) return MITHRA_E_ARG;
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  free_power(d);
  const mithra_grid& g = d->g;
  d->pN = p->N; d->pNl = p->Nl; d->pNf = p->Nf; d->pc = p->pc;
  d->pgamma = p->gamma; d->pbeta = p->beta; d->pc0 = p->c0;
  d->plane_of = new int[p->N > 0 ? p->N : 1];
  std::vector<int> hk;
  std::vector<double> hdzr, hz;
  unsigned nz = 0;
  for (unsigned k = 0; k < p->N; ++k) {
    if (p->z[k] < g.zp1 && p->z[k] >= g.zp0) {  // src/radiation.cpp:157
      double c;
      double dzr = std::modf((p->z[k] - g.zmin) / g.dz, &c);
      if ((int)c < 0 || (int)c + 1 >= g.N2) return MITHRA_E_ARG;
      hk.push_back((int)c);
      hdzr.push_back(dzr);
      hz.push_back(p->z[k]);
      d->plane_of[k] = (int)nz++;
    } else {
      d->plane_of[k] = -1;
    }
  }
  d->pNz = nz;
  d->pk_lo = 1;
  d->pk_hi = 0;
  for (int c : hk) {
    if (d->pk_lo > d->pk_hi) d->pk_lo = d->pk_hi = c;
    d->pk_lo = std::min(d->pk_lo, c);
    d->pk_hi = std::max(d->pk_hi, c);
  }
  // DFT phases exactly as the reference builds ep/em (src/radiation.cpp:110-117)
  std::vector<double> dft((size_t)p->Nl * p->Nf * 2);
  for (unsigned i = 0; i < p->Nl; ++i)
    for (unsigned j = 0; j < p->Nf; ++j) {
      dft[((size_t)i * p->Nf + j) * 2] = std::cos(p->w[i] * j * g.dt);
      dft[((size_t)i * p->Nf + j) * 2 + 1] = std::sin(p->w[i] * j * g.dt);
    }
  const long ni = (long)(g.N0 - 4) * (g.N1 - 4);
  d->pblocks = ni > 0 ? (int)((ni + 255) / 256) : 1;
  std::vector<int> kz(hk);
  for (unsigned k = 0; k < p->N; ++k) kz.push_back(d->plane_of[k]);
  const size_t ring = (size_t)p->Nf * 4 * (nz ? nz : 1) * d->N1N0;
  if (hipMalloc(&d->pk, sizeof(int) * (kz.size() ? kz.size() : 1)) != hipSuccess ||
      hipMalloc(&d->pdzr, sizeof(double) * (nz ? nz : 1)) != hipSuccess ||
      hipMalloc(&d->pz, sizeof(double) * (nz ? nz : 1)) != hipSuccess ||
      hipMalloc(&d->dft, sizeof(double) * (dft.size() ? dft.size() : 1)) != hipSuccess ||
      hipMalloc(&d->ring, sizeof(double) * ring) != hipSuccess ||
      hipMalloc(&d->partial, sizeof(double) * (size_t)(nz ? nz : 1) * d->pblocks * (p->Nl ? p->Nl : 1)) != hipSuccess ||
      hipMalloc(&d->history, sizeof(double) * (size_t)rows * (p->N ? p->N : 1) * (p->Nl ? p->Nl : 1)) != hipSuccess) {
    free_power(d);
    return MITHRA_E_NOMEM;
  }
  if (!kz.empty()) MITHRA_TRY(hipMemcpy(d->pk, kz.data(), sizeof(int) * kz.size(), hipMemcpyHostToDevice));
  if (nz) {
    MITHRA_TRY(hipMemcpy(d->pdzr, hdzr.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
    MITHRA_TRY(hipMemcpy(d->pz, hz.data(), sizeof(double) * nz, hipMemcpyHostToDevice));
  }
  if (!dft.empty()) MITHRA_TRY(hipMemcpy(d->dft, dft.data(), sizeof(double) * dft.size(), hipMemcpyHostToDevice));
  MITHRA_TRY(launch_fill_zero(d, d->ring, sizeof(double) * ring));
  MITHRA_TRY(launch_fill_zero(d, d->history, sizeof(double) * (size_t)rows * (p->N ? p->N : 1) * (p->Nl ? p->Nl : 1)));
  d->prows = rows;
  return hip_status(hipStreamSynchronize(d->stream));
}

int mithra_dev_power_sample(mithra_dev* d, unsigned n_time, unsigned row) {
  if (!d) return MITHRA_E_ARG;
  if (!d->prows || row >= d->prows) return MITHRA_E_STATE;
  return hip_status(launch_power(d, n_time, row));
}

int mithra_dev_power_fetch(mithra_dev* d, unsigned first, unsigned count, double* out) {
  if (!d || (count && !out)) return MITHRA_E_ARG;
  if (!d->prows || first + count > d->prows) return MITHRA_E_STATE;
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  const size_t w = (size_t)d->pN * d->pNl;
  if (count && w)
    MITHRA_TRY(hipMemcpy(out, d->history + first * w, sizeof(double) * w * count, hipMemcpyDeviceToHost));
  return MITHRA_OK;
}

int mithra_dev_field_shift(mithra_dev* d) {
  if (!d) return MITHRA_E_ARG;
  int rc = settle(d);
  if (rc) return rc;
  const int t = d->i_nm1;
  d->i_nm1 = d->i_n;
  d->i_n = d->i_np1;
  d->i_np1 = t;
  return MITHRA_OK;
}

int mithra_dev_current_reset(mithra_dev* d) {
  if (!d) return MITHRA_E_ARG;
  if (d->pipe) return MITHRA_E_STATE;
  d->zero_pending = 1;
  return MITHRA_OK;
}

int mithra_dev_current_update(mithra_dev* d) {
  if (!d) return MITHRA_E_ARG;
  if (d->pipe) return MITHRA_E_STATE;
  const int accumulate = d->zero_pending ? 0 : 1;
  d->zero_pending = 0;
  const hipError_t e = launch_deposit(d, accumulate);
  if (e == hipErrorOutOfMemory) return MITHRA_E_NOMEM;
  return hip_status(e);
}

int mithra_dev_set_lazy_fields(mithra_dev* d, int lazy) {
  if (!d || (lazy != 0 && lazy != 1)) return MITHRA_E_ARG;
  d->lazy_fields = lazy;
  return MITHRA_OK;
}

int mithra_dev_set_reorder_interval(mithra_dev* d, int interval) {
  if (!d || interval < 0) return MITHRA_E_ARG;
  d->sort_interval = interval;
  d->sort_count = 0;
  return MITHRA_OK;
}

int mithra_dev_set_deterministic(mithra_dev* d, int on) {
  if (!d || (on != 0 && on != 1)) return MITHRA_E_ARG;
  if (d->pipe) return MITHRA_E_STATE;  // a pipelined run holds X + a4 J, not J
  if (on && fixed_order_cells(d) == 0) return MITHRA_E_UNSUPPORTED;
  if (on == d->deterministic) return MITHRA_OK;
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  d->deterministic = on;
  // the orders are keyed by cell in one mode and by row in the other: list them afresh
  d->rm_order_valid = 0;
  d->order1_is_rnp = 0;
  if (!on) {
    (void)hipFree(d->det_block);
    d->det_block = nullptr;
    d->det_bytes = 0;
    d->ord[0].cell = d->ord[1].cell = nullptr;
  }
  return MITHRA_OK;
}

int mithra_dev_step(mithra_dev* d, double time, const mithra_seed* seed, const mithra_bunch_params* p,
                    const mithra_undulator* und, int n_und, double time_bunch, double bunch_time_step, int n_sub,
                    unsigned n_time, int power_row) {
  if (!d) return MITHRA_E_ARG;
  return run_step(d, time, seed, p, und, n_und, time_bunch, bunch_time_step, n_sub, n_time, power_row, 0, nullptr);
}

int mithra_dev_step_pipelined(mithra_dev* d, double time, const mithra_seed* seed, const mithra_bunch_params* p,
                              const mithra_undulator* und, int n_und, double time_bunch, double bunch_time_step,
                              int n_sub, unsigned n_time, int power_row, int more, float* ms) {
  if (!d || (more != 0 && more != 1)) return MITHRA_E_ARG;
  return run_step(d, time, seed, p, und, n_und, time_bunch, bunch_time_step, n_sub, n_time, power_row, more, ms);
}

int mithra_dev_pipeline_pending(const mithra_dev* d) { return d ? d->pipe : MITHRA_E_ARG; }

int mithra_tile_shape(int variant, int* rows, int* planes) {
  if (variant < 0 || variant >= kTileVariants) return MITHRA_E_ARG;
  if (rows) *rows = kTileShapes[variant][0];
  if (planes) *planes = kTileShapes[variant][1];
  return MITHRA_OK;
}

int mithra_dev_set_tiles(mithra_dev* d, int interior, int eval) {
  if (!d || interior < 0 || interior >= kTileVariants || eval < 0 || eval >= kTileVariants) return MITHRA_E_ARG;
  d->tile_interior = interior;
  d->tile_eval = eval;
  return MITHRA_OK;
}

int mithra_dev_autotune(mithra_dev* d, int reps, int* chosen, float* ms) {
  if (!d || reps < 1) return MITHRA_E_ARG;
  if (int rc = settle(d)) return rc;
  hipStream_t s = d->stream;
  hipEvent_t e0, e1;
  MITHRA_TRY(hipEventCreate(&e0));
  MITHRA_TRY(hipEventCreate(&e1));
  // The interior update overwrites the J it reads: time it on a saved copy of that level.
  const size_t level_bytes = (size_t)d->ncomp * ((d->nodes * d->real_size + 255) / 256 * 256);
  void* saved = nullptr;
  int rc = MITHRA_OK;
  if (hipMalloc(&saved, level_bytes) != hipSuccess) rc = MITHRA_E_NOMEM;
  void* np1 = d->lvl[d->i_np1].c[0];  // the components of one level are contiguous
  if (!rc) rc = hip_status(hipMemcpyAsync(saved, np1, level_bytes, hipMemcpyDeviceToDevice, s));
  const int keep[2] = {d->tile_interior, d->tile_eval};
  int best[2] = {keep[0], keep[1]};
  for (int kernel = 0; kernel < 2 && !rc; ++kernel) {
    float best_ms = 0.0f;
    for (int v = 0; v < kTileVariants && !rc; ++v) {
      if (kernel == 0) d->tile_interior = v;
      else d->tile_eval = v;
      float total = 0.0f;
      for (int r = 0; r <= reps && !rc; ++r) {  // r = 0 warms up
        if (kernel == 0) rc = hip_status(hipMemcpyAsync(np1, saved, level_bytes, hipMemcpyDeviceToDevice, s));
        if (!rc) rc = hip_status(hipEventRecord(e0, s));
        // the interior is timed with the Mur boundaries: the LDS forms take over the y faces
        if (!rc) rc = hip_status(kernel == 0 ? launch_interior(d, true) : launch_eval_eb(d));
        if (!rc && kernel == 0) rc = hip_status(launch_boundaries(d));
        if (!rc) rc = hip_status(hipEventRecord(e1, s));
        if (!rc) rc = hip_status(hipEventSynchronize(e1));
        float t = 0.0f;
        if (!rc) rc = hip_status(hipEventElapsedTime(&t, e0, e1));
        if (r > 0) total += t;
      }
      if (!rc && (v == 0 || total < best_ms)) {
        best_ms = total;
        best[kernel] = v;
      }
      if (ms) ms[kernel * kTileVariants + v] = total / reps;
    }
  }
  if (saved) {
    hipError_t e = hipMemcpyAsync(np1, saved, level_bytes, hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (!rc && e != hipSuccess) rc = hip_status(e);
    (void)hipFree(saved);
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  d->tile_interior = rc ? keep[0] : best[0];
  d->tile_eval = rc ? keep[1] : best[1];
  if (!rc && chosen) {
    chosen[0] = best[0];
    chosen[1] = best[1];
  }
  return rc;
}

int mithra_dev_bunching(mithra_dev* d, double k, double* b_out) {
  if (!d || !b_out) return MITHRA_E_ARG;
  const int blocks = kBunchingBlocks;
  double* partial = nullptr;
  if (hipMalloc(&partial, sizeof(double) * 3 * blocks) != hipSuccess) return MITHRA_E_NOMEM;
  hipError_t e = launch_fill_zero(d, partial, sizeof(double) * 3 * blocks);
  if (e == hipSuccess) e = launch_bunching(d, k, partial, blocks);
  std::vector<double> h(3 * blocks);
  if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
  if (e == hipSuccess) e = hipMemcpy(h.data(), partial, sizeof(double) * h.size(), hipMemcpyDeviceToHost);
  hipFree(partial);
  if (e != hipSuccess) return hip_status(e);
  *b_out = bunching_from_partials(h.data(), blocks);
  return MITHRA_OK;
}

int mithra_dev_bunching_noalloc(mithra_dev* d, double k, double* b_out) {
  if (!d || !b_out) return MITHRA_E_ARG;
  // the partial sums of k_bunching: one per workgroup of the fixed grid, whatever the particle
  // count, so every domain the library accepts fits
  constexpr long need = 3L * kBunchingBlocks;
  static_assert(need <= kBunchingScratch, "bunching partials exceed the static scratch");
  if (d->device < 0 || d->device >= kMaxDevices) return MITHRA_E_ARG;
  double* partial = nullptr;
  if (int rc = bunching_scratch(d->device, &partial)) return rc;
  MITHRA_TRY(launch_fill_zero(d, partial, sizeof(double) * need));
  MITHRA_TRY(launch_bunching(d, k, partial, kBunchingBlocks));
  double h[need];
  MITHRA_TRY(hipMemcpyAsync(h, partial, sizeof h, hipMemcpyDeviceToHost, d->stream));
  MITHRA_TRY(hipStreamSynchronize(d->stream));
  *b_out = bunching_from_partials(h, kBunchingBlocks);
  return MITHRA_OK;
}

int mithra_dev_step_timed(mithra_dev* d, double time, const mithra_seed* seed, const mithra_bunch_params* p,
                          const mithra_undulator* und, int n_und, double time_bunch, double bunch_time_step, int n_sub,
                          unsigned n_time, int power_row, float* ms) {
  if (!d || !ms) return MITHRA_E_ARG;
  return run_step(d, time, seed, p, und, n_und, time_bunch, bunch_time_step, n_sub, n_time, power_row, 0, ms);
}

int mithra_dev_set_fused_tile(mithra_dev* d, int variant) {
  if (!d || variant < 0 || variant >= kFusedVariants) return MITHRA_E_ARG;
  d->tile_fused = variant;
  return MITHRA_OK;
}

int mithra_fused_tile_shape(int variant, int* rows, int* planes) {
  if (variant < 0 || variant >= kFusedVariants) return MITHRA_E_ARG;
  if (rows) *rows = kFusedShapes[variant][0];
  if (planes) *planes = kFusedShapes[variant][1];
  return MITHRA_OK;
}

int mithra_dev_autotune_fused(mithra_dev* d, int reps, int* chosen, float* ms) {
  if (!d || reps < 1) return MITHRA_E_ARG;
  if (int rc = settle(d)) return rc;
  hipStream_t s = d->stream;
  // the kernel writes its X into the A^{n-1} level and E/B into the field array: time it there
  // and restore both after
  const size_t level_bytes = (size_t)d->ncomp * ((d->nodes * d->real_size + 255) / 256 * 256);
  const size_t eb_bytes = sizeof(float) * 2 * kEBPairs * (size_t)d->nodes;
  char* saved = nullptr;
  if (hipMalloc(&saved, level_bytes + eb_bytes) != hipSuccess) return MITHRA_E_NOMEM;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = hip_status(hipEventCreate(&e0));
  if (!rc) rc = hip_status(hipEventCreate(&e1));
  void* nm1 = d->lvl[d->i_nm1].c[0];  // the components of one level are contiguous
  if (!rc) rc = hip_status(hipMemcpyAsync(saved, nm1, level_bytes, hipMemcpyDeviceToDevice, s));
  if (!rc) rc = hip_status(hipMemcpyAsync(saved + level_bytes, d->eb, eb_bytes, hipMemcpyDeviceToDevice, s));
  const int keep = d->tile_fused;
  int best = keep;
  float best_ms = 0.0f;
  for (int v = 0; v < kFusedVariants && !rc; ++v) {
    d->tile_fused = v;
    float total = 0.0f;
    for (int r = 0; r <= reps && !rc; ++r) {  // r = 0 warms up
      rc = hip_status(hipEventRecord(e0, s));
      if (!rc) rc = hip_status(launch_stencil_eb(d));
      if (!rc) rc = hip_status(hipEventRecord(e1, s));
      if (!rc) rc = hip_status(hipEventSynchronize(e1));
      float t = 0.0f;
      if (!rc) rc = hip_status(hipEventElapsedTime(&t, e0, e1));
      if (r > 0) total += t;
    }
    if (!rc && (v == 0 || total < best_ms)) {
      best_ms = total;
      best = v;
    }
    if (ms) ms[v] = total / reps;
  }
  hipError_t e = hipMemcpyAsync(nm1, saved, level_bytes, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d->eb, saved + level_bytes, eb_bytes, hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (!rc && e != hipSuccess) rc = hip_status(e);
  (void)hipFree(saved);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  d->tile_fused = rc ? keep : best;
  if (!rc && chosen) *chosen = best;
  return rc;
}

}  // extern "C"

namespace {
template <typename T>
__global__ void k_copy(const T* __restrict__ a, T* __restrict__ b, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) b[t] = a[t];
}

template <typename T>
int copy_rate(unsigned long bytes, int reps, double* gbps) {
  const size_t n = bytes / sizeof(T);
  T *a = nullptr, *b = nullptr;
  hipEvent_t e0, e1;
  if (hipMalloc(&a, n * sizeof(T)) != hipSuccess || hipMalloc(&b, n * sizeof(T)) != hipSuccess) {
    (void)hipFree(a);
    return MITHRA_E_NOMEM;
  }
  MITHRA_TRY(hipMemset(a, 0, n * sizeof(T)));
  MITHRA_TRY(hipEventCreate(&e0));
  MITHRA_TRY(hipEventCreate(&e1));
  const unsigned blocks = 256 * 32;
  hipLaunchKernelGGL(k_copy<T>, dim3(blocks), dim3(256), 0, 0, a, b, n);  // warm-up
  MITHRA_TRY(hipEventRecord(e0, 0));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy<T>, dim3(blocks), dim3(256), 0, 0, a, b, n);
  MITHRA_TRY(hipEventRecord(e1, 0));
  MITHRA_TRY(hipEventSynchronize(e1));
  float ms = 0;
  MITHRA_TRY(hipEventElapsedTime(&ms, e0, e1));
  *gbps = 2.0 * n * sizeof(T) * reps / (ms * 1e-3) / 1e9;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  return MITHRA_OK;
}
// The interior stencil's stream pattern without its arithmetic: three arrays read, one written,
// 16 bytes per thread per array, the read-once streams and the written one nontemporal.
typedef double d2v __attribute__((ext_vector_type(2)));

__global__ __launch_bounds__(512) void k_stream3r1w(const d2v* __restrict__ a, const d2v* __restrict__ b,
                                                    const d2v* __restrict__ c, d2v* __restrict__ d, size_t n) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
    __builtin_nontemporal_store(a[t] + __builtin_nontemporal_load(b + t) + __builtin_nontemporal_load(c + t), d + t);
}

int stream_rate(unsigned long bytes, int reps, double* gbps) {
  const size_t n = bytes / sizeof(d2v);
  d2v* buf[4] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MITHRA_OK;
  for (int i = 0; i < 4 && !rc; ++i)
    if (hipMalloc(&buf[i], n * sizeof(d2v)) != hipSuccess) rc = MITHRA_E_NOMEM;
  for (int i = 0; i < 3 && !rc; ++i) rc = hip_status(hipMemset(buf[i], 0, n * sizeof(d2v)));
  if (!rc) rc = hip_status(hipEventCreate(&e0));
  if (!rc) rc = hip_status(hipEventCreate(&e1));
  const unsigned blocks = 65536;
  float ms = 0.0f;
  if (!rc) {
    hipLaunchKernelGGL(k_stream3r1w, dim3(blocks), dim3(512), 0, 0, buf[0], buf[1], buf[2], buf[3], n);  // warm-up
    rc = hip_status(hipEventRecord(e0, 0));
  }
  for (int r = 0; r < reps && !rc; ++r)
    hipLaunchKernelGGL(k_stream3r1w, dim3(blocks), dim3(512), 0, 0, buf[0], buf[1], buf[2], buf[3], n);
  if (!rc) rc = hip_status(hipGetLastError());
  if (!rc) rc = hip_status(hipEventRecord(e1, 0));
  if (!rc) rc = hip_status(hipEventSynchronize(e1));
  if (!rc) rc = hip_status(hipEventElapsedTime(&ms, e0, e1));
  if (!rc) *gbps = 4.0 * n * sizeof(d2v) * reps / (ms * 1e-3) / 1e9;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  for (d2v* p : buf) (void)hipFree(p);
  return rc;
}
}  // namespace

extern "C" {

int mithra_stream_bandwidth(unsigned long bytes, int reps, double* gbps) {
  if (!gbps || reps < 1 || bytes < 4096) return MITHRA_E_ARG;
  return stream_rate(bytes, reps, gbps);
}

int mithra_copy_bandwidth(unsigned long bytes, int reps, double* gbps) {
  return mithra_copy_bandwidth_elem(bytes, reps, 32, gbps);
}

int mithra_copy_bandwidth_elem(unsigned long bytes, int reps, int elem_bytes, double* gbps) {
  if (!gbps || reps < 1 || bytes < 4096) return MITHRA_E_ARG;
  switch (elem_bytes) {
    case 8: return copy_rate<double>(bytes, reps, gbps);
    case 16: return copy_rate<double2>(bytes, reps, gbps);
    case 32: return copy_rate<double4>(bytes, reps, gbps);
  }
  return MITHRA_E_ARG;
}

}  // extern "C"

static_assert(MITHRA_FUSED_TILE_VARIANTS == mithra::kFusedVariants, "header and library disagree on the fused shapes");
