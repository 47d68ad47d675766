// Batched parameter scans: many independent FEL domains of uneven size marched together.
//
// Each instance is an ordinary mithra_dev (created, filled and read back through the single-
// domain calls).  A batch adds, per instance, a device descriptor (BatchView) holding its
// pointers, launch constants and the state a step advances (potential levels, particle-order
// parity, time, time_bunch, step), and per kernel a table of (instance, local tile) entries over
// the unfinished instances.  One step is a fixed sequence of batched launches, each a set of
// persistent workgroups walking its table; since every launch reads its per-step values from
// the descriptors, one HIP graph, captured once when the scan starts, replays every step to the
// end of the scan; a finished instance's table entries are skipped on the device by its
// descriptor's `active` flag, so nothing is re-captured or re-instantiated mid-scan (an earlier
// form rebuilt the tables and re-instantiated the graph each time an instance finished, some 60
// times per 64-instance scan).  The deposition's particle orders come from a
// counting sort by cell row (graph-safe: kernels only).  Every instance computes with the same
// device functions as its single-domain run, so its P(z) and bunching equal that run's to the
// rounding of the deposit's summation order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "common.h"

using namespace mithra;

namespace {

inline int status(hipError_t e) { return e == hipSuccess ? MITHRA_OK : MITHRA_E_HIP - (int)e; }

#define BATCH_TRY(expr)                       \
  do {                                        \
    hipError_t _e = (expr);                   \
    if (_e != hipSuccess) return status(_e);  \
  } while (0)

__global__ void kb_shift(BatchView* V, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !V[i].active) return;
  BatchView& v = V[i];
  const int t = v.i_nm1;  // FdTd::fieldShift
  v.i_nm1 = v.i_n;
  v.i_n = v.i_np1;
  v.i_np1 = t;
}

__global__ void kb_advance(BatchView* V, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !V[i].active) return;
  BatchView& v = V[i];
  if (v.reorder > 0 && v.step % v.reorder == 0) {  // kb_reorder gathered into the spare arrays
    double* t = v.part;
    v.part = v.part_alt;
    v.part_alt = t;
    unsigned* ti = v.pid;
    v.pid = v.pid_alt;
    v.pid_alt = ti;
  }
  v.parity ^= 1;        // this step's rnp order is the next step's rnm order
  v.time += v.dt;       // Solver::solve: time_ += dt, timeBunch_ += dtb per sub-step
  for (int s = 0; s < v.n_sub; ++s) v.time_bunch += v.dtb;
  v.step += 1;
  v.active = v.step < v.n_steps;
}

}  // namespace

namespace mithra {

int batch_grid(int ntiles, int per_cu) {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) !=
                                                hipSuccess || n <= 0)
      n = 256;
    cus = n;
  }
  return std::max(1, std::min(ntiles, cus * per_cu));
}

}  // namespace mithra

struct mithra_batch {
  std::vector<mithra_dev*> dev;
  int n = 0, precision = 0;
  hipStream_t stream = nullptr;
  BatchView* V = nullptr;           // device descriptors
  std::vector<BatchView> hv;        // host copies
  std::vector<int> ready;           // instance described by mithra_batch_instance
  std::vector<void*> owned;         // batch-only device buffers
  int variant = 0;                  // tile shape of the field kernels (kTileShapes; 7..9 LDS-staged)
  // tables
  int2* tiles = nullptr;
  size_t tiles_cap = 0;
  BatchTiles t_int[2], t_eval, t_faces, t_edges, t_corners, t_part, t_rows, t_inst, t_dep, t_pnodes, t_pred;
  BatchTiles t_tfsf[3];             // TF/SF lines of the seeded instances, per axis
  size_t dep_lds = 0;
  bool prepared = false;
  std::vector<long> step;           // host mirror of each instance's step
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
};

namespace {

void free_batch_buffers(mithra_batch* b) {
  for (void* p : b->owned) (void)hipFree(p);
  b->owned.clear();
  (void)hipFree(b->tiles);
  b->tiles = nullptr;
  b->tiles_cap = 0;
  (void)hipFree(b->V);
  b->V = nullptr;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
}

int alloc_owned(mithra_batch* b, size_t bytes, void** out) {
  *out = nullptr;
  if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) return MITHRA_E_NOMEM;
  b->owned.push_back(*out);
  return MITHRA_OK;
}

// Builds every table over the instances still running; the largest instances come first so
// the persistent workgroups end together.  The tables are built once per scan (and per tile
// shape while autotuning): an instance that finishes keeps its entries, which the kernels then
// skip by its descriptor's `active` flag.
int build_tables(mithra_batch* b) {
  std::vector<int> order;
  for (int i = 0; i < b->n; ++i)
    if (b->step[i] < b->hv[i].n_steps) order.push_back(i);
  std::sort(order.begin(), order.end(), [&](int x, int y) {
    const BatchView &a = b->hv[x], &c = b->hv[y];
    return (double)a.N0 * a.N1 * a.N2 + 40.0 * a.np > (double)c.N0 * c.N1 * c.N2 + 40.0 * c.np;
  });
  std::vector<int2> all;
  auto table = [&](BatchTiles& T, auto each) {
    const size_t start = all.size();
    for (int i : order) each(i, b->hv[i]);
    T.t = reinterpret_cast<const int2*>((size_t)start);  // offset, fixed up below
    T.n = (int)(all.size() - start);
  };
  auto add = [&](int i, long count) {
    for (long l = 0; l < count; ++l) all.push_back(make_int2(i, (int)l));
  };
  for (int s = 0; s < 2; ++s)
    table(b->t_int[s], [&](int i, const BatchView& v) {
      if (v.nsfd == s) add(i, (long)v.tilesJ * v.tilesI * v.chunks * v.ncomp);
    });
  table(b->t_eval, [&](int i, const BatchView& v) { add(i, (long)v.tilesJ * v.tilesI * v.chunks); });
  table(b->t_faces, [&](int i, const BatchView& v) {
    add(i, v.ncomp * ((mur_face_count(v.N0, v.N1, v.N2, v.yfaces) + 255) / 256));
  });
  table(b->t_edges, [&](int i, const BatchView& v) {
    const long ne = 4L * ((v.N2 - 2) + (v.N0 - 2) + (v.N1 - 2));
    if (v.trunc2) add(i, v.ncomp * ((ne + 255) / 256));
  });
  table(b->t_corners, [&](int i, const BatchView& v) {
    if (v.trunc2) add(i, v.ncomp);
  });
  table(b->t_part, [&](int i, const BatchView& v) { add(i, (v.np + 255) / 256); });
  table(b->t_rows, [&](int i, const BatchView& v) { add(i, ((long)v.rows + 1 + 1023) / 1024); });
  table(b->t_inst, [&](int i, const BatchView&) { add(i, 1); });
  table(b->t_dep, [&](int i, const BatchView& v) { add(i, deposit_blocks(v.N2, v.dtilesI, deposit_chunk())); });
  table(b->t_pnodes, [&](int i, const BatchView& v) {
    if (v.pNz > 0 && v.N0 > 4 && v.N1 > 4) add(i, (long)v.pNz * v.pblocks);
  });
  table(b->t_pred, [&](int i, const BatchView& v) {
    if (v.prows > 0) add(i, ((long)v.pN * v.pNl + 63) / 64);
  });
  for (int axis = 0; axis < 3; ++axis)
    table(b->t_tfsf[axis], [&](int i, const BatchView& v) {
      if (!v.seeded) return;
      const long lines = axis == 0 ? (long)(v.N1 - 4) * (v.N2 - 4)
                                   : axis == 1 ? (long)(v.N0 - 4) * (v.N2 - 4) : (long)(v.N0 - 4) * (v.N1 - 4);
      add(i, (lines + 63) / 64);  // 256 work items (four injections of 64 lines) per entry
    });
  if (all.size() > b->tiles_cap) {
    (void)hipFree(b->tiles);
    b->tiles = nullptr;
    b->tiles_cap = 0;
    if (hipMalloc(&b->tiles, sizeof(int2) * all.size()) != hipSuccess) return MITHRA_E_NOMEM;
    b->tiles_cap = all.size();
  }
  if (!all.empty()) {  // `all` is a local: wait for the copy
    BATCH_TRY(hipMemcpyAsync(b->tiles, all.data(), sizeof(int2) * all.size(), hipMemcpyHostToDevice, b->stream));
    BATCH_TRY(hipStreamSynchronize(b->stream));
  }
  BatchTiles* ts[] = {&b->t_int[0], &b->t_int[1], &b->t_eval, &b->t_faces,    &b->t_edges,   &b->t_corners,
                      &b->t_part,   &b->t_rows,   &b->t_inst, &b->t_dep,      &b->t_pnodes,  &b->t_pred,
                      &b->t_tfsf[0], &b->t_tfsf[1], &b->t_tfsf[2]};
  for (BatchTiles* T : ts) T->t = b->tiles + (size_t)T->t;
  return MITHRA_OK;
}

int enqueue_step(mithra_batch* b) {
  hipStream_t s = b->stream;
  BatchView* V = b->V;
  const int p = b->precision;
  BATCH_TRY(launch_batch_interior(V, b->t_int[1], b->variant, p, 1, s));
  BATCH_TRY(launch_batch_interior(V, b->t_int[0], b->variant, p, 0, s));
  BATCH_TRY(launch_batch_tfsf(V, b->t_tfsf, p, s));
  BATCH_TRY(launch_batch_mur(V, b->t_faces, b->t_edges, b->t_corners, p, s));
  BATCH_TRY(launch_batch_eval(V, b->t_eval, b->variant, p, s));
  BATCH_TRY(launch_batch_push(V, b->t_part, s));
  BATCH_TRY(launch_batch_bunching(V, b->t_part, b->t_inst, s));
  BATCH_TRY(launch_batch_power(V, b->t_pnodes, b->t_pred, s));
  hipLaunchKernelGGL(kb_shift, dim3((b->n + 63) / 64), dim3(64), 0, s, V, b->n);
  BATCH_TRY(hipGetLastError());
  BATCH_TRY(launch_batch_sort(V, b->t_part, b->t_rows, b->t_inst, 0, 0, s));
  BATCH_TRY(launch_batch_deposit(V, b->t_dep, p, b->dep_lds, s));
  BATCH_TRY(launch_batch_reorder(V, b->t_part, s));
  hipLaunchKernelGGL(kb_advance, dim3((b->n + 63) / 64), dim3(64), 0, s, V, b->n);
  BATCH_TRY(hipGetLastError());
  return MITHRA_OK;
}

int capture(mithra_batch* b) {
  if (b->exec) {
    BATCH_TRY(hipGraphExecDestroy(b->exec));
    b->exec = nullptr;
  }
  if (!b->use_graph) return MITHRA_OK;
  hipGraph_t g = nullptr;
  BATCH_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  const int rc = enqueue_step(b);
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  BATCH_TRY(e);
  const hipError_t ei = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  BATCH_TRY(ei);
  return MITHRA_OK;
}

int upload_views(mithra_batch* b) {  // the host copies may change right after: wait for the copy
  BATCH_TRY(hipMemcpyAsync(b->V, b->hv.data(), sizeof(BatchView) * b->n, hipMemcpyHostToDevice, b->stream));
  BATCH_TRY(hipStreamSynchronize(b->stream));
  return MITHRA_OK;
}

void set_tile_geometry(BatchView& v, int variant) {
  const int TI = kTileShapes[variant][0], K = kTileShapes[variant][1];
  v.tilesJ = (v.N1 + 63) / 64;
  v.tilesI = (v.N0 - 2 + TI - 1) / TI;
  v.chunks = (v.N2 - 2 + K - 1) / K;
  // without a TF/SF write between them, the interior kernels take the y faces
  v.yfaces = v.seeded ? 0 : yface_mask(variant, v.N1);
}

int prepare(mithra_batch* b) {
  b->dep_lds = 0;
  for (int i = 0; i < b->n; ++i) {
    if (!b->ready[i]) return MITHRA_E_STATE;
    set_tile_geometry(b->hv[i], b->variant);
    const BatchView& v = b->hv[i];
    b->dep_lds = std::max(b->dep_lds, deposit_lds_bytes(v.ncomp, v.dTI, v.N1));
  }
  if (int rc = upload_views(b)) return rc;
  if (int rc = build_tables(b)) return rc;
  // The rnm halves of the first deposit: the first push sets rnm to the present rnp, so the
  // present rnp is sorted once, into the slot the first step reads as its rnm order.
  BATCH_TRY(launch_batch_sort(b->V, b->t_part, b->t_rows, b->t_inst, 0, 1, b->stream));
  if (int rc = capture(b)) return rc;
  b->prepared = true;
  return MITHRA_OK;
}

// Copies each instance's advanced state back into its mithra_dev, so the single-domain calls
// (power fetch, particles, potentials, further steps) continue from where the batch stopped.
int write_back(mithra_batch* b) {
  BATCH_TRY(hipStreamSynchronize(b->stream));
  BATCH_TRY(hipMemcpy(b->hv.data(), b->V, sizeof(BatchView) * b->n, hipMemcpyDeviceToHost));
  for (int i = 0; i < b->n; ++i) {
    mithra_dev* d = b->dev[i];
    const BatchView& v = b->hv[i];
    d->i_np1 = v.i_np1;
    d->i_n = v.i_n;
    d->i_nm1 = v.i_nm1;
    d->part = v.part;
    d->part_alt = v.part_alt;
    d->pid = v.pid;
    d->pid_alt = v.pid_alt;
    d->ord[0] = v.ord[v.parity];
    d->ord[1] = v.ord[1 - v.parity];  // the last deposit's rnp order
    d->order1_is_rnp = v.step > 0 ? 1 : d->order1_is_rnp;
    d->rm_order_valid = 0;
    d->zero_pending = 0;
    d->zrange_valid = 0;
  }
  return MITHRA_OK;
}

}  // namespace

extern "C" {

int mithra_batch_create(mithra_dev* const* devs, int n, mithra_batch** out) {
  if (!devs || n < 1 || !out) return MITHRA_E_ARG;
  *out = nullptr;
  for (int i = 0; i < n; ++i) {
    const mithra_dev* d = devs[i];
    if (!d) return MITHRA_E_ARG;
    if (d->precision != devs[0]->precision) return MITHRA_E_ARG;
    if (!deposit_tiled(d) || d->g.N0 < 5 || d->g.N1 < 5 || d->g.N2 < 5 || d->np >= (1L << 31))
      return MITHRA_E_UNSUPPORTED;
  }
  mithra_batch* b = new (std::nothrow) mithra_batch();
  if (!b) return MITHRA_E_NOMEM;
  b->dev.assign(devs, devs + n);
  b->n = n;
  b->precision = devs[0]->precision;
  b->hv.assign(n, BatchView());
  std::memset(b->hv.data(), 0, sizeof(BatchView) * n);
  b->ready.assign(n, 0);
  b->step.assign(n, 0);
  int rc = status(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  if (!rc && hipMalloc(&b->V, sizeof(BatchView) * n) != hipSuccess) rc = MITHRA_E_NOMEM;
  if (rc) {
    mithra_batch_destroy(b);
    return rc;
  }
  *out = b;
  return MITHRA_OK;
}

int mithra_batch_seed(mithra_batch* b, int i, const mithra_seed* seed) {
  if (!b || i < 0 || i >= b->n || !seed) return MITHRA_E_ARG;
  if (b->prepared || !b->ready[i]) return MITHRA_E_STATE;
  if (seed->seed_type < MITHRA_SEED_PLANEWAVE || seed->seed_type > MITHRA_SEED_STANDING_SUPERGAUSSIANBEAM ||
      seed->signal_type < MITHRA_SIGNAL_NEUMANN || seed->signal_type > MITHRA_SIGNAL_INVGAUSSIAN)
    return MITHRA_E_ARG;
  // the super-gaussian seed's running vector is a single-domain path
  if (std::fabs(seed->amplitude) > 1.0e-50 && seed->seed_type == MITHRA_SEED_SUPERGAUSSIANBEAM)
    return MITHRA_E_UNSUPPORTED;
  BatchView& v = b->hv[i];
  v.seed = *seed;
  // as the single-domain field update decides (src/fdtd.cpp:307); the standing types inject nothing
  v.seeded = std::fabs(seed->amplitude) > 1.0e-50 && v.N0 > 4 && v.N1 > 4 && v.N2 > 4 &&
             seed->seed_type < MITHRA_SEED_STANDING_PLANEWAVE;
  return MITHRA_OK;
}

int mithra_batch_destroy(mithra_batch* b) {
  if (!b) return MITHRA_E_ARG;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  free_batch_buffers(b);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return MITHRA_OK;
}

int mithra_batch_instance(mithra_batch* b, int i, double time, double time_bunch, double bunch_time_step, int n_sub,
                          long n_steps, const mithra_bunch_params* p, const mithra_undulator* und, int n_und,
                          double bunching_k, long bunching_every, int bunching_slots) {
  if (!b || i < 0 || i >= b->n || !p || n_sub < 1 || n_steps < 0 || bunching_slots < 0) return MITHRA_E_ARG;
  if (b->prepared) return MITHRA_E_STATE;
  mithra_dev* d = b->dev[i];
  // The batched sort and deposit walk every instance's particle orders, which exist only for
  // a domain holding particles; an instance without a bunch radiates nothing either.
  if (d->np <= 0) return MITHRA_E_UNSUPPORTED;
  // the batched push adds static undulators only; the batched deposit is the row-tiled one
  if (d->n_opt + d->n_ext > 0 || d->deterministic) return MITHRA_E_UNSUPPORTED;
  BATCH_TRY(hipStreamSynchronize(d->stream));
  if (int rc = settle(d)) return rc;
  if (int rc = upload_undulators(d, und, n_und, p->c0)) return rc;
  const mithra_grid& g = d->g;
  const unsigned rows = (unsigned)((g.N2 - 1) * (long)g.N0);
  if (hipError_t e = ensure_orders(d, rows); e != hipSuccess) return e == hipErrorOutOfMemory ? MITHRA_E_NOMEM : status(e);
  BATCH_TRY(hipStreamSynchronize(d->stream));
  BatchView& v = b->hv[i];
  std::memset(&v, 0, sizeof v);
  v.active = n_steps > 0;
  v.i_np1 = d->i_np1;
  v.i_n = d->i_n;
  v.i_nm1 = d->i_nm1;
  v.parity = 0;
  v.step = 0;
  v.n_steps = n_steps;
  v.time = time;
  v.time_bunch = time_bunch;
  v.N0 = g.N0;
  v.N1 = g.N1;
  v.N2 = g.N2;
  v.ncomp = d->ncomp;
  v.grid = g;
  v.seeded = 0;  // mithra_batch_seed sets the seed
  v.nsfd = g.solver == MITHRA_SOLVER_NSFD;
  v.trunc2 = g.truncation_order == 2;
  for (int l = 0; l < 3; ++l) v.lvl[l] = d->lvl[l];
  v.eb = d->eb;
  for (int t = 0; t < 6; ++t) v.a[t] = d->co.a[t];
  v.alpha = d->co.alpha;
  v.beta = d->co.beta;
  for (int t = 0; t < 5; ++t) {
    v.mur.bB[t] = d->co.bB[t]; v.mur.cB[t] = d->co.cB[t]; v.mur.dB[t] = d->co.dB[t];
    v.mur.eE[t] = d->co.eE[t]; v.mur.fE[t] = d->co.fE[t]; v.mur.gE[t] = d->co.gE[t];
  }
  for (int t = 0; t < 17; ++t) v.mur.hC[t] = d->co.hC[t];
  v.ev.mdt = -g.dt; v.ev.dx2 = 2.0 * g.dx; v.ev.dy2 = 2.0 * g.dy; v.ev.dz2 = 2.0 * g.dz;
  v.ev.rmdt = 1.0 / v.ev.mdt; v.ev.rdx2 = 1.0 / v.ev.dx2; v.ev.rdy2 = 1.0 / v.ev.dy2; v.ev.rdz2 = 1.0 / v.ev.dz2;
  v.dt = g.dt;
  if (!d->part_alt) {  // spare arrays of the periodic reorder
    if (hipMalloc(&d->part_alt, sizeof(double) * PCOUNT * (size_t)d->cap) != hipSuccess) return MITHRA_E_NOMEM;
    if (hipMalloc(&d->pid_alt, sizeof(unsigned) * (size_t)d->cap) != hipSuccess) {
      (void)hipFree(d->part_alt);
      d->part_alt = nullptr;
      return MITHRA_E_NOMEM;
    }
  }
  v.part = d->part;
  v.part_alt = d->part_alt;
  v.pid = d->pid;
  v.pid_alt = d->pid_alt;
  v.reorder = d->sort_interval;
  v.cap = d->cap;
  v.np = d->np;
  PushConst& c = v.push;  // as launch_push builds it
  c.xmin = g.xmin; c.xmax = g.xmax; c.ymin = g.ymin; c.ymax = g.ymax; c.zmin = g.zmin;
  c.zp0 = g.zp0; c.zp1 = g.zp1; c.dx = g.dx; c.dy = g.dy; c.dz = g.dz; c.Lz = g.Lz;
  c.N0 = g.N0; c.N1 = g.N1; c.N2 = g.N2;
  c.ix = 1.0 / g.dx; c.iy = 1.0 / g.dy; c.iz = 1.0 / g.dz;
  c.gamma = p->gamma; c.beta = p->beta; c.c0 = p->c0; c.dt_shift = p->dt_shift;
  c.r1 = p->r1; c.r2 = p->r2; c.dtb = p->dtb;
  c.tb0 = time_bunch; c.dtb_time = bunch_time_step;
  c.n_sub = n_sub; c.n_und = n_und; c.save_rnm = 1;
  c.dist0 = n_und > 0 ? und[0].dist : 0.0;
  v.und = d->und_dev;
  v.counters = d->counters;
  v.n_sub = n_sub;
  v.dtb = bunch_time_step;
  v.dep = make_deposit_const(d);
  v.ord[0] = d->ord[0];
  v.ord[1] = d->ord[1];
  v.rows = rows;
  v.dTI = deposit_tile_rows(d);
  v.dtilesI = (g.N0 + v.dTI - 1) / v.dTI;
  void* mem;
  if (int rc = alloc_owned(b, sizeof(unsigned) * ((size_t)rows + 1), &mem)) return rc;
  v.cnt = (unsigned*)mem;
  BATCH_TRY(launch_fill_zero(d, v.cnt, sizeof(unsigned) * ((size_t)rows + 1)));
  if (int rc = alloc_owned(b, sizeof(unsigned) * ((size_t)rows / 1024 + 2), &mem)) return rc;
  v.bsum = (unsigned*)mem;
  v.ring = d->ring;
  v.dft = d->dft;
  v.pk = d->pk;
  v.pdzr = d->pdzr;
  v.partial = d->partial;
  v.history = d->history;
  v.plane_of = d->pk ? d->pk + d->pNz : nullptr;
  v.pNz = d->prows ? d->pNz : 0;
  v.pN = d->pN;
  v.pNl = d->pNl;
  v.pNf = d->pNf;
  v.prows = d->prows;
  v.pblocks = d->pblocks;
  v.pgamma = d->pgamma; v.pbeta = d->pbeta; v.pc0 = d->pc0; v.pc = d->pc;
  v.bk = bunching_k;
  v.bevery = bunching_every;
  v.bslots = bunching_slots;
  if (int rc = alloc_owned(b, sizeof(double) * (bunching_slots ? bunching_slots : 1), &mem)) return rc;
  v.bout = (double*)mem;
  BATCH_TRY(launch_fill_zero(d, v.bout, sizeof(double) * (bunching_slots ? bunching_slots : 1)));
  if (int rc = alloc_owned(b, sizeof(double) * 3 * (size_t)((d->np + 255) / 256 + 1), &mem)) return rc;
  v.bpart = (double*)mem;
  BATCH_TRY(hipStreamSynchronize(d->stream));
  b->ready[i] = 1;
  b->step[i] = 0;
  return MITHRA_OK;
}

int mithra_batch_set_tiles(mithra_batch* b, int variant, int use_graph) {
  if (!b || variant < 0 || variant >= kTileVariants || (use_graph != 0 && use_graph != 1)) return MITHRA_E_ARG;
  if (b->prepared) return MITHRA_E_STATE;
  b->variant = variant;
  b->use_graph = use_graph != 0;
  return MITHRA_OK;
}

int mithra_batch_run(mithra_batch* b, long steps, long* done) {
  if (!b || steps < 0) return MITHRA_E_ARG;
  if (!b->prepared)
    if (int rc = prepare(b)) return rc;
  long k = 0;
  for (; k < steps; ++k) {
    bool any = false;
    for (int i = 0; i < b->n; ++i) any = any || b->step[i] < b->hv[i].n_steps;
    if (!any) break;
    // the graph captured at preparation replays every step to the end of the scan: each
    // kernel skips the table entries of instances whose descriptor says they are finished
    if (b->exec) {
      BATCH_TRY(hipGraphLaunch(b->exec, b->stream));
    } else if (int rc = enqueue_step(b)) {
      return rc;
    }
    for (int i = 0; i < b->n; ++i)
      if (b->step[i] < b->hv[i].n_steps) ++b->step[i];
  }
  if (done) *done = k;
  return write_back(b);
}

int mithra_batch_steps(const mithra_batch* b, int i, long* step) {
  if (!b || i < 0 || i >= b->n || !step) return MITHRA_E_ARG;
  *step = b->step[i];
  return MITHRA_OK;
}

int mithra_batch_bunching(mithra_batch* b, int i, double* out) {
  if (!b || i < 0 || i >= b->n || !out) return MITHRA_E_ARG;
  if (!b->ready[i]) return MITHRA_E_STATE;
  BATCH_TRY(hipStreamSynchronize(b->stream));
  if (b->hv[i].bslots)
    BATCH_TRY(hipMemcpy(out, b->hv[i].bout, sizeof(double) * b->hv[i].bslots, hipMemcpyDeviceToHost));
  return MITHRA_OK;
}

int mithra_batch_autotune(mithra_batch* b, int reps, int* chosen, float* ms) {
  if (!b || reps < 1) return MITHRA_E_ARG;
  if (b->prepared) return MITHRA_E_STATE;
  for (int i = 0; i < b->n; ++i)
    if (!b->ready[i]) return MITHRA_E_STATE;
  // The interior update overwrites the J it reads: save every instance's A^{n+1} level first.
  std::vector<void*> saved(b->n, nullptr);
  std::vector<size_t> bytes(b->n);
  int rc = MITHRA_OK;
  for (int i = 0; i < b->n && !rc; ++i) {
    mithra_dev* d = b->dev[i];
    bytes[i] = (size_t)d->ncomp * ((d->nodes * d->real_size + 255) / 256 * 256);
    if (hipMalloc(&saved[i], bytes[i]) != hipSuccess) rc = MITHRA_E_NOMEM;
    else rc = status(hipMemcpyAsync(saved[i], d->lvl[d->i_np1].c[0], bytes[i], hipMemcpyDeviceToDevice, b->stream));
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc) rc = status(hipEventCreate(&e0));
  if (!rc) rc = status(hipEventCreate(&e1));
  const int keep = b->variant;
  int best = keep;
  float best_ms = 0.0f;
  for (int v = 0; v < kTileVariants && !rc; ++v) {
    b->variant = v;
    for (int i = 0; i < b->n; ++i) set_tile_geometry(b->hv[i], v);
    rc = upload_views(b);
    if (!rc) rc = build_tables(b);
    float total = 0.0f;
    for (int r = 0; r <= reps && !rc; ++r) {  // r = 0 warms up
      for (int i = 0; i < b->n && !rc; ++i) {
        mithra_dev* d = b->dev[i];
        rc = status(hipMemcpyAsync(d->lvl[d->i_np1].c[0], saved[i], bytes[i], hipMemcpyDeviceToDevice, b->stream));
      }
      if (!rc) rc = status(hipEventRecord(e0, b->stream));
      if (!rc) rc = status(launch_batch_interior(b->V, b->t_int[1], v, b->precision, 1, b->stream));
      if (!rc) rc = status(launch_batch_interior(b->V, b->t_int[0], v, b->precision, 0, b->stream));
      if (!rc) rc = status(launch_batch_mur(b->V, b->t_faces, b->t_edges, b->t_corners, b->precision, b->stream));
      if (!rc) rc = status(launch_batch_eval(b->V, b->t_eval, v, b->precision, b->stream));
      if (!rc) rc = status(hipEventRecord(e1, b->stream));
      if (!rc) rc = status(hipEventSynchronize(e1));
      float t = 0.0f;
      if (!rc) rc = status(hipEventElapsedTime(&t, e0, e1));
      if (r > 0) total += t;
    }
    if (!rc && (v == 0 || total < best_ms)) {
      best_ms = total;
      best = v;
    }
    if (ms && !rc) ms[v] = total / reps;
  }
  for (int i = 0; i < b->n; ++i) {
    if (saved[i]) {
      mithra_dev* d = b->dev[i];
      hipError_t e = hipMemcpyAsync(d->lvl[d->i_np1].c[0], saved[i], bytes[i], hipMemcpyDeviceToDevice, b->stream);
      if (!rc && e != hipSuccess) rc = status(e);
    }
  }
  hipError_t e = hipStreamSynchronize(b->stream);
  if (!rc && e != hipSuccess) rc = status(e);
  for (void* p : saved) (void)hipFree(p);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  b->variant = rc ? keep : best;
  for (int i = 0; i < b->n; ++i) set_tile_geometry(b->hv[i], b->variant);
  if (!rc && chosen) *chosen = best;
  return rc;
}

}  // extern "C"
