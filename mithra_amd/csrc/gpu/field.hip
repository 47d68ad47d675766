// Field kernels: interior NSFD/FD update, TF/SF seed injection, Mur absorbing boundaries
// and the E/B evaluation, for one domain resident in HBM.
//
// All arithmetic keeps the reference's source order and the library is built with
// -ffp-contract=off, so on identical inputs the FP64 path reproduces the reference's
// potentials bit for bit.
#include <algorithm>
#include <cmath>
#include <initializer_list>
#include <utility>
#include <vector>

#include "common.h"

namespace mithra {

namespace {

constexpr int kTileJ = 64;      // one wavefront along y (contiguous in memory)
constexpr int kNumXcd = 8;      // MI355X: 8 accelerator complex dies, each with its own L2

// Workgroups are dispatched round-robin over the XCDs (workgroup b runs on XCD b % 8).
// Returns the logical tile of workgroup b such that every XCD receives one contiguous range
// of logical tiles: neighbouring tiles, which share their halo rows, then meet in one L2.
__device__ inline int xcd_logical_block(int b, int nblk) {
  int q = nblk / kNumXcd, r = nblk % kNumXcd;
  int xcd = b % kNumXcd, local = b / kNumXcd;
  int base = (xcd < r) ? xcd * (q + 1) : r * (q + 1) + (xcd - r) * q;
  return base + local;
}

struct Plane5 {  // one z plane of a column and its four transverse neighbours
  double c, xp, xm, yp, ym;
};

template <typename PT>
__device__ inline Plane5 load_plane(PT a, long m, long N1) {
  Plane5 p;
  p.c = (double)a[m];
  p.xp = (double)a[m + N1];
  p.xm = (double)a[m - N1];
  p.yp = (double)a[m + 1];
  p.ym = (double)a[m - 1];
  return p;
}

// Interior update (AdvanceField::advance*PotentialNSFD / FD, reference src/database.cpp:44-134,
// loops src/fdtd.cpp:270-303 and src/fdtdSC.cpp:317-366).  Each thread owns one (i, j) column
// and marches kChunkK planes along z with the three planes it needs held in registers; the
// x/y neighbours it reads are the centres of neighbouring lanes and rows, served by L1/L2.
// J (or rho) lives in the A^{n+1} buffer and is overwritten in place, as in the reference.
template <typename R, bool NSFD, int TI, int K, bool FACES = false>
__device__ inline void interior_body(FieldSet an_set, FieldSet anm1_set, FieldSet anp1_set, int ncomp, int N0,
                                     int N1, int N2, int tilesJ, int tilesI, int chunks, double a0, double a1,
                                     double a2, double a3, double a4, double a5, double alpha, double beta, int lb,
                                     const MurYFace& yf = MurYFace{}) {
  const int tj = lb % tilesJ;
  lb /= tilesJ;
  const int ti = lb % tilesI;
  lb /= tilesI;
  const int kc = lb % chunks;
  const int comp = lb / chunks;
  const int lane = threadIdx.x % kTileJ;
  const int j = tj * kTileJ + lane;
  const int i = 1 + ti * TI + (threadIdx.x / kTileJ);
  const int kb = 1 + kc * K;
  const int ke = min(kb + K, N2 - 1);
  if (comp >= ncomp || kb >= ke) return;  // uniform over the workgroup
  // FACES: as in interior_lds_body, the lanes of the face nodes j = 0 and N1-1 update the Mur y
  // faces in `yf.mask`; they take the inward neighbour's new value, its A^{n-1} and its x
  // neighbours of A^n by lane shuffle, and their own A^n column from their register queue.
  const bool face_tile = FACES && (tj == 0 || tj == (N1 - 1) / kTileJ);
  const bool active = i <= N0 - 2 && j >= 1 && j <= N1 - 2;
  const bool face = FACES && i <= N0 - 2 && (((yf.mask & 1) && j == 0) || ((yf.mask & 2) && j == N1 - 1));
  if (!active && !face_tile) return;
  const bool load = active || face;
  const int fs = j == 0 ? 1 : -1;
  const int nb_lane = max(lane + fs, 0);
  const gptr<const R> an = G((const R*)component(an_set, comp));
  const gptr<const R> anm1 = G((const R*)component(anm1_set, comp));
  const gptr<R> anp1 = G((R*)component(anp1_set, comp));
  const double as = comp < 3 ? a4 : a5;
  const long N1N0 = (long)N1 * N0;
  long m = N1N0 * kb + (long)N1 * i + j;
  Plane5 pm = {}, p0 = {};
  if (load) {
    pm = load_plane(an, m - N1N0, N1);
    p0 = load_plane(an, m, N1);
  }
  for (int k = kb; k < ke; ++k, m += N1N0) {
    Plane5 pp = {};
    double v1 = 0.0, r = 0.0;
    if (load) {
      pp = load_plane(an, m + N1N0, N1);
      v1 = (double)ld_stream(anm1, m);
    }
    if (active) {
      const double src = (double)ld_stream(G((const R*)component(anp1_set, comp)), m);
      if (NSFD) {
        double tx = p0.xp + p0.xm + beta * (pp.xp + pm.xp + pp.xm + pm.xm);
        double ty = p0.yp + p0.ym + beta * (pp.yp + pm.yp + pp.ym + pm.ym);
        r = a0 * p0.c - v1 + alpha * (a1 * tx + a2 * ty) + a3 * (pp.c + pm.c) + as * src;
      } else {
        r = a0 * p0.c - v1 + a1 * (p0.xp + p0.xm) + a2 * (p0.yp + p0.ym) + a3 * (pp.c + pm.c) + as * src;
      }
      st_stream(anp1, m, (R)r);
    }
    if (face_tile) {
      const double pn = __shfl((double)(R)r, nb_lane);
      const double mn = __shfl(v1, nb_lane);
      const double nxp = __shfl(p0.xp, nb_lane), nxm = __shfl(p0.xm, nb_lane);
      if (face) {
        const double* B = yf.B;
        const double ns = fs > 0 ? p0.yp : p0.ym, nsp = fs > 0 ? pp.yp : pp.ym, nsm = fs > 0 ? pm.yp : pm.ym;
        anp1[m] = (R)(B[0] * (v1 + pn) + B[1] * mn + B[2] * (p0.c + ns) + B[3] * (nxp + nxm + p0.xp + p0.xm) +
                      B[4] * (nsp + nsm + pp.c + pm.c));
      }
    }
    pm = p0;
    p0 = pp;
  }
}

template <typename R, bool NSFD, int TI, int K, bool FACES>
__global__ __launch_bounds__(kTileJ* TI) void k_interior(
    FieldSet an_set, FieldSet anm1_set, FieldSet anp1_set, int ncomp, int N0, int N1, int N2,
    int tilesJ, int tilesI, int chunks, double a0, double a1, double a2, double a3, double a4, double a5,
    double alpha, double beta, MurYFace yf) {
  interior_body<R, NSFD, TI, K, FACES>(an_set, anm1_set, anp1_set, ncomp, N0, N1, N2, tilesJ, tilesI, chunks, a0,
                                       a1, a2, a3, a4, a5, alpha, beta, xcd_logical_block(blockIdx.x, gridDim.x), yf);
}

// The interior update with the A^n planes staged in LDS (2.5-D blocking).  The workgroup (TI
// wavefronts, one x-row of 64 y-nodes each) marches its K planes in lockstep: plane k+1 of its
// tile plus a one-node halo ring is fetched once, coalesced, into a three-plane LDS ring, and
// every neighbour of the stencil is read from LDS.  Same arithmetic, in the same order, as
// k_interior.
//
// FACES: the workgroups of the first and last tile columns also update the Mur y faces in
// `yf.mask` (mur_face below with s = +-1 along y, t along x, u along z): the lane of face node
// j = 0 (N1-1) takes the new value and the A^{n-1} value of node j = 1 (N1-2) from the
// neighbouring lane and the A^n values from the LDS ring, so the faces cost no extra HBM pass.
// Valid only when nothing writes A^{n+1} between the interior update and the boundaries.
template <typename R, bool NSFD, int TI, int K, bool FACES = false>
__device__ inline void interior_lds_body(FieldSet an_set, FieldSet anm1_set, FieldSet anp1_set, int ncomp, int N0,
                                         int N1, int N2, int tilesJ, int tilesI, int chunks, double a0, double a1,
                                         double a2, double a3, double a4, double a5, double alpha, double beta,
                                         int lb, const MurYFace& yf = MurYFace{}) {
  constexpr int W = kTileJ + 2;           // LDS row: y-halo, 64 nodes, y-halo
  constexpr int P = (TI + 2) * W;         // one plane of the tile with its halo ring
  __shared__ double L[3 * P];
  const int tj = lb % tilesJ;
  lb /= tilesJ;
  const int ti = lb % tilesI;
  lb /= tilesI;
  const int kc = lb % chunks;
  const int comp = lb / chunks;
  const int kb = 1 + kc * K;
  const int ke = min(kb + K, N2 - 1);
  if (comp >= ncomp || kb >= ke) return;  // uniform over the workgroup
  const int lane = threadIdx.x % kTileJ, w = threadIdx.x / kTileJ;
  const int j0 = tj * kTileJ, i0 = 1 + ti * TI;
  const int j = j0 + lane, i = i0 + w;
  const bool active = i <= N0 - 2 && j >= 1 && j <= N1 - 2;
  const gptr<const R> an = G((const R*)component(an_set, comp));
  const gptr<const R> anm1 = G((const R*)component(anm1_set, comp));
  const gptr<R> anp1 = G((R*)component(anp1_set, comp));
  const double as = comp < 3 ? a4 : a5;
  const long N1N0 = (long)N1 * N0;
  // What this thread stages per plane: its own node, and (wave 0 / wave 1) the halo rows
  // i0-1 / i0+TI, and (lanes < 2 TI of wave 0) the halo columns j0-1 / j0+64 of rows i0..i0+TI-1.
  const bool own_ok = i <= N0 - 1 && j <= N1 - 1;
  const long own_off = (long)N1 * i + j;
  const int own_slot = (w + 1) * W + lane + 1;
  int hrow_i = -1, hrow_slot = 0;
  if (w == 0) { hrow_i = i0 - 1; hrow_slot = lane + 1; }
  if (w == 1) { hrow_i = i0 + TI; hrow_slot = (TI + 1) * W + lane + 1; }
  const bool hrow_ok = hrow_i >= 0 && hrow_i <= N0 - 1 && j <= N1 - 1;
  const long hrow_off = (long)N1 * hrow_i + j;
  int hcol_j = -1, hcol_i = 0, hcol_slot = 0;
  if (w == 0 && lane < TI) { hcol_i = i0 + lane; hcol_j = j0 - 1; hcol_slot = (lane + 1) * W; }
  if (w == 0 && lane >= TI && lane < 2 * TI) {
    hcol_i = i0 + lane - TI; hcol_j = j0 + kTileJ; hcol_slot = (lane - TI + 1) * W + kTileJ + 1;
  }
  const bool hcol_ok = hcol_j >= 0 && hcol_j <= N1 - 1 && hcol_i <= N0 - 1;
  const long hcol_off = (long)N1 * hcol_i + hcol_j;
  // y faces: tile columns that hold a face node (uniform over the workgroup), the face lanes,
  // and the lane each reads its inward neighbour from
  const bool face_tile = FACES && (tj == 0 || tj == (N1 - 1) / kTileJ);
  const int fs = j == 0 ? 1 : -1;
  const bool face = FACES && i <= N0 - 2 && (((yf.mask & 1) && j == 0) || ((yf.mask & 2) && j == N1 - 1));
  const int nb_lane = max(lane + fs, 0);
  // Prefetched values stay in the storage type until they are staged or used: a conversion
  // right after a load would make the wave wait for that load there (FP32-bulk).
  R r_own = 0, r_row = 0, r_col = 0;
  auto fetch = [&](int k) {
    const long b = N1N0 * k;
    if (own_ok) r_own = an[b + own_off];
    if (hrow_ok) r_row = an[b + hrow_off];
    if (hcol_ok) r_col = an[b + hcol_off];
  };
  auto stage = [&](int slot) {
    double* p = L + slot * P;
    if (own_ok) p[own_slot] = (double)r_own;
    if (hrow_ok) p[hrow_slot] = (double)r_row;
    if (hcol_ok) p[hcol_slot] = (double)r_col;
  };
  // A^{n-1} and J of the next plane are loaded one plane ahead, like the A^n planes, so their
  // latency overlaps the current plane's work.
  const gptr<const R> jsrc = G((const R*)component(anp1_set, comp));
  const bool need_m1 = active || face;
  R v1_next = 0, src_next = 0;
  fetch(kb - 1);
  stage(0);
  fetch(kb);
  stage(1);
  fetch(kb + 1);
  long m = N1N0 * kb + own_off;
  if (need_m1) v1_next = anm1[m];
  if (active) src_next = jsrc[m];
  for (int k = kb; k < ke; ++k, m += N1N0) {
    const int s0 = (k - kb + 1) % 3, sp = (k - kb + 2) % 3, sm = (k - kb) % 3;
    const double v1 = (double)v1_next, src = (double)src_next;
    stage(sp);
    __syncthreads();
    if (k + 2 <= ke) fetch(k + 2);
    if (k + 1 < ke) {
      if (need_m1) v1_next = anm1[m + N1N0];
      if (active) src_next = jsrc[m + N1N0];
    }
    const double* c0 = L + s0 * P + own_slot;
    const double* cp = L + sp * P + own_slot;
    const double* cm = L + sm * P + own_slot;
    double r = 0.0;
    if (active) {
      if (NSFD) {
        double tx = c0[W] + c0[-W] + beta * (cp[W] + cm[W] + cp[-W] + cm[-W]);
        double ty = c0[1] + c0[-1] + beta * (cp[1] + cm[1] + cp[-1] + cm[-1]);
        r = a0 * c0[0] - v1 + alpha * (a1 * tx + a2 * ty) + a3 * (cp[0] + cm[0]) + as * src;
      } else {
        r = a0 * c0[0] - v1 + a1 * (c0[W] + c0[-W]) + a2 * (c0[1] + c0[-1]) + a3 * (cp[0] + cm[0]) + as * src;
      }
      anp1[m] = (R)r;
    }
    if (face_tile) {
      // the inward neighbour's A^{n+1} as stored, and its A^{n-1}
      const double pn = __shfl((double)(R)r, nb_lane);
      const double mn = __shfl(v1, nb_lane);
      if (face) {
        const double* B = yf.B;
        anp1[m] = (R)(B[0] * (v1 + pn) + B[1] * mn + B[2] * (c0[0] + c0[fs]) +
                      B[3] * (c0[fs + W] + c0[fs - W] + c0[W] + c0[-W]) +
                      B[4] * (cp[fs] + cm[fs] + cp[0] + cm[0]));
      }
    }
    __syncthreads();
  }
}

template <typename R, bool NSFD, int TI, int K, bool FACES>
__global__ __launch_bounds__(kTileJ* TI) void k_interior_lds(
    FieldSet an_set, FieldSet anm1_set, FieldSet anp1_set, int ncomp, int N0, int N1, int N2,
    int tilesJ, int tilesI, int chunks, double a0, double a1, double a2, double a3, double a4, double a5,
    double alpha, double beta, MurYFace yf) {
  interior_lds_body<R, NSFD, TI, K, FACES>(an_set, anm1_set, anp1_set, ncomp, N0, N1, N2, tilesJ, tilesI, chunks,
                                           a0, a1, a2, a3, a4, a5, alpha, beta,
                                           xcd_logical_block(blockIdx.x, gridDim.x), yf);
}

// ---- analytic seed (Signal::self src/classes.cpp:534-575, Seed::fields :740-855) --------

// Plane-wave, truncated plane-wave and Gaussian-beam seeds (the types whose output does not
// depend on the previous call).  Writes the boosted-frame vector potential a[3].
__device__ void seed_fields(const mithra_seed& s, double r0, double r1, double r2, double time, double a[3]) {
  double rv[3];
  rv[0] = r0 - s.position[0];
  rv[1] = r1 - s.position[1];
  rv[2] = s.gamma * (r2 + s.beta * s.c0 * (time + s.dt_shift)) - s.position[2];
  double tl = s.gamma * (time + s.dt_shift + s.beta / s.c0 * r2);
  const double z = rv[0] * s.direction[0] + rv[1] * s.direction[1] + rv[2] * s.direction[2];
  tl -= z / s.c0;
  const double* pol = s.polarization;
  const double* dir = s.direction;
  double amp = 0.0;
  double ts = signal(s, tl, 0.0);
  bool zero = fabs(ts) < 1.0e-6;
  if (s.seed_type == MITHRA_SEED_PLANEWAVE) {
    amp = s.amplitude * ts;
  } else {
    double x = rv[0] * pol[0] + rv[1] * pol[1] + rv[2] * pol[2];
    double yv0 = dir[1] * pol[2] - dir[2] * pol[1];
    double yv1 = dir[2] * pol[0] - dir[0] * pol[2];
    double yv2 = dir[0] * pol[1] - dir[1] * pol[0];
    double y = rv[0] * yv0 + rv[1] * yv1 + rv[2] * yv2;
    if (s.seed_type == MITHRA_SEED_PLANEWAVE_TRUNCATED) {
      zero = zero || fabs(x) > s.radius[0] || fabs(y) > s.radius[1];
      amp = s.amplitude * ts;
    } else if (!zero) {
      const double l = s.c0 / s.f0;
      const double zRp = kPI * s.radius[0] * s.radius[0] / l;
      const double wrp = sqrt(1.0 + z * z / (zRp * zRp));
      const double zRs = kPI * s.radius[1] * s.radius[1] / l;
      const double wrs = sqrt(1.0 + z * z / (zRs * zRs));
      const double p = 0.5 * (atan(z / zRp) + atan(z / zRs) - kPI) -
                       kPI * z / l * (pow(x / (zRp * wrp), 2) + pow(y / (zRs * wrs), 2));
      ts = signal(s, tl, p);
      const double t = exp(-pow(x / (s.radius[0] * wrp), 2) - pow(y / (s.radius[1] * wrs), 2)) / sqrt(wrs * wrp) *
                       s.amplitude;
      amp = t * ts;
    }
  }
  if (zero) {
    a[0] = a[1] = a[2] = 0.0;
  } else {
    a[0] = amp * pol[0];
    a[1] = amp * pol[1];
    a[2] = amp * pol[2];
  }
  a[2] *= s.gamma;
}

template <typename R>
__device__ inline void tfsf_apply(FieldSet A, long m, double px, double py, double pz, double time,
                                  const mithra_seed& s, double coef, bool minus) {
  double a[3];
  seed_fields(s, px, py, pz, time, a);
  for (int c = 0; c < 3; ++c) {
    R* p = (R*)A.c[c];
    double v = (double)p[m];
    p[m] = (R)(minus ? v - coef * a[c] : v + coef * a[c]);
  }
}

// TF/SF injection (src/fdtd.cpp:307-373).  axis 0: planes i = 1, 2, N0-2, N0-3 over (j, k);
// axis 1: planes j = 1, 2, N1-2, N1-3 over (i, k); axis 2: planes k = 1, 2, np-2, np-3 over (i, j).
// The three axes run as three launches in the reference's loop order.  Launches carry four work
// items per line.  Where the axis has 6 nodes or more its four planes are distinct, and work item
// t = 64 (4 g + q) + l applies injection q of line 64 g + l alone (q is uniform over a
// wavefront, so each wavefront follows one branch).  On a shorter axis two planes coincide, and
// work item t < lines applies all four injections of line t in the reference's order, so the node
// touched twice receives its contributions in the same sequence.
template <typename R>
__device__ inline void tfsf_body(FieldSet A, int axis, int N0, int N1, int N2, const mithra_grid& g,
                                 const mithra_seed& s, double time, double coef, long t) {
  const long N1N0 = (long)N1 * N0;
  const bool split = (axis == 0 ? N0 : axis == 1 ? N1 : N2) >= 6;
  int q = 0;
  if (split) {
    q = (int)((t >> 6) & 3);
    t = ((t >> 8) << 6) + (t & 63);
  }
  const bool q0 = !split || q == 0, q1 = !split || q == 1, q2 = !split || q == 2, q3 = !split || q == 3;
  if (axis == 0) {
    const long nj = N1 - 4, nk = N2 - 4;
    if (t >= nj * nk) return;
    const long j = 2 + t % nj, k = 2 + t / nj;
    const double y = g.ymin + j * g.dy, z = g.zmin + k * g.dz;
    const long base = N1N0 * k + j;
    if (q0) tfsf_apply<R>(A, base + N1 * 1, g.xmin + 2 * g.dx, y, z, time, s, coef, true);
    if (q1) tfsf_apply<R>(A, base + N1 * 2, g.xmin + 1 * g.dx, y, z, time, s, coef, false);
    if (q2) tfsf_apply<R>(A, base + (long)N1 * (N0 - 2), g.xmin + (N0 - 3) * g.dx, y, z, time, s, coef, true);
    if (q3) tfsf_apply<R>(A, base + (long)N1 * (N0 - 3), g.xmin + (N0 - 2) * g.dx, y, z, time, s, coef, false);
  } else if (axis == 1) {
    const long ni = N0 - 4, nk = N2 - 4;
    if (t >= ni * nk) return;
    const long i = 2 + t % ni, k = 2 + t / ni;
    const double x = g.xmin + i * g.dx, z = g.zmin + k * g.dz;
    const long base = N1N0 * k + (long)N1 * i;
    if (q0) tfsf_apply<R>(A, base + 1, x, g.ymin + 2 * g.dy, z, time, s, coef, true);
    if (q1) tfsf_apply<R>(A, base + 2, x, g.ymin + 1 * g.dy, z, time, s, coef, false);
    if (q2) tfsf_apply<R>(A, base + N1 - 2, x, g.ymin + (N1 - 3) * g.dy, z, time, s, coef, true);
    if (q3) tfsf_apply<R>(A, base + N1 - 3, x, g.ymin + (N1 - 2) * g.dy, z, time, s, coef, false);
  } else {
    const long ni = N0 - 4, nj = N1 - 4;
    if (t >= ni * nj) return;
    const long j = 2 + t % nj, i = 2 + t / nj;
    const double x = g.xmin + i * g.dx, y = g.ymin + j * g.dy;
    const long base = (long)N1 * i + j;
    if (q0) tfsf_apply<R>(A, base + N1N0 * 1, x, y, g.zmin + 2 * g.dz, time, s, coef, true);
    if (q1) tfsf_apply<R>(A, base + N1N0 * 2, x, y, g.zmin + 1 * g.dz, time, s, coef, false);
    if (q2) tfsf_apply<R>(A, base + N1N0 * (N2 - 2), x, y, g.zmin + (N2 - 3) * g.dz, time, s, coef, true);
    if (q3) tfsf_apply<R>(A, base + N1N0 * (N2 - 3), x, y, g.zmin + (N2 - 2) * g.dz, time, s, coef, false);
  }
}

template <typename R>
__global__ void k_tfsf(FieldSet A, int axis, int N0, int N1, int N2, mithra_grid g, mithra_seed s, double time,
                       double coef) {
  tfsf_body<R>(A, axis, N0, N1, N2, g, s, time, coef, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// ---- the super-gaussian seed's TF/SF injection ------------------------------------------------
// Seed::fields adds the super-gaussian beam into its output vector instead of overwriting it
// (src/classes.cpp:814-851), and FdTd::fieldUpdate passes the same vector, zeroed once per
// step, to every call of the injection (src/fdtd.cpp:239, :307-373).  Call c therefore injects
// the running result of calls 0..c in the reference's loop order: reset to zero by a call whose
// signal is below the threshold, otherwise the beam's (2 o0 + 1)(2 o1 + 1) equal terms added one
// by one, then the z component times gamma.  The calls are numbered in that loop order:
//   x planes  (j outer, k inner, four planes each), y planes (i outer, k inner, four planes each),
//   z = 1, 2 planes (i outer, j inner, two planes each), z = np-2, np-3 planes (likewise).
struct SgCall {
  long m;            // node that receives the injection
  double r[3];       // point the seed is evaluated at (the neighbour across the TF/SF plane)
  int block;         // 0: x planes, 1: y planes, 2: z = 1, 2, 3: z = np-2, np-3
  bool minus;        // mmv (subtract) or pmv (add)
};

__device__ inline SgCall sg_call(const mithra_grid& g, long c) {
  const int N0 = g.N0, N1 = g.N1, N2 = g.N2;
  const long N1N0 = (long)N1 * N0, ni = N0 - 4, nj = N1 - 4, nk = N2 - 4;
  SgCall q;
  const long cx = 4 * nj * nk, cy = 4 * ni * nk, cz = 2 * ni * nj;
  if (c < cx) {
    const long t = c / 4, s = c % 4, j = 2 + t / nk, k = 2 + t % nk;
    const long i = s == 0 ? 1 : s == 1 ? 2 : s == 2 ? N0 - 2 : N0 - 3;
    const long ip = s == 0 ? 2 : s == 1 ? 1 : s == 2 ? N0 - 3 : N0 - 2;  // i of the evaluated node
    q.m = N1N0 * k + N1 * i + j;
    q.r[0] = g.xmin + ip * g.dx; q.r[1] = g.ymin + j * g.dy; q.r[2] = g.zmin + k * g.dz;
    q.block = 0;
    q.minus = (s % 2) == 0;
  } else if (c < cx + cy) {
    const long u = c - cx, t = u / 4, s = u % 4, i = 2 + t / nk, k = 2 + t % nk;
    const long j = s == 0 ? 1 : s == 1 ? 2 : s == 2 ? N1 - 2 : N1 - 3;
    const long jp = s == 0 ? 2 : s == 1 ? 1 : s == 2 ? N1 - 3 : N1 - 2;
    q.m = N1N0 * k + N1 * i + j;
    q.r[0] = g.xmin + i * g.dx; q.r[1] = g.ymin + jp * g.dy; q.r[2] = g.zmin + k * g.dz;
    q.block = 1;
    q.minus = (s % 2) == 0;
  } else {
    const long u = c - cx - cy, upper = u >= cz, v = upper ? u - cz : u, t = v / 2, s = v % 2;
    const long i = 2 + t / nj, j = 2 + t % nj;
    const long k = upper ? (s == 0 ? N2 - 2 : N2 - 3) : (s == 0 ? 1 : 2);
    const long kp = upper ? (s == 0 ? N2 - 3 : N2 - 2) : (s == 0 ? 2 : 1);
    q.m = N1N0 * k + N1 * i + j;
    q.r[0] = g.xmin + i * g.dx; q.r[1] = g.ymin + j * g.dy; q.r[2] = g.zmin + kp * g.dz;
    q.block = upper ? 3 : 2;
    q.minus = s == 0;
  }
  return q;
}

// One call's beam term t * tsignal at boosted-frame point r (every one of its (2 o0 + 1)(2 o1 + 1)
// terms is the same), or a reset (returns true) when the signal is below the threshold
// (src/classes.cpp:742-755, :816-850).
__device__ bool sg_term(const mithra_seed& s, double r0, double r1, double r2, double time, double* term) {
  const double rl2 = s.gamma * (r2 + s.beta * s.c0 * (time + s.dt_shift));
  double tl = s.gamma * (time + s.dt_shift + s.beta / s.c0 * r2);
  const double rv[3] = {r0 - s.position[0], r1 - s.position[1], rl2 - s.position[2]};
  const double z = rv[0] * s.direction[0] + rv[1] * s.direction[1] + rv[2] * s.direction[2];
  tl -= z / s.c0;
  double ts = signal(s, tl, 0.0);
  if (fabs(ts) < 1.0e-6) {
    *term = 0.0;
    return true;
  }
  const double* pol = s.polarization;
  const double* dir = s.direction;
  const double x = rv[0] * pol[0] + rv[1] * pol[1] + rv[2] * pol[2];
  const double yv0 = dir[1] * pol[2] - dir[2] * pol[1];
  const double yv1 = dir[2] * pol[0] - dir[0] * pol[2];
  const double yv2 = dir[0] * pol[1] - dir[1] * pol[0];
  const double y = rv[0] * yv0 + rv[1] * yv1 + rv[2] * yv2;
  const double l = s.c0 / s.f0;
  const double zRp = kPI * s.radius[0] * s.radius[0] / l;
  const double wrp = sqrt(1.0 + z * z / (zRp * zRp));
  const double zRs = kPI * s.radius[1] * s.radius[1] / l;
  const double wrs = sqrt(1.0 + z * z / (zRs * zRs));
  const double p = 0.5 * (atan(z / zRp) + atan(z / zRs) - kPI) -
                   kPI * z / l * (pow(x / (zRp * wrp), 2) + pow(y / (zRs * wrs), 2));
  ts = signal(s, tl, p);
  const double t = exp(-pow(x / (s.radius[0] * wrp), 2) - pow(y / (s.radius[1] * wrs), 2)) / sqrt(wrs * wrp) *
                   s.amplitude;
  *term = t * ts;
  return false;
}

__global__ void k_sg_terms(mithra_grid g, mithra_seed s, double time, long n, double* __restrict__ term,
                           int* __restrict__ reset) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  const SgCall q = sg_call(g, c);
  reset[c] = sg_term(s, q.r[0], q.r[1], q.r[2], time, &term[c]) ? 1 : 0;
}

// The same terms at listed points and times (the point evaluator's call order).
__global__ void k_sg_point_terms(mithra_seed s, const double* __restrict__ r, const double* __restrict__ time,
                                 long n, double* __restrict__ term, int* __restrict__ reset) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= n) return;
  reset[c] = sg_term(s, r[3 * c], r[3 * c + 1], r[3 * c + 2], time[c], &term[c]) ? 1 : 0;
}

// The running vector in call order: strictly sequential, one lane.
__global__ void k_sg_running(mithra_seed s, long n, const double* __restrict__ term, const int* __restrict__ reset,
                             double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const int reps = (2 * s.order[0] + 1) * (2 * s.order[1] + 1);
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (long c = 0; c < n; ++c) {
    if (reset[c]) {
      a0 = a1 = a2 = 0.0;
    } else {
      const double v = term[c];
      for (int r = 0; r < reps; ++r) {
        a0 += v * s.polarization[0];
        a1 += v * s.polarization[1];
        a2 += v * s.polarization[2];
      }
    }
    a2 *= s.gamma;
    out[3 * c] = a0;
    out[3 * c + 1] = a1;
    out[3 * c + 2] = a2;
  }
}

// The injection of one block of calls (nodes are distinct within a block; the blocks run in
// the reference's order, so a node two blocks touch gets its two updates in sequence).
template <typename R>
__global__ void k_sg_apply(FieldSet A, mithra_grid g, long first, long count, const double* __restrict__ run,
                           double coef) {
  const long u = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= count) return;
  const long c = first + u;
  const SgCall q = sg_call(g, c);
  for (int k = 0; k < 3; ++k) {
    R* p = (R*)A.c[k];
    const double v = (double)p[q.m];
    p[q.m] = (R)(q.minus ? v - coef * run[3 * c + k] : v + coef * run[3 * c + k]);
  }
}

// The analytic seed on its own, at listed points: one thread per point.
__global__ void k_signal_eval(mithra_seed s, const double* __restrict__ t, const double* __restrict__ phase, long n,
                              double* __restrict__ out) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = signal(s, t[q], phase[q]);
}

__global__ void k_seed_eval(mithra_seed s, const double* __restrict__ r, const double* __restrict__ time, long n,
                            double* __restrict__ a) {
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  double v[3];
  seed_fields(s, r[3 * q], r[3 * q + 1], r[3 * q + 2], time[q], v);
  a[3 * q] = v[0];
  a[3 * q + 1] = v[1];
  a[3 * q + 2] = v[2];
}

// ---- Mur absorbing boundaries (src/database.cpp:137-288, loops src/fdtd.cpp:377-725) ------
// Offsets are in nodes.  For a face node, s is the offset of its inward neighbour and t, u the
// two tangential axes (B[3] multiplies the t-neighbours, B[4] the u-neighbours).  For an edge
// node, o1 and o2 are the two inward offsets and w the offset along the edge.  For a corner,
// sx, sy, sz are the three inward offsets.

template <typename PT>
__device__ inline double rd(PT p, long m) { return (double)p[m]; }

template <typename TP, typename TN>
__device__ inline double mur_face(const double* B, TP P, TN N, TN M, long m, long s, long t, long u) {
  return B[0] * (rd(M, m) + rd(P, m + s)) + B[1] * rd(M, m + s) + B[2] * (rd(N, m) + rd(N, m + s)) +
         B[3] * (rd(N, m + s + t) + rd(N, m + s - t) + rd(N, m + t) + rd(N, m - t)) +
         B[4] * (rd(N, m + s + u) + rd(N, m + s - u) + rd(N, m + u) + rd(N, m - u));
}

template <typename TP, typename TN>
__device__ inline double mur_edge(const double* B, TP P, TN N, TN M, long m, long o1, long o2, long w) {
  const long o3 = o1 + o2;
  return B[0] * (rd(P, m + o1) + rd(M, m + o2)) + B[1] * (rd(M, m + o1) + rd(P, m + o2)) +
         B[2] * (rd(M, m) + rd(P, m + o3)) + B[3] * (rd(N, m) + rd(N, m + o1) + rd(N, m + o2) + rd(N, m + o3)) +
         B[4] * (rd(N, m - w) + rd(N, m - w + o1) + rd(N, m - w + o2) + rd(N, m - w + o3) + rd(N, m + w) +
                 rd(N, m + w + o1) + rd(N, m + w + o2) + rd(N, m + w + o3)) -
         rd(M, m + o3);
}

template <typename TP, typename TN>
__device__ inline double mur_corner(const double* h, TP P, TN N, TN M, long m, long sx, long sy,
                                    long sz) {
  const long o[7] = {sx, sy, sz, sx + sy, sz + sx, sz + sy, sz + sx + sy};
  double s = rd(N, m) * h[16] + rd(M, m) * h[8];
  for (int t = 0; t < 7; ++t) s = s + rd(P, m + o[t]) * h[1 + t] + rd(N, m + o[t]) * h[16] + rd(M, m + o[t]) * h[9 + t];
  return -s / h[0];
}


// All six faces in one launch: x faces (j, k), y faces (i, k), z faces (i, j), each min/max.
// The y faces in `ymask` (yface_mask) were updated by the interior kernel and are skipped.
template <typename R>
__device__ inline void mur_faces_body(FieldSet P, FieldSet N, FieldSet M, int ncomp, int N0, int N1, int N2,
                                      const MurCoeffs& c, long t, int comp, int ymask) {
  const long N1N0 = (long)N1 * N0;
  const long X = N1, Y = 1, Z = N1N0;
  const long nx = (long)(N1 - 2) * (N2 - 2), ny = (long)(N0 - 2) * (N2 - 2), nz = (long)(N0 - 2) * (N1 - 2);
  const long nylo = (ymask & 1) ? 0 : ny, nyhi = (ymask & 2) ? 0 : ny;
  if (comp >= ncomp) return;
  const gptr<const R> pn = G((const R*)component(N, comp));
  const gptr<const R> pm = G((const R*)component(M, comp));
  const gptr<R> pp = G((R*)component(P, comp));
  long m, s, tt, u;
  int axis;
  if (t < 2 * nx) {
    const bool hi = t >= nx;
    t -= hi ? nx : 0;
    const long j = 1 + t % (N1 - 2), k = 1 + t / (N1 - 2);
    m = N1N0 * k + (hi ? N1N0 - N1 : 0) + j;
    s = hi ? -X : X; tt = Y; u = Z; axis = 0;
  } else if ((t -= 2 * nx) < nylo + nyhi) {
    const bool hi = t >= nylo;
    t -= hi ? nylo : 0;
    const long i = 1 + t % (N0 - 2), k = 1 + t / (N0 - 2);
    m = N1N0 * k + (long)N1 * i + (hi ? N1 - 1 : 0);
    s = hi ? -Y : Y; tt = X; u = Z; axis = 1;
  } else if ((t -= nylo + nyhi) < 2 * nz) {
    const bool hi = t >= nz;
    t -= hi ? nz : 0;
    const long j = 1 + t % (N1 - 2), i = 1 + t / (N1 - 2);
    m = (hi ? N1N0 * (N2 - 1) : 0) + (long)N1 * i + j;
    s = hi ? -Z : Z; tt = X; u = Y; axis = 2;
  } else {
    return;
  }
  // The coefficients by selects into registers: a pointer into the coefficient struct would
  // copy the whole struct to scratch memory in every thread.
  double B[5];
#pragma unroll
  for (int q = 0; q < 5; ++q) B[q] = axis == 0 ? c.bB[q] : axis == 1 ? c.cB[q] : c.dB[q];
  pp[m] = (R)mur_face(B, pp, pn, pm, m, s, tt, u);
}

template <typename R>
__global__ void k_mur_faces(FieldSet P, FieldSet N, FieldSet M, int ncomp, int N0, int N1, int N2, MurCoeffs c,
                           int ymask) {
  mur_faces_body<R>(P, N, M, ncomp, N0, N1, N2, c, (long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y, ymask);
}

// The twelve edges (second-order Mur only).
template <typename R>
__device__ inline void mur_edges_body(FieldSet P, FieldSet N, FieldSet M, int ncomp, int N0, int N1, int N2,
                                      const MurCoeffs& c, long t, int comp) {
  const long N1N0 = (long)N1 * N0;
  const long X = N1, Y = 1, Z = N1N0;
  const long ez = N2 - 2, ex = N0 - 2, ey = N1 - 2;
  if (comp >= ncomp) return;
  const gptr<const R> pn = G((const R*)component(N, comp));
  const gptr<const R> pm = G((const R*)component(M, comp));
  const gptr<R> pp = G((R*)component(P, comp));
  long m, o1, o2, w;
  int axis;
  if (t < 4 * ez) {  // along z at the four (x, y) corners
    const int which = t / ez;
    const long k = 1 + t % ez;
    const bool xh = which & 1, yh = which & 2;
    m = N1N0 * k + (xh ? (long)N1 * (N0 - 1) : 0) + (yh ? N1 - 1 : 0);
    o1 = xh ? -X : X; o2 = yh ? -Y : Y; w = Z; axis = 0;
  } else if ((t -= 4 * ez) < 4 * ex) {  // along x at the four (y, z) corners
    const int which = t / ex;
    const long i = 1 + t % ex;
    const bool yh = which & 1, zh = which & 2;
    m = (long)N1 * i + (yh ? N1 - 1 : 0) + (zh ? N1N0 * (N2 - 1) : 0);
    o1 = yh ? -Y : Y; o2 = zh ? -Z : Z; w = X; axis = 1;
  } else if ((t -= 4 * ex) < 4 * ey) {  // along y at the four (z, x) corners
    const int which = t / ey;
    const long j = 1 + t % ey;
    const bool zh = which & 1, xh = which & 2;
    m = j + (zh ? N1N0 * (N2 - 1) : 0) + (xh ? N1N0 - N1 : 0);
    o1 = zh ? -Z : Z; o2 = xh ? -X : X; w = Y; axis = 2;
  } else {
    return;
  }
  double B[5];  // by selects, as in mur_faces_body
#pragma unroll
  for (int q = 0; q < 5; ++q) B[q] = axis == 0 ? c.eE[q] : axis == 1 ? c.fE[q] : c.gE[q];
  pp[m] = (R)mur_edge(B, pp, pn, pm, m, o1, o2, w);
}

template <typename R>
__global__ void k_mur_edges(FieldSet P, FieldSet N, FieldSet M, int ncomp, int N0, int N1, int N2, MurCoeffs c) {
  mur_edges_body<R>(P, N, M, ncomp, N0, N1, N2, c, (long)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

template <typename R>
__device__ inline void mur_corners_body(FieldSet P, FieldSet N, FieldSet M, int ncomp, int N0, int N1, int N2,
                                        const MurCoeffs& c, int t, int comp) {
  if (t >= 8 || comp >= ncomp) return;
  const long N1N0 = (long)N1 * N0;
  const bool xh = t & 1, yh = t & 2, zh = t & 4;
  const long m = (xh ? N1N0 - N1 : 0) + (yh ? N1 - 1 : 0) + (zh ? N1N0 * (N2 - 1) : 0);
  const gptr<const R> pn = G((const R*)component(N, comp));
  const gptr<const R> pm = G((const R*)component(M, comp));
  const gptr<R> pp = G((R*)component(P, comp));
  pp[m] = (R)mur_corner(c.hC, pp, pn, pm, m, xh ? -(long)N1 : N1, yh ? -1L : 1L, zh ? -N1N0 : N1N0);
}

template <typename R>
__global__ void k_mur_corners(FieldSet P, FieldSet N, FieldSet M, int ncomp, int N0, int N1, int N2, MurCoeffs c) {
  mur_corners_body<R>(P, N, M, ncomp, N0, N1, N2, c, threadIdx.x, blockIdx.y);
}

// E/B on every interior node (FdTd::fieldEvaluate src/fdtd.cpp:818-845, FdTdSC::fieldEvaluate
// src/fdtdSC.cpp:1110-1145); planes k = 0 and N2-1 take copies of k = 1 and N2-2
// (src/fdtd.cpp:742-774).  E_x and E_y are one double expression rounded once to float and E_z
// keeps the float rounding after each term: that is what the reference computes as compiled by
// g++ -O3 (its vectorised x/y lanes drop the intermediate float round trip of
// FieldVector<float>::dv/mdv, src/fieldvector.h:87-108), and the parity fixtures record it.
// Lazy evaluation (the reference evaluates E/B only where particles and samplers read them,
// src/solver.h:241, src/fdtd.cpp:818-845): planes kb..ke-1 are needed if they meet the nodes the
// bunch reads during the coming push (its cells zr[0]..zr[1] at the start, at most one cell of
// motion per field step, nodes k and k+1 of a cell) or the power planes pb0..pb1.
__device__ inline bool lazy_needed(int kb, int ke, const int* zr, int pb0, int pb1) {
  const int a0 = zr[0] - 2, a1 = zr[1] + 2;
  return (ke - 1 >= a0 && kb <= a1) || (ke - 1 >= pb0 && kb <= pb1);
}


// Each thread owns one (i, j) column and marches K planes; the z neighbours of A_x and A_y (and
// phi) at both time levels ride in a three-plane register queue, the transverse neighbours are
// the centres of neighbouring lanes and rows.  Planes outside [klo, khi] are skipped.
template <typename R, bool SC, int TI, int K>
__device__ inline void eval_body(FieldSet An, FieldSet Ap, float* __restrict__ eb, int N0, int N1, int N2,
                                 int tilesJ, int tilesI, int klo, int khi, const EvalConst& q,
                                 const int* __restrict__ zr, int pb0, int pb1, int lb) {
  const gptr<f2> ebg = G(reinterpret_cast<f2*>(eb));
  const int tj = lb % tilesJ;
  lb /= tilesJ;
  const int ti = lb % tilesI;
  const int kc = lb / tilesI;
  const int j = tj * kTileJ + (threadIdx.x % kTileJ);
  const int i = 1 + ti * TI + (threadIdx.x / kTileJ);
  const int kb = klo + kc * K;
  const int ke = min(kb + K, khi + 1);
  if (zr && !lazy_needed(kb, ke, zr, pb0, pb1)) return;
  if (i > N0 - 2 || j < 1 || j > N1 - 2 || kb >= ke) return;
  const long N1N0 = (long)N1 * N0;
  const gptr<const R> n0 = G((const R*)An.c[0]);
  const gptr<const R> n1 = G((const R*)An.c[1]);
  const gptr<const R> n2 = G((const R*)An.c[2]);
  const gptr<const R> p0 = G((const R*)Ap.c[0]);
  const gptr<const R> p1 = G((const R*)Ap.c[1]);
  const gptr<const R> p2 = G((const R*)Ap.c[2]);
  const gptr<const R> f = G(SC ? (const R*)An.c[3] : (const R*)nullptr);
  const double mdt = q.mdt, dx2 = q.dx2, dy2 = q.dy2, dz2 = q.dz2;
  auto dt_ = [&](double x) { return div_rn(x, mdt, q.rmdt); };
  auto dx_ = [&](double x) { return div_rn(x, dx2, q.rdx2); };
  auto dy_ = [&](double x) { return div_rn(x, dy2, q.rdy2); };
  auto dz_ = [&](double x) { return div_rn(x, dz2, q.rdz2); };
  long m = N1N0 * kb + (long)N1 * i + j;
  // queue: [0] plane k-1, [1] plane k, [2] plane k+1
  double qn0[3], qn1[3], qp0[3], qp1[3], qf[3];
  for (int t = 0; t < 2; ++t) {
    const long mm = m + (t - 1) * N1N0;
    qn0[t] = rd(n0, mm); qn1[t] = rd(n1, mm); qp0[t] = rd(p0, mm); qp1[t] = rd(p1, mm);
    if (SC) qf[t] = rd(f, mm);
  }
  for (int k = kb; k < ke; ++k, m += N1N0) {
    qn0[2] = rd(n0, m + N1N0); qn1[2] = rd(n1, m + N1N0); qp0[2] = rd(p0, m + N1N0); qp1[2] = rd(p1, m + N1N0);
    if (SC) qf[2] = rd(f, m + N1N0);
    double ex = dt_(qp0[1]) - dt_(qn0[1]);
    double ey = dt_(qp1[1]) - dt_(qn1[1]);
    float ez = (float)dt_(rd(p2, m));
    ez = (float)((double)ez - dt_(rd(n2, m)));
    if (SC) {
      ex = ex - dx_(rd(f, m + N1) - rd(f, m - N1));
      ey = ey - dy_(rd(f, m + 1) - rd(f, m - 1));
      ez = (float)((double)ez - dz_(qf[2] - qf[0]));
    }
    const float bx = (float)(0.5 * (dy_(rd(n2, m + 1) - rd(n2, m - 1)) - dz_(qn1[2] - qn1[0]) +
                                    dy_(rd(p2, m + 1) - rd(p2, m - 1)) - dz_(qp1[2] - qp1[0])));
    const float by = (float)(0.5 * (dz_(qn0[2] - qn0[0]) - dx_(rd(n2, m + N1) - rd(n2, m - N1)) +
                                    dz_(qp0[2] - qp0[0]) - dx_(rd(p2, m + N1) - rd(p2, m - N1))));
    const float bz = (float)(0.5 * (dx_(rd(n1, m + N1) - rd(n1, m - N1)) - dy_(rd(n0, m + 1) - rd(n0, m - 1)) +
                                    dx_(rd(p1, m + N1) - rd(p1, m - N1)) - dy_(rd(p0, m + 1) - rd(p0, m - 1))));
    const f2 v0 = {(float)ex, (float)ey}, v1 = {ez, bx}, v2 = {by, bz};
    auto put = [&](long node) {  // E/B is read by later kernels only: stream past L2
      st_stream(ebg, kEBPairs * node, v0);
      st_stream(ebg, kEBPairs * node + 1, v1);
      st_stream(ebg, kEBPairs * node + 2, v2);
    };
    put(m);
    if (k == 1) put(m - N1N0);
    if (k == N2 - 2) put(m + N1N0);
    qn0[0] = qn0[1]; qn1[0] = qn1[1]; qp0[0] = qp0[1]; qp1[0] = qp1[1];
    qn0[1] = qn0[2]; qn1[1] = qn1[2]; qp0[1] = qp0[2]; qp1[1] = qp1[2];
    if (SC) { qf[0] = qf[1]; qf[1] = qf[2]; }
  }
}

template <typename R, bool SC, int TI, int K>
__global__ __launch_bounds__(kTileJ* TI) void k_eval_eb(FieldSet An, FieldSet Ap, float* __restrict__ eb, int N0,
                                                        int N1, int N2, int tilesJ, int tilesI, int klo, int khi,
                                                        EvalConst q, const int* __restrict__ zr, int pb0, int pb1) {
  eval_body<R, SC, TI, K>(An, Ap, eb, N0, N1, N2, tilesJ, tilesI, klo, khi, q, zr, pb0, pb1,
                          xcd_logical_block(blockIdx.x, gridDim.x));
}

// The E/B evaluation with plane k of A^n and A^{n+1} staged in LDS (2.5-D blocking, as
// k_interior_lds): the workgroup marches its planes in lockstep, every transverse neighbour is
// read from LDS, the z neighbours of A_x, A_y (and phi) ride in a register queue, and the
// next plane is fetched while the current one is computed.  Same expressions as k_eval_eb.
template <typename R, bool SC, int TI, int K>
__device__ inline void eval_lds_body(FieldSet An, FieldSet Ap, float* __restrict__ eb, int N0, int N1, int N2,
                                     int tilesJ, int tilesI, int klo, int khi, const EvalConst& q,
                                     const int* __restrict__ zr, int pb0, int pb1, int lb) {
  const gptr<f2> ebg = G(reinterpret_cast<f2*>(eb));
  constexpr int NA = SC ? 7 : 6;      // staged arrays: n0 n1 n2 p0 p1 p2 [phi]
  constexpr int NZ = SC ? 5 : 4;      // arrays with z neighbours: n0 n1 p0 p1 [phi]
  constexpr int W = kTileJ + 2, P = (TI + 2) * W;
  __shared__ double L[NA * P];
  const int tj = lb % tilesJ;
  lb /= tilesJ;
  const int ti = lb % tilesI;
  const int kc = lb / tilesI;
  const int kb = klo + kc * K;
  const int ke = min(kb + K, khi + 1);
  if (kb >= ke) return;  // uniform over the workgroup
  if (zr && !lazy_needed(kb, ke, zr, pb0, pb1)) return;
  const int lane = threadIdx.x % kTileJ, w = threadIdx.x / kTileJ;
  const int j0 = tj * kTileJ, i0 = 1 + ti * TI;
  const int j = j0 + lane, i = i0 + w;
  const bool active = i <= N0 - 2 && j >= 1 && j <= N1 - 2;
  const long N1N0 = (long)N1 * N0;
  const gptr<const R> A[7] = {G((const R*)An.c[0]), G((const R*)An.c[1]), G((const R*)An.c[2]),
                              G((const R*)Ap.c[0]), G((const R*)Ap.c[1]), G((const R*)Ap.c[2]),
                              G(SC ? (const R*)An.c[3] : (const R*)nullptr)};
  constexpr int zarr[5] = {0, 1, 3, 4, 6};  // n0 n1 p0 p1 phi
  const bool own_ok = i <= N0 - 1 && j <= N1 - 1;
  const long own_off = (long)N1 * i + j;
  const int own_slot = (w + 1) * W + lane + 1;
  int hrow_i = -1, hrow_slot = 0;
  if (w == 0) { hrow_i = i0 - 1; hrow_slot = lane + 1; }
  if (w == 1) { hrow_i = i0 + TI; hrow_slot = (TI + 1) * W + lane + 1; }
  const bool hrow_ok = hrow_i >= 0 && hrow_i <= N0 - 1 && j <= N1 - 1;
  const long hrow_off = (long)N1 * hrow_i + j;
  int hcol_j = -1, hcol_i = 0, hcol_slot = 0;
  if (w == 0 && lane < TI) { hcol_i = i0 + lane; hcol_j = j0 - 1; hcol_slot = (lane + 1) * W; }
  if (w == 0 && lane >= TI && lane < 2 * TI) {
    hcol_i = i0 + lane - TI; hcol_j = j0 + kTileJ; hcol_slot = (lane - TI + 1) * W + kTileJ + 1;
  }
  const bool hcol_ok = hcol_j >= 0 && hcol_j <= N1 - 1 && hcol_i <= N0 - 1;
  const long hcol_off = (long)N1 * hcol_i + hcol_j;
  // qz[a][s]: own value of z-array a at planes k-1, k, k+1, k+2 (s = 0..3) of iteration k.
  // Prefetched values stay in the storage type until they are staged or used: a conversion
  // right after a load would make the wave wait for that load there (FP32-bulk).
  R qz[NZ][4];
  R pre_n2 = 0, pre_p2 = 0, pre_row[NA], pre_col[NA];
  auto fetch_z = [&](int k, int s) {
#pragma unroll
    for (int a = 0; a < NZ; ++a) qz[a][s] = own_ok ? A[zarr[a]][N1N0 * k + own_off] : (R)0;
  };
  // The curl reads A_x (arrays 0, 3) only at y +- 1 and A_y (arrays 1, 4) only at x +- 1: they
  // need no halo rows and no halo columns respectively.
  auto need_row = [](int a) { return a != 0 && a != 3; };
  auto need_col = [](int a) { return a != 1 && a != 4; };
  auto fetch_plane = [&](int k) {  // what plane k stages besides the queued own values
    const long b = N1N0 * k;
    if (own_ok) {
      pre_n2 = A[2][b + own_off];
      pre_p2 = A[5][b + own_off];
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      if (need_row(a) && hrow_ok) pre_row[a] = A[a][b + hrow_off];
      if (need_col(a) && hcol_ok) pre_col[a] = A[a][b + hcol_off];
    }
  };
  fetch_z(kb - 1, 0);
  fetch_z(kb, 1);
  fetch_z(kb + 1, 2);
  fetch_plane(kb);
  const double mdt = q.mdt, dx2 = q.dx2, dy2 = q.dy2, dz2 = q.dz2;
  auto dt_ = [&](double x) { return div_rn(x, mdt, q.rmdt); };
  auto dx_ = [&](double x) { return div_rn(x, dx2, q.rdx2); };
  auto dy_ = [&](double x) { return div_rn(x, dy2, q.rdy2); };
  auto dz_ = [&](double x) { return div_rn(x, dz2, q.rdz2); };
  long m = N1N0 * kb + own_off;
  for (int k = kb; k < ke; ++k, m += N1N0) {
    // stage plane k
    if (own_ok) {
      L[0 * P + own_slot] = (double)qz[0][1];
      L[1 * P + own_slot] = (double)qz[1][1];
      L[2 * P + own_slot] = (double)pre_n2;
      L[3 * P + own_slot] = (double)qz[2][1];
      L[4 * P + own_slot] = (double)qz[3][1];
      L[5 * P + own_slot] = (double)pre_p2;
      if (SC) L[6 * P + own_slot] = (double)qz[NZ - 1][1];
    }
#pragma unroll
    for (int a = 0; a < NA; ++a) {
      if (need_row(a) && hrow_ok) L[a * P + hrow_slot] = (double)pre_row[a];
      if (need_col(a) && hcol_ok) L[a * P + hcol_slot] = (double)pre_col[a];
    }
    __syncthreads();
    if (k + 1 < ke) fetch_plane(k + 1);
    if (k + 2 <= ke) fetch_z(k + 2, 3);
    if (active) {
      const double* n1l = L + 1 * P + own_slot;
      const double* n2l = L + 2 * P + own_slot;
      const double* p1l = L + 4 * P + own_slot;
      const double* p2l = L + 5 * P + own_slot;
      const double* n0l = L + 0 * P + own_slot;
      const double* p0l = L + 3 * P + own_slot;
      double ex = dt_((double)qz[2][1]) - dt_((double)qz[0][1]);
      double ey = dt_((double)qz[3][1]) - dt_((double)qz[1][1]);
      float ez = (float)dt_(p2l[0]);
      ez = (float)((double)ez - dt_(n2l[0]));
      if (SC) {
        const double* fl = L + 6 * P + own_slot;
        ex = ex - dx_(fl[W] - fl[-W]);
        ey = ey - dy_(fl[1] - fl[-1]);
        ez = (float)((double)ez - dz_((double)qz[NZ - 1][2] - (double)qz[NZ - 1][0]));
      }
      const float bx = (float)(0.5 * (dy_(n2l[1] - n2l[-1]) - dz_((double)qz[1][2] - (double)qz[1][0]) +
                                      dy_(p2l[1] - p2l[-1]) - dz_((double)qz[3][2] - (double)qz[3][0])));
      const float by = (float)(0.5 * (dz_((double)qz[0][2] - (double)qz[0][0]) - dx_(n2l[W] - n2l[-W]) +
                                      dz_((double)qz[2][2] - (double)qz[2][0]) - dx_(p2l[W] - p2l[-W])));
      const float bz = (float)(0.5 * (dx_(n1l[W] - n1l[-W]) - dy_(n0l[1] - n0l[-1]) +
                                      dx_(p1l[W] - p1l[-W]) - dy_(p0l[1] - p0l[-1])));
      const f2 v0 = {(float)ex, (float)ey}, v1 = {ez, bx}, v2 = {by, bz};
      auto put = [&](long node) {  // E/B is read by later kernels only: stream past L2
        st_stream(ebg, kEBPairs * node, v0);
        st_stream(ebg, kEBPairs * node + 1, v1);
        st_stream(ebg, kEBPairs * node + 2, v2);
      };
      put(m);
      if (k == 1) put(m - N1N0);
      if (k == N2 - 2) put(m + N1N0);
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < NZ; ++a) {
      qz[a][0] = qz[a][1];
      qz[a][1] = qz[a][2];
      qz[a][2] = qz[a][3];
    }
  }
}

template <typename R, bool SC, int TI, int K>
__global__ __launch_bounds__(kTileJ* TI) void k_eval_lds(FieldSet An, FieldSet Ap, float* __restrict__ eb, int N0,
                                                         int N1, int N2, int tilesJ, int tilesI, int klo, int khi,
                                                         EvalConst q, const int* __restrict__ zr, int pb0, int pb1) {
  eval_lds_body<R, SC, TI, K>(An, Ap, eb, N0, N1, N2, tilesJ, tilesI, klo, khi, q, zr, pb0, pb1,
                              xcd_logical_block(blockIdx.x, gridDim.x));
}

// ---- pipelined step: E/B of step n fused with the J-free part of the next interior update ----
//
// Once A^{n+1} is complete (interior, seed injection and Mur boundaries), the two field passes
// that remain before the next deposit read the same two levels: the E/B evaluation of step n
// reads A^n and A^{n+1}, and the interior update of step n+1 reads A^{n+1} (stencil) and A^n
// (its A^{n-1}).  Only the update's last term, a4 J^{n+1}, waits for the deposit.  The update
// (src/database.cpp:44-134) sums its terms left to right with the source last, so
//   A^{n+2} = X + a4 J,  X = a0 A^{n+1} - A^n + (transverse terms) + a3 (z terms)
// splits exactly: this kernel writes X (into the A^{n-1} buffer, free once the boundaries of
// A^{n+1} are done) and the E/B of step n; the row-tiled deposit then writes X + a4 J in place
// of J.  In FP64 the result is the unsplit update's bit for bit; in FP32 X is rounded to float
// once before the source is added.
//
// A workgroup owns TI x-rows x 64 y-nodes of all three components and marches K planes: A^{n+1}
// with a one-node halo ring in a three-plane LDS ring per component (the stencil's planes k-1,
// k, k+1), A^n of plane k with its ring in one LDS plane per component (the curl's transverse
// neighbours), and the z neighbours of A^n_x, A^n_y in a register queue.  Same expressions, in
// the same order, as k_interior_lds (without the source) and k_eval_lds (without phi).
struct StencilConst {
  double a0, a1, a2, a3, alpha, beta;
};
#ifndef MITHRA_STENCIL_WAVES
#define MITHRA_STENCIL_WAVES 4
#endif
constexpr int kStencilWaves = MITHRA_STENCIL_WAVES;  // waves per SIMD the register budget allows

template <typename R, bool NSFD, int TI, int K, bool SC = false>
__global__ __launch_bounds__(kTileJ* TI, SC ? 3 : kStencilWaves) void k_stencil_eb(
    FieldSet An1, FieldSet An, FieldSet X, float* __restrict__ eb, int N0, int N1, int N2, int tilesJ, int tilesI,
    StencilConst s, EvalConst q, const int* __restrict__ zr, int pb0, int pb1) {
  // Components: A_x, A_y, A_z, and phi under space charge (the same stencil; its E/B terms are
  // the gradient of phi^n, as in eval_lds_body).
  constexpr int NC = SC ? 4 : 3;
  constexpr int W = kTileJ + 2, P = (TI + 2) * W;
  __shared__ double L1[3 * NC * P];  // A^{n+1}: [component][slot][P]
  __shared__ double L0[NC * P];      // A^n, plane k: [component][P]
  int lb = xcd_logical_block(blockIdx.x, gridDim.x);
  const int tj = lb % tilesJ;
  lb /= tilesJ;
  const int ti = lb % tilesI;
  const int kc = lb / tilesI;
  const int kb = 1 + kc * K;
  const int ke = min(kb + K, N2 - 1);
  if (kb >= ke) return;  // uniform over the workgroup
  const bool ebneed = !zr || lazy_needed(kb, ke, zr, pb0, pb1);
  const gptr<f2> ebg = G(reinterpret_cast<f2*>(eb));
  const int lane = threadIdx.x % kTileJ, w = threadIdx.x / kTileJ;
  const int j0 = tj * kTileJ, i0 = 1 + ti * TI;
  const int j = j0 + lane, i = i0 + w;
  const bool active = i <= N0 - 2 && j >= 1 && j <= N1 - 2;
  const long N1N0 = (long)N1 * N0;
  gptr<const R> p[NC], n[NC];
  gptr<R> x[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    p[c] = G((const R*)component(An1, c));
    n[c] = G((const R*)component(An, c));
    x[c] = G((R*)component(X, c));
  }
  // staged per plane by this thread: its own node, and (as in k_interior_lds) a halo row node
  // (waves 0 and 1) or a halo column node (lanes < 2 TI of wave 0)
  const bool own_ok = i <= N0 - 1 && j <= N1 - 1;
  const long own_off = (long)N1 * i + j;
  const int own_slot = (w + 1) * W + lane + 1;
  int hrow_i = -1, hrow_slot = 0;
  if (w == 0) { hrow_i = i0 - 1; hrow_slot = lane + 1; }
  if (w == 1) { hrow_i = i0 + TI; hrow_slot = (TI + 1) * W + lane + 1; }
  const bool hrow_ok = hrow_i >= 0 && hrow_i <= N0 - 1 && j <= N1 - 1;
  const long hrow_off = (long)N1 * hrow_i + j;
  int hcol_j = -1, hcol_i = 0, hcol_slot = 0;
  if (w == 0 && lane < TI) { hcol_i = i0 + lane; hcol_j = j0 - 1; hcol_slot = (lane + 1) * W; }
  if (w == 0 && lane >= TI && lane < 2 * TI) {
    hcol_i = i0 + lane - TI; hcol_j = j0 + kTileJ; hcol_slot = (lane - TI + 1) * W + kTileJ + 1;
  }
  const bool hcol_ok = hcol_j >= 0 && hcol_j <= N1 - 1 && hcol_i <= N0 - 1;
  const long hcol_off = (long)N1 * hcol_i + hcol_j;
  // A^{n+1} of the plane to stage next.  Prefetched values stay in the storage type until they
  // are used: a conversion right after a load would make the wave wait for that load there,
  // and the prefetch would overlap nothing (FP32-bulk).
  R p_own[NC] = {}, p_row[NC] = {}, p_col[NC] = {};
  auto fetch_p = [&](int k) {
    const long b = N1N0 * k;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (own_ok) p_own[c] = p[c][b + own_off];
      if (hrow_ok) p_row[c] = p[c][b + hrow_off];
      if (hcol_ok) p_col[c] = p[c][b + hcol_off];
    }
  };
  auto stage_p = [&](int slot) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double* d = L1 + (c * 3 + slot) * P;
      if (own_ok) d[own_slot] = (double)p_own[c];
      if (hrow_ok) d[hrow_slot] = (double)p_row[c];
      if (hcol_ok) d[hcol_slot] = (double)p_col[c];
    }
  };
  // A^n: own A_x, A_y (and phi) at planes k-1 .. k+2 (qx, qy, qf), own A_z at k and k+1, halo
  // of the next plane.  The curl reads A^n_x only at y +- 1 and A^n_y only at x +- 1, so A^n_x
  // needs no halo rows and A^n_y no halo columns; the gradient of phi needs both.
  R qx[4] = {}, qy[4] = {}, qf[4] = {}, qz = 0, qz_next = 0, n_row[NC] = {}, n_col[NC] = {};
  auto fetch_n_halo = [&](int k) {
    const long b = N1N0 * k;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c != 0 && hrow_ok) n_row[c] = n[c][b + hrow_off];
      if (c != 1 && hcol_ok) n_col[c] = n[c][b + hcol_off];
    }
  };
  fetch_p(kb - 1);
  stage_p(0);
  fetch_p(kb);
  stage_p(1);
  fetch_p(kb + 1);
  if (own_ok) {
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      qx[t] = n[0][N1N0 * (kb - 1 + t) + own_off];
      qy[t] = n[1][N1N0 * (kb - 1 + t) + own_off];
      if (SC) qf[t] = n[NC - 1][N1N0 * (kb - 1 + t) + own_off];
    }
    qz = n[2][N1N0 * kb + own_off];
  }
  fetch_n_halo(kb);
  const double mdt = q.mdt, dx2 = q.dx2, dy2 = q.dy2, dz2 = q.dz2;
  auto dt_ = [&](double v) { return div_rn(v, mdt, q.rmdt); };
  auto dx_ = [&](double v) { return div_rn(v, dx2, q.rdx2); };
  auto dy_ = [&](double v) { return div_rn(v, dy2, q.rdy2); };
  auto dz_ = [&](double v) { return div_rn(v, dz2, q.rdz2); };
  long m = N1N0 * kb + own_off;
  for (int k = kb; k < ke; ++k, m += N1N0) {
    const int s0 = (k - kb + 1) % 3, sp = (k - kb + 2) % 3, sm = (k - kb) % 3;
    stage_p(sp);
    {
      const double own[4] = {(double)qx[1], (double)qy[1], (double)qz, (double)qf[1]};
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double* d = L0 + c * P;
        if (own_ok) d[own_slot] = own[c];
        if (c != 0 && hrow_ok) d[hrow_slot] = (double)n_row[c];
        if (c != 1 && hcol_ok) d[hcol_slot] = (double)n_col[c];
      }
    }
    __syncthreads();
    if (k + 2 <= ke) fetch_p(k + 2);
    if (k + 1 < ke) {
      fetch_n_halo(k + 1);
      if (own_ok) {
        qx[3] = n[0][m + 2 * N1N0];
        qy[3] = n[1][m + 2 * N1N0];
        if (SC) qf[3] = n[NC - 1][m + 2 * N1N0];
        qz_next = n[2][m + N1N0];
      }
    }
    if (active) {
      const double v1[4] = {(double)qx[1], (double)qy[1], (double)qz, (double)qf[1]};
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const double* c0 = L1 + (c * 3 + s0) * P + own_slot;
        const double* cp = L1 + (c * 3 + sp) * P + own_slot;
        const double* cm = L1 + (c * 3 + sm) * P + own_slot;
        double r;
        if (NSFD) {
          double tx = c0[W] + c0[-W] + s.beta * (cp[W] + cm[W] + cp[-W] + cm[-W]);
          double ty = c0[1] + c0[-1] + s.beta * (cp[1] + cm[1] + cp[-1] + cm[-1]);
          r = s.a0 * c0[0] - v1[c] + s.alpha * (s.a1 * tx + s.a2 * ty) + s.a3 * (cp[0] + cm[0]);
        } else {
          r = s.a0 * c0[0] - v1[c] + s.a1 * (c0[W] + c0[-W]) + s.a2 * (c0[1] + c0[-1]) + s.a3 * (cp[0] + cm[0]);
        }
        st_stream(x[c], m, (R)r);  // X and E/B are not read again here: keep L2 for the halos
      }
      if (ebneed) {
        const double* n0l = L0 + own_slot;
        const double* n1l = L0 + P + own_slot;
        const double* n2l = L0 + 2 * P + own_slot;
        const double* p0l = L1 + (0 * 3 + s0) * P + own_slot;
        const double* p1l = L1 + (1 * 3 + s0) * P + own_slot;
        const double* p2l = L1 + (2 * 3 + s0) * P + own_slot;
        const double p0p = L1[(0 * 3 + sp) * P + own_slot], p0m = L1[(0 * 3 + sm) * P + own_slot];
        const double p1p = L1[(1 * 3 + sp) * P + own_slot], p1m = L1[(1 * 3 + sm) * P + own_slot];
        double ex = dt_(p0l[0]) - dt_((double)qx[1]);
        double ey = dt_(p1l[0]) - dt_((double)qy[1]);
        float ez = (float)dt_(p2l[0]);
        ez = (float)((double)ez - dt_(n2l[0]));
        if (SC) {  // minus the gradient of phi^n (eval_lds_body's expressions)
          const double* fl = L0 + (NC - 1) * P + own_slot;
          ex = ex - dx_(fl[W] - fl[-W]);
          ey = ey - dy_(fl[1] - fl[-1]);
          ez = (float)((double)ez - dz_((double)qf[2] - (double)qf[0]));
        }
        const float bx = (float)(0.5 * (dy_(n2l[1] - n2l[-1]) - dz_((double)qy[2] - (double)qy[0]) +
                                        dy_(p2l[1] - p2l[-1]) - dz_(p1p - p1m)));
        const float by = (float)(0.5 * (dz_((double)qx[2] - (double)qx[0]) - dx_(n2l[W] - n2l[-W]) +
                                        dz_(p0p - p0m) - dx_(p2l[W] - p2l[-W])));
        const float bz = (float)(0.5 * (dx_(n1l[W] - n1l[-W]) - dy_(n0l[1] - n0l[-1]) +
                                        dx_(p1l[W] - p1l[-W]) - dy_(p0l[1] - p0l[-1])));
        const f2 v0 = {(float)ex, (float)ey}, vv1 = {ez, bx}, v2 = {by, bz};
        auto put = [&](long node) {
          st_stream(ebg, kEBPairs * node, v0);
          st_stream(ebg, kEBPairs * node + 1, vv1);
          st_stream(ebg, kEBPairs * node + 2, v2);
        };
        put(m);
        if (k == 1) put(m - N1N0);
        if (k == N2 - 2) put(m + N1N0);
      }
    }
    __syncthreads();
    qx[0] = qx[1]; qx[1] = qx[2]; qx[2] = qx[3];
    qy[0] = qy[1]; qy[1] = qy[2]; qy[2] = qy[3];
    qf[0] = qf[1]; qf[1] = qf[2]; qf[2] = qf[3];
    qz = qz_next;
  }
}

// ---- batched forms: persistent workgroups walking a table of (instance, local tile) ---------
// Workgroup b takes entries r * G + l(b), r = 0, 1, ..., where l gives every XCD a contiguous
// run of entries, so neighbouring tiles of one instance meet in one L2.
__device__ inline int batch_entry(int r) { return r * (int)gridDim.x + xcd_logical_block(blockIdx.x, gridDim.x); }

template <typename R, bool NSFD, int TI, int K, bool LDS>
__global__ __launch_bounds__(kTileJ* TI) void kb_interior(const BatchView* __restrict__ V, BatchTiles T) {
  for (int r = 0;; ++r) {
    const int t = batch_entry(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (v.active) {
      if constexpr (LDS) {
        MurYFace yf;  // a local copy: stores to the potentials cannot alias it
        for (int q = 0; q < 5; ++q) yf.B[q] = v.mur.cB[q];
        yf.mask = v.yfaces;
        interior_lds_body<R, NSFD, TI, K, true>(v.lvl[v.i_n], v.lvl[v.i_nm1], v.lvl[v.i_np1], v.ncomp, v.N0, v.N1,
                                                v.N2, v.tilesJ, v.tilesI, v.chunks, v.a[0], v.a[1], v.a[2], v.a[3],
                                                v.a[4], v.a[5], v.alpha, v.beta, e.y, yf);
      } else {
        MurYFace yf;
        for (int q = 0; q < 5; ++q) yf.B[q] = v.mur.cB[q];
        yf.mask = v.yfaces;
        interior_body<R, NSFD, TI, K, true>(v.lvl[v.i_n], v.lvl[v.i_nm1], v.lvl[v.i_np1], v.ncomp, v.N0, v.N1, v.N2,
                                            v.tilesJ, v.tilesI, v.chunks, v.a[0], v.a[1], v.a[2], v.a[3], v.a[4],
                                            v.a[5], v.alpha, v.beta, e.y, yf);
      }
    }
    if constexpr (LDS) __syncthreads();  // the LDS ring is reused by the next tile
  }
}

template <typename R, int TI, int K, bool LDS>
__global__ __launch_bounds__(kTileJ* TI) void kb_eval(const BatchView* __restrict__ V, BatchTiles T) {
  for (int r = 0;; ++r) {
    const int t = batch_entry(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (v.active) {
      const EvalConst q = v.ev;  // a local copy: stores to eb cannot alias it
      if constexpr (LDS) {
        if (v.ncomp == 4)  // space charge: E includes -grad phi^n
          eval_lds_body<R, true, TI, K>(v.lvl[v.i_n], v.lvl[v.i_np1], v.eb, v.N0, v.N1, v.N2, v.tilesJ, v.tilesI, 1,
                                        v.N2 - 2, q, nullptr, 0, -1, e.y);
        else
          eval_lds_body<R, false, TI, K>(v.lvl[v.i_n], v.lvl[v.i_np1], v.eb, v.N0, v.N1, v.N2, v.tilesJ, v.tilesI,
                                         1, v.N2 - 2, q, nullptr, 0, -1, e.y);
      } else {
        if (v.ncomp == 4)
          eval_body<R, true, TI, K>(v.lvl[v.i_n], v.lvl[v.i_np1], v.eb, v.N0, v.N1, v.N2, v.tilesJ, v.tilesI, 1,
                                    v.N2 - 2, q, nullptr, 0, -1, e.y);
        else
          eval_body<R, false, TI, K>(v.lvl[v.i_n], v.lvl[v.i_np1], v.eb, v.N0, v.N1, v.N2, v.tilesJ, v.tilesI, 1,
                                     v.N2 - 2, q, nullptr, 0, -1, e.y);
      }
    }
    if constexpr (LDS) __syncthreads();
  }
}

// TF/SF injection of the seeded instances, one axis per launch (the reference's order);
// entry .y = block of 256 lines of the axis.
template <typename R>
__global__ __launch_bounds__(256) void kb_tfsf(const BatchView* __restrict__ V, BatchTiles T, int axis) {
  for (int r = 0;; ++r) {
    const int t = batch_entry(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (!v.active || !v.seeded) continue;
    const double coef = axis == 0 ? v.a[1] : axis == 1 ? v.a[2] : v.a[3];
    tfsf_body<R>(v.lvl[v.i_np1], axis, v.N0, v.N1, v.N2, v.grid, v.seed, v.time, coef,
                 (long)e.y * 256 + threadIdx.x);
  }
}

// Mur faces / edges / corners: entry .y = comp * blocks_per_comp + block
template <typename R, int WHICH>
__global__ __launch_bounds__(256) void kb_mur(const BatchView* __restrict__ V, BatchTiles T) {
  for (int r = 0;; ++r) {
    const int t = batch_entry(r);
    if (t >= T.n) break;
    const int2 e = T.t[t];
    const BatchView& v = V[e.x];
    if (!v.active) continue;
    const FieldSet P = v.lvl[v.i_np1], N = v.lvl[v.i_n], M = v.lvl[v.i_nm1];
    if (WHICH == 0) {
      const long nf = mur_face_count(v.N0, v.N1, v.N2, v.yfaces);
      const int nb = (int)((nf + 255) / 256);
      mur_faces_body<R>(P, N, M, v.ncomp, v.N0, v.N1, v.N2, v.mur, (long)(e.y % nb) * 256 + threadIdx.x, e.y / nb,
                        v.yfaces);
    } else if (WHICH == 1) {
      const long ne = 4L * ((v.N2 - 2) + (v.N0 - 2) + (v.N1 - 2));
      const int nb = (int)((ne + 255) / 256);
      mur_edges_body<R>(P, N, M, v.ncomp, v.N0, v.N1, v.N2, v.mur, (long)(e.y % nb) * 256 + threadIdx.x, e.y / nb);
    } else {
      mur_corners_body<R>(P, N, M, v.ncomp, v.N0, v.N1, v.N2, v.mur, threadIdx.x, e.y);
    }
  }
}

__global__ void k_fill_zero(uint4* p, size_t n16) {
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n16; t += (size_t)gridDim.x * blockDim.x)
    p[t] = make_uint4(0u, 0u, 0u, 0u);
}

__global__ void k_fill_zero_bytes(unsigned char* p, size_t n) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0;
}

// Tile shapes of the marched field kernels: rows of x per workgroup (one wavefront of 64 y
// nodes per row) and z planes per workgroup.  Every shape computes the same values; the
// autotuner (capi) picks the fastest per domain.
template <typename R, bool NSFD, int TI, int K, bool LDS = false>
hipError_t interior_launch(Domain* d, int ymask) {
  const mithra_grid& g = d->g;
  const int tilesJ = (g.N1 + kTileJ - 1) / kTileJ;
  const int tilesI = (g.N0 - 2 + TI - 1) / TI;
  const int chunks = (g.N2 - 2 + K - 1) / K;
  const long nblk = (long)tilesJ * tilesI * chunks * d->ncomp;
  const double* a = d->co.a;
  MurYFace yf;
  for (int q = 0; q < 5; ++q) yf.B[q] = d->co.cB[q];
  yf.mask = ymask;
  if constexpr (LDS) {
    if (ymask)
      hipLaunchKernelGGL((k_interior_lds<R, NSFD, TI, K, true>), dim3((unsigned)nblk), dim3(kTileJ * TI), 0,
                         d->stream, d->lvl[d->i_n], d->lvl[d->i_nm1], d->lvl[d->i_np1], d->ncomp, g.N0, g.N1, g.N2,
                         tilesJ, tilesI, chunks, a[0], a[1], a[2], a[3], a[4], a[5], d->co.alpha, d->co.beta, yf);
    else
      hipLaunchKernelGGL((k_interior_lds<R, NSFD, TI, K, false>), dim3((unsigned)nblk), dim3(kTileJ * TI), 0,
                         d->stream, d->lvl[d->i_n], d->lvl[d->i_nm1], d->lvl[d->i_np1], d->ncomp, g.N0, g.N1, g.N2,
                         tilesJ, tilesI, chunks, a[0], a[1], a[2], a[3], a[4], a[5], d->co.alpha, d->co.beta, yf);
  } else if (ymask) {
    hipLaunchKernelGGL((k_interior<R, NSFD, TI, K, true>), dim3((unsigned)nblk), dim3(kTileJ * TI), 0, d->stream,
                       d->lvl[d->i_n], d->lvl[d->i_nm1], d->lvl[d->i_np1], d->ncomp, g.N0, g.N1, g.N2, tilesJ, tilesI,
                       chunks, a[0], a[1], a[2], a[3], a[4], a[5], d->co.alpha, d->co.beta, yf);
  } else {
    hipLaunchKernelGGL((k_interior<R, NSFD, TI, K, false>), dim3((unsigned)nblk), dim3(kTileJ * TI), 0, d->stream,
                       d->lvl[d->i_n], d->lvl[d->i_nm1], d->lvl[d->i_np1], d->ncomp, g.N0, g.N1, g.N2, tilesJ, tilesI,
                       chunks, a[0], a[1], a[2], a[3], a[4], a[5], d->co.alpha, d->co.beta, yf);
  }
  return hipGetLastError();
}

template <typename R, bool SC, int TI, int K, bool LDS = false>
hipError_t eval_launch(Domain* d, int klo, int khi, const int* zr, int pb0, int pb1) {
  const mithra_grid& g = d->g;
  const int tilesJ = (g.N1 + kTileJ - 1) / kTileJ;
  const int tilesI = (g.N0 - 2 + TI - 1) / TI;
  const int chunks = (khi - klo + 1 + K - 1) / K;
  const long nblk = (long)tilesJ * tilesI * chunks;
  EvalConst q;
  q.mdt = -g.dt; q.dx2 = 2.0 * g.dx; q.dy2 = 2.0 * g.dy; q.dz2 = 2.0 * g.dz;
  q.rmdt = 1.0 / q.mdt; q.rdx2 = 1.0 / q.dx2; q.rdy2 = 1.0 / q.dy2; q.rdz2 = 1.0 / q.dz2;
  if constexpr (LDS)
    hipLaunchKernelGGL((k_eval_lds<R, SC, TI, K>), dim3((unsigned)nblk), dim3(kTileJ * TI), 0, d->stream,
                       d->lvl[d->i_n], d->lvl[d->i_np1], d->eb, g.N0, g.N1, g.N2, tilesJ, tilesI, klo, khi, q, zr, pb0, pb1);
  else
    hipLaunchKernelGGL((k_eval_eb<R, SC, TI, K>), dim3((unsigned)nblk), dim3(kTileJ * TI), 0, d->stream,
                       d->lvl[d->i_n], d->lvl[d->i_np1], d->eb, g.N0, g.N1, g.N2, tilesJ, tilesI, klo, khi, q, zr, pb0, pb1);
  return hipGetLastError();
}

template <typename R, bool FLAG>
hipError_t field_variant(Domain* d, int kernel, int v, int klo, int khi, const int* zr = nullptr, int pb0 = 0,
                         int pb1 = -1, int ymask = 0) {
  // kernel 0: interior (FLAG = NSFD), kernel 1: E/B evaluation (FLAG = space charge)
#define MITHRA_VARIANT(idx, TI, K)                                                            \
  case idx:                                                                                    \
    return kernel == 0 ? interior_launch<R, FLAG, TI, K>(d, ymask) : eval_launch<R, FLAG, TI, K>(d, klo, khi, zr, pb0, pb1);
  switch (v) {
    MITHRA_VARIANT(0, 4, 64)
    MITHRA_VARIANT(1, 2, 64)
    MITHRA_VARIANT(2, 4, 32)
    MITHRA_VARIANT(3, 4, 128)
    MITHRA_VARIANT(4, 8, 32)
    MITHRA_VARIANT(5, 8, 64)
    MITHRA_VARIANT(6, 16, 32)
  }
  static_assert(kTileVariants == 10, "variant table and switch disagree");
  // variants 7..9: the LDS-staged (2.5-D blocked) forms of both kernels
  switch (v) {
    case 7: return kernel == 0 ? interior_launch<R, FLAG, 4, 64, true>(d, ymask) : eval_launch<R, FLAG, 4, 64, true>(d, klo, khi, zr, pb0, pb1);
    case 8: return kernel == 0 ? interior_launch<R, FLAG, 8, 64, true>(d, ymask) : eval_launch<R, FLAG, 8, 64, true>(d, klo, khi, zr, pb0, pb1);
    case 9: return kernel == 0 ? interior_launch<R, FLAG, 16, 64, true>(d, ymask) : eval_launch<R, FLAG, 16, 64, true>(d, klo, khi, zr, pb0, pb1);
  }
#undef MITHRA_VARIANT
  return hipErrorInvalidValue;
}

template <typename R, bool NSFD, int TI, int K, bool SC>
hipError_t stencil_eb_launch(Domain* d, const int* zr, int pb0, int pb1) {
  const mithra_grid& g = d->g;
  const int tilesJ = (g.N1 + kTileJ - 1) / kTileJ;
  const int tilesI = (g.N0 - 2 + TI - 1) / TI;
  const int chunks = (g.N2 - 2 + K - 1) / K;
  const long nblk = (long)tilesJ * tilesI * chunks;
  const double* a = d->co.a;
  const StencilConst s = {a[0], a[1], a[2], a[3], d->co.alpha, d->co.beta};
  EvalConst q;
  q.mdt = -g.dt; q.dx2 = 2.0 * g.dx; q.dy2 = 2.0 * g.dy; q.dz2 = 2.0 * g.dz;
  q.rmdt = 1.0 / q.mdt; q.rdx2 = 1.0 / q.dx2; q.rdy2 = 1.0 / q.dy2; q.rdz2 = 1.0 / q.dz2;
  hipLaunchKernelGGL((k_stencil_eb<R, NSFD, TI, K, SC>), dim3((unsigned)nblk), dim3(kTileJ * TI), 0, d->stream,
                     d->lvl[d->i_np1], d->lvl[d->i_n], d->lvl[d->i_nm1], d->eb, g.N0, g.N1, g.N2, tilesJ, tilesI, s,
                     q, zr, pb0, pb1);
  return hipGetLastError();
}

template <typename R, bool NSFD, bool SC = false>
hipError_t stencil_eb_variant(Domain* d, const int* zr, int pb0, int pb1) {
  static_assert(kFusedVariants == 4, "variant table and switch disagree");
  switch (d->tile_fused) {
    case 0: return stencil_eb_launch<R, NSFD, 4, 64, SC>(d, zr, pb0, pb1);
    case 1: return stencil_eb_launch<R, NSFD, 8, 64, SC>(d, zr, pb0, pb1);
    // under space charge 16-row tiles need more than the 128 registers a 1024-thread
    // workgroup allows (they spill): that variant runs the 8-row form
    case 2:
      if constexpr (SC) return stencil_eb_launch<R, NSFD, 8, 64, SC>(d, zr, pb0, pb1);
      else return stencil_eb_launch<R, NSFD, 16, 64, SC>(d, zr, pb0, pb1);
    case 3: return stencil_eb_launch<R, NSFD, 8, 128, SC>(d, zr, pb0, pb1);
  }
  return hipErrorInvalidValue;
}

template <typename R>
hipError_t stencil_eb_impl(Domain* d) {
  const int* zr = nullptr;
  int pb0 = 0, pb1 = -1;
  if (d->lazy_fields && d->zrange_valid) {
    zr = d->zrange;
    lazy_planes(d, &pb0, &pb1);
  }
  if (d->ncomp == 4)
    return d->g.solver == MITHRA_SOLVER_NSFD ? stencil_eb_variant<R, true, true>(d, zr, pb0, pb1)
                                             : stencil_eb_variant<R, false, true>(d, zr, pb0, pb1);
  return d->g.solver == MITHRA_SOLVER_NSFD ? stencil_eb_variant<R, true>(d, zr, pb0, pb1)
                                           : stencil_eb_variant<R, false>(d, zr, pb0, pb1);
}

template <typename R>
hipError_t interior_impl(Domain* d, bool fuse_yfaces) {
  const int ymask = fuse_yfaces ? yface_mask(d->tile_interior, d->g.N1) : 0;
  const hipError_t e = d->g.solver == MITHRA_SOLVER_NSFD
                           ? field_variant<R, true>(d, 0, d->tile_interior, 0, 0, nullptr, 0, -1, ymask)
                           : field_variant<R, false>(d, 0, d->tile_interior, 0, 0, nullptr, 0, -1, ymask);
  d->yfaces_pending = e == hipSuccess ? ymask : 0;
  return e;
}

template <typename R>
hipError_t tfsf_sg_impl(Domain* d, double time, const mithra_seed* seed) {
  const mithra_grid& g = d->g;
  const long ni = g.N0 - 4, nj = g.N1 - 4, nk = g.N2 - 4;
  const long blocks[4] = {4 * nj * nk, 4 * ni * nk, 2 * ni * nj, 2 * ni * nj};
  const long n = blocks[0] + blocks[1] + blocks[2] + blocks[3];
  if (n <= 0) return hipSuccess;
  if (d->sg_cap < n) {
    hipError_t e = hipStreamSynchronize(d->stream);
    if (e != hipSuccess) return e;
    (void)hipFree(d->sg_buf);
    d->sg_buf = nullptr;
    d->sg_cap = 0;
    if ((e = hipMalloc(&d->sg_buf, (size_t)n * (4 * sizeof(double) + sizeof(int)))) != hipSuccess) return e;
    d->sg_cap = n;
  }
  double* term = (double*)d->sg_buf;
  double* run = term + n;
  int* reset = (int*)(run + 3 * n);
  const unsigned grid = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_sg_terms, dim3(grid), dim3(256), 0, d->stream, g, *seed, time, n, term, reset);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_sg_running, dim3(1), dim3(64), 0, d->stream, *seed, n, term, reset, run);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  const double coef[4] = {d->co.a[1], d->co.a[2], d->co.a[3], d->co.a[3]};
  long first = 0;
  for (int b = 0; b < 4; ++b) {
    if (blocks[b] > 0) {
      hipLaunchKernelGGL((k_sg_apply<R>), dim3((unsigned)((blocks[b] + 255) / 256)), dim3(256), 0, d->stream,
                         d->lvl[d->i_np1], g, first, blocks[b], run, coef[b]);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    first += blocks[b];
  }
  return hipSuccess;
}

template <typename R>
hipError_t tfsf_impl(Domain* d, double time, const mithra_seed* seed) {
  const mithra_grid& g = d->g;
  // the standing types are not handled by Seed::fields: its output stays the zero vector the
  // update starts from (src/classes.cpp:758-851, src/fdtd.cpp:239), so nothing is injected
  if (seed->seed_type >= MITHRA_SEED_STANDING_PLANEWAVE) return hipSuccess;
  if (seed->seed_type == MITHRA_SEED_SUPERGAUSSIANBEAM) return tfsf_sg_impl<R>(d, time, seed);
  const long n[3] = {(long)(g.N1 - 4) * (g.N2 - 4), (long)(g.N0 - 4) * (g.N2 - 4), (long)(g.N0 - 4) * (g.N1 - 4)};
  const double coef[3] = {d->co.a[1], d->co.a[2], d->co.a[3]};
  for (int axis = 0; axis < 3; ++axis) {
    if (n[axis] <= 0) continue;
    hipLaunchKernelGGL((k_tfsf<R>), dim3((unsigned)((n[axis] + 63) / 64)), dim3(256), 0, d->stream,
                       d->lvl[d->i_np1], axis, g.N0, g.N1, g.N2, g, *seed, time, coef[axis]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename R>
hipError_t boundaries_impl(Domain* d) {
  const mithra_grid& g = d->g;
  MurCoeffs c;
  for (int t = 0; t < 5; ++t) {
    c.bB[t] = d->co.bB[t]; c.cB[t] = d->co.cB[t]; c.dB[t] = d->co.dB[t];
    c.eE[t] = d->co.eE[t]; c.fE[t] = d->co.fE[t]; c.gE[t] = d->co.gE[t];
  }
  for (int t = 0; t < 17; ++t) c.hC[t] = d->co.hC[t];
  FieldSet P = d->lvl[d->i_np1], N = d->lvl[d->i_n], M = d->lvl[d->i_nm1];
  const int ymask = d->yfaces_pending;  // y faces the interior launch already updated
  d->yfaces_pending = 0;
  const long nf = mur_face_count(g.N0, g.N1, g.N2, ymask);
  hipLaunchKernelGGL((k_mur_faces<R>), dim3((unsigned)((nf + 255) / 256), d->ncomp), dim3(256), 0, d->stream, P, N, M,
                     d->ncomp, g.N0, g.N1, g.N2, c, ymask);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || g.truncation_order != 2) return e;
  const long ne = 4L * ((g.N2 - 2) + (g.N0 - 2) + (g.N1 - 2));
  hipLaunchKernelGGL((k_mur_edges<R>), dim3((unsigned)((ne + 255) / 256), d->ncomp), dim3(256), 0, d->stream, P, N, M,
                     d->ncomp, g.N0, g.N1, g.N2, c);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL((k_mur_corners<R>), dim3(1, d->ncomp), dim3(64), 0, d->stream, P, N, M, d->ncomp, g.N0, g.N1,
                     g.N2, c);
  return hipGetLastError();
}

template <typename R>
hipError_t eval_impl(Domain* d, int klo, int khi, const int* zr = nullptr, int pb0 = 0, int pb1 = -1) {
  klo = std::max(klo, 1);
  khi = std::min(khi, d->g.N2 - 2);
  if (klo > khi) return hipSuccess;
  return d->g.space_charge ? field_variant<R, true>(d, 1, d->tile_eval, klo, khi, zr, pb0, pb1)
                           : field_variant<R, false>(d, 1, d->tile_eval, klo, khi, zr, pb0, pb1);
}

}  // namespace

extern const int kTileShapes[kTileVariants][2] = {{4, 64}, {2, 64}, {4, 32}, {4, 128}, {8, 32},
                                                  {8, 64}, {16, 32}, {4, 64}, {8, 64}, {16, 64}};

hipError_t launch_interior(Domain* d, bool fuse_yfaces) {
  return d->precision == MITHRA_PRECISION_FP64 ? interior_impl<double>(d, fuse_yfaces)
                                               : interior_impl<float>(d, fuse_yfaces);
}

extern const int kFusedShapes[kFusedVariants][2] = {{4, 64}, {8, 64}, {16, 64}, {8, 128}};

hipError_t launch_stencil_eb(Domain* d) {
  if ((d->ncomp != 3 && d->ncomp != 4) || d->tile_fused < 0 || d->tile_fused >= kFusedVariants)
    return hipErrorInvalidValue;
  return d->precision == MITHRA_PRECISION_FP64 ? stencil_eb_impl<double>(d) : stencil_eb_impl<float>(d);
}

hipError_t launch_tfsf(Domain* d, double time, const mithra_seed* seed) {
  return d->precision == MITHRA_PRECISION_FP64 ? tfsf_impl<double>(d, time, seed) : tfsf_impl<float>(d, time, seed);
}

hipError_t launch_boundaries(Domain* d) {
  return d->precision == MITHRA_PRECISION_FP64 ? boundaries_impl<double>(d) : boundaries_impl<float>(d);
}

hipError_t launch_eval_eb(Domain* d) { return launch_eval_eb_range(d, 1, d->g.N2 - 2); }

hipError_t launch_eval_eb_step(Domain* d) {
  if (!d->lazy_fields || !d->zrange_valid) return launch_eval_eb(d);
  int pb0, pb1;
  lazy_planes(d, &pb0, &pb1);
  return d->precision == MITHRA_PRECISION_FP64 ? eval_impl<double>(d, 1, d->g.N2 - 2, d->zrange, pb0, pb1)
                                               : eval_impl<float>(d, 1, d->g.N2 - 2, d->zrange, pb0, pb1);
}

hipError_t launch_eval_eb_range(Domain* d, int klo, int khi) {
  return d->precision == MITHRA_PRECISION_FP64 ? eval_impl<double>(d, klo, khi) : eval_impl<float>(d, klo, khi);
}

hipError_t launch_fill_zero(Domain* d, void* ptr, size_t bytes) {
  // A fill kernel rather than hipMemsetAsync, so the step stays safe to capture in a graph.
  const size_t n16 = bytes / 16, tail = bytes % 16;
  if (n16) {
    unsigned blocks = (unsigned)((n16 + 255) / 256);
    if (blocks > 65536u) blocks = 65536u;
    hipLaunchKernelGGL(k_fill_zero, dim3(blocks), dim3(256), 0, d->stream, (uint4*)ptr, n16);
  }
  if (tail)
    hipLaunchKernelGGL(k_fill_zero_bytes, dim3(1), dim3(64), 0, d->stream, (unsigned char*)ptr + n16 * 16, tail);
  return hipGetLastError();
}


// ---- batched launchers --------------------------------------------------------------------
template <typename R, int TI, int K, bool LDS>
hipError_t batch_field(BatchView* V, BatchTiles T, int kernel, int nsfd, hipStream_t s) {
  // workgroups resident per CU, by shape: LDS forms are register- or LDS-bound, the others small
  const int per_cu = LDS ? (TI >= 16 ? 1 : TI >= 8 ? 2 : 4) : (TI >= 16 ? 2 : TI >= 8 ? 4 : 8);
  const int g = batch_grid(T.n, per_cu);
  if (kernel == 0) {
    if (nsfd) hipLaunchKernelGGL((kb_interior<R, true, TI, K, LDS>), dim3(g), dim3(kTileJ * TI), 0, s, V, T);
    else hipLaunchKernelGGL((kb_interior<R, false, TI, K, LDS>), dim3(g), dim3(kTileJ * TI), 0, s, V, T);
  } else {
    hipLaunchKernelGGL((kb_eval<R, TI, K, LDS>), dim3(g), dim3(kTileJ * TI), 0, s, V, T);
  }
  return hipGetLastError();
}

template <typename R>
hipError_t batch_field_variant(BatchView* V, BatchTiles T, int kernel, int variant, int nsfd, hipStream_t s) {
  if (T.n == 0) return hipSuccess;
  switch (variant) {
    case 0: return batch_field<R, 4, 64, false>(V, T, kernel, nsfd, s);
    case 1: return batch_field<R, 2, 64, false>(V, T, kernel, nsfd, s);
    case 2: return batch_field<R, 4, 32, false>(V, T, kernel, nsfd, s);
    case 3: return batch_field<R, 4, 128, false>(V, T, kernel, nsfd, s);
    case 4: return batch_field<R, 8, 32, false>(V, T, kernel, nsfd, s);
    case 5: return batch_field<R, 8, 64, false>(V, T, kernel, nsfd, s);
    case 6: return batch_field<R, 16, 32, false>(V, T, kernel, nsfd, s);
    case 7: return batch_field<R, 4, 64, true>(V, T, kernel, nsfd, s);
    case 8: return batch_field<R, 8, 64, true>(V, T, kernel, nsfd, s);
    case 9: return batch_field<R, 16, 64, true>(V, T, kernel, nsfd, s);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_batch_interior(BatchView* V, BatchTiles T, int variant, int precision, int nsfd, hipStream_t s) {
  return precision == MITHRA_PRECISION_FP64 ? batch_field_variant<double>(V, T, 0, variant, nsfd, s)
                                            : batch_field_variant<float>(V, T, 0, variant, nsfd, s);
}

hipError_t launch_batch_eval(BatchView* V, BatchTiles T, int variant, int precision, hipStream_t s) {
  return precision == MITHRA_PRECISION_FP64 ? batch_field_variant<double>(V, T, 1, variant, 0, s)
                                            : batch_field_variant<float>(V, T, 1, variant, 0, s);
}

hipError_t launch_batch_tfsf(BatchView* V, const BatchTiles* axes, int precision, hipStream_t s) {
  for (int axis = 0; axis < 3; ++axis) {
    const BatchTiles& T = axes[axis];
    if (!T.n) continue;
    if (precision == MITHRA_PRECISION_FP64)
      hipLaunchKernelGGL((kb_tfsf<double>), dim3(batch_grid(T.n, 8)), dim3(256), 0, s, V, T, axis);
    else
      hipLaunchKernelGGL((kb_tfsf<float>), dim3(batch_grid(T.n, 8)), dim3(256), 0, s, V, T, axis);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template <typename R>
hipError_t batch_mur_impl(BatchView* V, BatchTiles F, BatchTiles E, BatchTiles C, hipStream_t s) {
  if (F.n) hipLaunchKernelGGL((kb_mur<R, 0>), dim3(batch_grid(F.n, 8)), dim3(256), 0, s, V, F);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (E.n) hipLaunchKernelGGL((kb_mur<R, 1>), dim3(batch_grid(E.n, 8)), dim3(256), 0, s, V, E);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (C.n) hipLaunchKernelGGL((kb_mur<R, 2>), dim3(batch_grid(C.n, 8)), dim3(64), 0, s, V, C);
  return hipGetLastError();
}

hipError_t launch_batch_mur(BatchView* V, BatchTiles faces, BatchTiles edges, BatchTiles corners, int precision,
                            hipStream_t s) {
  return precision == MITHRA_PRECISION_FP64 ? batch_mur_impl<double>(V, faces, edges, corners, s)
                                            : batch_mur_impl<float>(V, faces, edges, corners, s);
}

}  // namespace mithra

namespace {

// Runs one of the analytic-seed kernels over n host points: inputs in (n_in arrays of per[i]
// doubles per point), one output of per_out doubles per point.
template <typename Launch>
int eval_points(long n, std::initializer_list<std::pair<const double*, int>> in, int per_out, double* out,
                Launch launch) {
  if (n < 0 || (n > 0 && !out)) return MITHRA_E_ARG;
  for (const auto& a : in)
    if (n > 0 && !a.first) return MITHRA_E_ARG;
  if (n == 0) return MITHRA_OK;
  std::vector<double*> dev;
  hipError_t e = hipSuccess;
  auto cleanup = [&]() {
    for (double* p : dev) (void)hipFree(p);
  };
  for (const auto& a : in) {
    double* p = nullptr;
    if ((e = hipMalloc(&p, sizeof(double) * a.second * n)) != hipSuccess) break;
    dev.push_back(p);
    if ((e = hipMemcpy(p, a.first, sizeof(double) * a.second * n, hipMemcpyHostToDevice)) != hipSuccess) break;
  }
  double* o = nullptr;
  if (e == hipSuccess && (e = hipMalloc(&o, sizeof(double) * per_out * n)) == hipSuccess) dev.push_back(o);
  if (e == hipSuccess) {
    launch(dev.data(), o, (unsigned)((n + 255) / 256));
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e == hipSuccess) e = hipMemcpy(out, o, sizeof(double) * per_out * n, hipMemcpyDeviceToHost);
  cleanup();
  if (e == hipErrorOutOfMemory) return MITHRA_E_NOMEM;
  return e == hipSuccess ? MITHRA_OK : MITHRA_E_HIP - (int)e;
}

bool signal_valid(const mithra_seed* s) { return s->signal_type >= MITHRA_SIGNAL_NEUMANN && s->signal_type <= MITHRA_SIGNAL_INVGAUSSIAN; }

}  // namespace

extern "C" {

int mithra_signal_evaluate(const mithra_seed* seed, const double* t, const double* phase, long n, double* out) {
  if (!seed || !signal_valid(seed)) return MITHRA_E_ARG;
  const mithra_seed s = *seed;
  return eval_points(n, {{t, 1}, {phase, 1}}, 1, out, [&](double** in, double* o, unsigned blocks) {
    hipLaunchKernelGGL(mithra::k_signal_eval, dim3(blocks), dim3(256), 0, 0, s, in[0], in[1], n, o);
  });
}

int mithra_seed_evaluate(const mithra_seed* seed, const double* r, const double* time, long n, double* a) {
  if (!seed || !signal_valid(seed)) return MITHRA_E_ARG;
  if (seed->seed_type < MITHRA_SEED_PLANEWAVE || seed->seed_type > MITHRA_SEED_STANDING_SUPERGAUSSIANBEAM)
    return MITHRA_E_ARG;
  if (n < 0 || (n > 0 && (!r || !time || !a))) return MITHRA_E_ARG;
  if (seed->seed_type >= MITHRA_SEED_STANDING_PLANEWAVE) {  // Seed::fields leaves the zero vector as it is
    for (long q = 0; q < 3 * n; ++q) a[q] = 0.0;
    return MITHRA_OK;
  }
  const mithra_seed s = *seed;
  if (s.seed_type == MITHRA_SEED_SUPERGAUSSIANBEAM) {
    // one output vector, zeroed once and passed to every point in turn, as the injection loop
    // passes it (src/fdtd.cpp:239, :307-373)
    if (n == 0) return MITHRA_OK;
    double* term = nullptr;
    int* reset = nullptr;
    if (hipMalloc(&term, sizeof(double) * n) != hipSuccess || hipMalloc(&reset, sizeof(int) * n) != hipSuccess) {
      (void)hipFree(term);
      return MITHRA_E_NOMEM;
    }
    const int rc = eval_points(n, {{r, 3}, {time, 1}}, 3, a, [&](double** in, double* o, unsigned blocks) {
      hipLaunchKernelGGL(mithra::k_sg_point_terms, dim3(blocks), dim3(256), 0, 0, s, in[0], in[1], n, term, reset);
      hipLaunchKernelGGL(mithra::k_sg_running, dim3(1), dim3(64), 0, 0, s, n, term, reset, o);
    });
    (void)hipFree(term);
    (void)hipFree(reset);
    return rc;
  }
  return eval_points(n, {{r, 3}, {time, 1}}, 3, a, [&](double** in, double* o, unsigned blocks) {
    hipLaunchKernelGGL(mithra::k_seed_eval, dim3(blocks), dim3(256), 0, 0, s, in[0], in[1], n, o);
  });
}

}  // extern "C"
