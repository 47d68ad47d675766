// Cell-row orders of the particle halves, and the physical reordering of the particle arrays.
//
// The current deposition (particles.hip) gives each workgroup the J values of one z-plane and a
// few x-rows of the mesh, and needs every particle half whose cell lies in the adjacent cell
// rows.  An "order" is the list of particle indices sorted by the row (k, i) of one half's cell,
// plus the start of each row in that list; a workgroup then reads two contiguous ranges per
// half.  The default order is a counting sort that walks the particles in the previous step's
// row order, where the new rows are nearly sorted (particles.hip: count, scan, scatter; the
// order within a row is arbitrary, as the deposit's LDS additions allow).  The fixed-order
// deposit keys by cell instead ((k N0 + i)(N1 - 1) + j) with rocPRIM's LSD radix sort of 32-bit
// (cell, index) pairs over the key's significant bits, which is stable, so the halves of one
// cell keep their particle-array order, and lists the start of every cell.
//
// The particle arrays themselves are permuted into row order now and then, so the sorted
// indices walk memory almost sequentially; the original index travels with each particle in
// `pid`, and the host always sees the records in load order.
#include "common.h"

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

namespace mithra {
namespace {

__global__ __launch_bounds__(256) void k_gather(const double* __restrict__ src, double* __restrict__ dst,
                                                const unsigned* __restrict__ id_src, unsigned* __restrict__ id_dst,
                                                unsigned* __restrict__ perm, long cap, long n) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const long s = perm[p];
#pragma unroll
  for (int a = 0; a < PCOUNT; ++a) dst[a * cap + p] = src[a * cap + s];
  id_dst[p] = id_src[s];
  perm[p] = (unsigned)p;  // the order now lists the reordered arrays in sequence
}

__global__ __launch_bounds__(256) void k_iota(unsigned* __restrict__ id, long n) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) id[p] = (unsigned)p;
}

// row_start[r] = number of sorted keys below r span (the first key of row r), for r in [0, rows]
__global__ __launch_bounds__(256) void k_row_start(const unsigned* __restrict__ key, long n, unsigned rows,
                                                   unsigned span, unsigned* __restrict__ row_start) {
  const long r = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > (long)rows) return;
  const unsigned long long first = (unsigned long long)r * span;
  long lo = 0, hi = n;
  while (lo < hi) {
    const long mid = (lo + hi) >> 1;
    if ((unsigned long long)key[mid] < first) lo = mid + 1;
    else hi = mid;
  }
  row_start[r] = (unsigned)lo;
}

// cell_start[c] = number of sorted cell keys below c, for c in [0, rows span]: a search within
// the keys of c's row only (row_start from k_row_start with the same span).
__global__ __launch_bounds__(256) void k_cell_start(const unsigned* __restrict__ key,
                                                    const unsigned* __restrict__ row_start, unsigned rows,
                                                    unsigned span, unsigned* __restrict__ cell_start) {
  const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long cells = (long)rows * span;
  if (c > cells) return;
  if (c == cells) {
    cell_start[c] = row_start[rows];
    return;
  }
  const long r = c / span;
  long lo = row_start[r], hi = row_start[r + 1];
  while (lo < hi) {
    const long mid = (lo + hi) >> 1;
    if ((long)key[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  cell_start[c] = (unsigned)lo;
}

}  // namespace

hipError_t launch_particle_ids(Domain* d) {
  if (d->np == 0) return hipSuccess;
  hipLaunchKernelGGL(k_iota, dim3((unsigned)((d->np + 255) / 256)), dim3(256), 0, d->stream, d->pid, d->np);
  return hipGetLastError();
}

void free_sort(Domain* d) {
  (void)hipFree(d->part_alt);
  (void)hipFree(d->pid_alt);
  (void)hipFree(d->sort_block);
  (void)hipFree(d->sort_tmp);
  d->part_alt = nullptr;
  d->pid_alt = nullptr;
  d->sort_block = nullptr;
  d->sort_tmp = nullptr;
  d->sort_tmp_bytes = 0;
  d->sort_cap = 0;
  for (int s = 0; s < 2; ++s) {
    d->ord[s].key = d->ord[s].idx = d->ord[s].row = d->ord[s].cell = nullptr;
  }
  d->key_alt = d->idx_alt = d->row_cnt = nullptr;
  d->rm_order_valid = 0;
  d->order1_is_rnp = 0;
}

hipError_t ensure_orders(Domain* d, unsigned rows) {
  if (d->sort_cap == d->cap && d->sort_rows == rows) return hipSuccess;
  free_sort(d);
  const size_t cap = (size_t)d->cap;
  const size_t per = (cap + 63) / 64 * 64, rper = ((size_t)rows + 1 + 63) / 64 * 64;
  hipError_t e = hipMalloc(&d->sort_block, sizeof(unsigned) * (6 * per + 3 * rper));
  if (e != hipSuccess) return e;
  unsigned* b = d->sort_block;
  for (int s = 0; s < 2; ++s) {
    d->ord[s].key = b; b += per;
    d->ord[s].idx = b; b += per;
    d->ord[s].row = b; b += rper;
  }
  d->key_alt = b; b += per;
  d->idx_alt = b; b += per;
  d->row_cnt = b;
  d->sort_cap = d->cap;
  d->sort_rows = rows;
  return launch_fill_zero(d, d->row_cnt, sizeof(unsigned) * rper);
}

// Sorts the keys of order o (key units: rows, or cells when span > 1 with `rows` * span cells)
// and lists the start of each row into o.row and, keyed by cell, of each cell into o.cell.
hipError_t sort_order(Domain* d, CellOrder& o, long n, unsigned rows, unsigned span) {
  const unsigned long long units = (unsigned long long)rows * span;
  if (units >= 0xffffffffull || (span > 1 && !o.cell)) return hipErrorInvalidValue;
  int bits = 1;
  while (bits < 32 && (units >> bits) != 0) ++bits;  // the sentinel `units` must fit
  rocprim::double_buffer<unsigned> k(o.key, d->key_alt), v(o.idx, d->idx_alt);
  size_t bytes = 0;
  hipError_t e = rocprim::radix_sort_pairs(nullptr, bytes, k, v, (unsigned)n, 0, bits, d->stream);
  if (e != hipSuccess) return e;
  if (bytes > d->sort_tmp_bytes || !d->sort_tmp) {
    (void)hipFree(d->sort_tmp);
    d->sort_tmp = nullptr;
    d->sort_tmp_bytes = 0;
    if ((e = hipMalloc(&d->sort_tmp, bytes ? bytes : 1)) != hipSuccess) return e;
    d->sort_tmp_bytes = bytes;
  }
  bytes = d->sort_tmp_bytes;
  if ((e = rocprim::radix_sort_pairs(d->sort_tmp, bytes, k, v, (unsigned)n, 0, bits, d->stream)) != hipSuccess)
    return e;
  if (k.current() != o.key) {  // keep the sorted pairs in the order's own arrays
    d->key_alt = o.key;
    o.key = k.current();
  }
  if (v.current() != o.idx) {
    d->idx_alt = o.idx;
    o.idx = v.current();
  }
  hipLaunchKernelGGL(k_row_start, dim3((unsigned)(((long)rows + 1 + 255) / 256)), dim3(256), 0, d->stream, o.key, n,
                     rows, span, o.row);
  if (span > 1)
    hipLaunchKernelGGL(k_cell_start, dim3((unsigned)((units + 1 + 255) / 256)), dim3(256), 0, d->stream, o.key, o.row,
                       rows, span, o.cell);
  return hipGetLastError();
}

// row[r] = cnt[0] + ... + cnt[r-1] for r in [0, rows]: the start of every row of a counting
// order (cnt has rows + 1 entries, the last one the bucket of halves that deposit nothing).
hipError_t scan_row_counts(Domain* d, const unsigned* cnt, unsigned* row, unsigned rows) {
  size_t bytes = 0;
  hipError_t e = rocprim::exclusive_scan(nullptr, bytes, cnt, row, 0u, (size_t)rows + 1, rocprim::plus<unsigned>(),
                                         d->stream);
  if (e != hipSuccess) return e;
  if (bytes > d->sort_tmp_bytes || !d->sort_tmp) {
    (void)hipFree(d->sort_tmp);
    d->sort_tmp = nullptr;
    d->sort_tmp_bytes = 0;
    if ((e = hipMalloc(&d->sort_tmp, bytes ? bytes : 1)) != hipSuccess) return e;
    d->sort_tmp_bytes = bytes;
  }
  bytes = d->sort_tmp_bytes;
  return rocprim::exclusive_scan(d->sort_tmp, bytes, cnt, row, 0u, (size_t)rows + 1, rocprim::plus<unsigned>(),
                                 d->stream);
}

hipError_t reorder_particles(Domain* d, CellOrder& o) {
  const long n = d->np;
  if (n < 2) return hipSuccess;
  if (!d->part_alt) {
    hipError_t e = hipMalloc(&d->part_alt, sizeof(double) * PCOUNT * (size_t)d->cap);
    if (e == hipSuccess) e = hipMalloc(&d->pid_alt, sizeof(unsigned) * (size_t)d->cap);
    if (e != hipSuccess) {
      (void)hipFree(d->part_alt);
      d->part_alt = nullptr;
      d->pid_alt = nullptr;
      return e;
    }
  }
  hipLaunchKernelGGL(k_gather, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, d->stream, d->part, d->part_alt,
                     d->pid, d->pid_alt, o.idx, d->cap, n);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  double* t = d->part;
  d->part = d->part_alt;
  d->part_alt = t;
  unsigned* ti = d->pid;
  d->pid = d->pid_alt;
  d->pid_alt = ti;
  return hipSuccess;
}

}  // namespace mithra
