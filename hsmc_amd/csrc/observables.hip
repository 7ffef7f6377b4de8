// observables.hip -- cell list, overlap tests, pair histograms and Widom
// insertions on one replica's configuration.
//
// Every neighbour search runs on an "observable grid": an unshifted cell grid
// whose side is at least the cutoff of the question asked (1 for overlaps and
// Widom, 1/sf for a compression by sf, rmax for the histograms), stored in
// CSR form (cell start offsets, particle ids and FP64 positions sorted by
// cell).  Because the side covers the cutoff, the set of pairs examined is
// every pair within the cutoff, which is also the set the reference examines
// through its own cell list whenever its cells are at least that wide
// (compute_press.c:61-69 enforces it for the contact histogram).
#include <math.h>

#include <algorithm>
#include <vector>

#include "ctx.h"
#include "device_common.h"

using namespace hsmc;

namespace {

enum Slot { S_CNT = 0, S_START, S_CELLOF, S_IDS, S_PX, S_PY, S_PZ, S_FILL, S_Q0, S_Q1, S_Q2, S_OUT0, S_OUT1,
            S_PART };

struct Grid {
  int nx, ny, nz;
  double sx, sy, sz;
  double lx, ly, lz, hx, hy, hz;
  const int *start;   // C + 1
  const int *ids;     // N, particle id by sorted position
  const double *px, *py, *pz;
};

__global__ void k_grid_count(const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, int n, int nx, int ny, int nz, double sx,
                             double sy, double sz, int *__restrict__ cnt, int *__restrict__ cellof) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = (plain_cell(x[i], sx, nx) * ny + plain_cell(y[i], sy, ny)) * nz + plain_cell(z[i], sz, nz);
  cellof[i] = c;
  atomicAdd(&cnt[c], 1);
}

// Exclusive prefix sum of cnt[0..n) into start[0..n], one workgroup.  Each
// pass covers 32 * 1024 entries: a thread loads its 32 consecutive entries
// at once (all loads in flight together), the block scans the thread sums,
// and a carry links the passes.
__global__ void __launch_bounds__(1024) k_scan1(const int *__restrict__ cnt, int n, int *__restrict__ start) {
  constexpr int K = 32;
  __shared__ int part[1024];
  const int t = threadIdx.x;
  int carry = 0;
  for (int base = 0; base < n; base += K * 1024) {
    const int b = base + t * K;
    int v[K];
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = (b + k < n) ? cnt[b + k] : 0;
    int s = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) s += v[k];
    part[t] = s;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
      const int u = (t >= d) ? part[t - d] : 0;
      __syncthreads();
      part[t] += u;
      __syncthreads();
    }
    int run = carry + part[t] - s;
    const int total = part[1023];
    __syncthreads();  // part is rewritten by the next pass
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (b + k < n) start[b + k] = run;
      run += v[k];
    }
    carry += total;
  }
  if (t == 0) start[n] = carry;
}

// Exclusive prefix sum over many workgroups: each 256-thread workgroup scans
// a tile of 4096 entries (16 consecutive per thread) and writes the tile
// total; the totals are scanned by k_scan1; k_scan_add adds each tile's
// offset.  Three short launches in place of one workgroup walking the whole
// array.
constexpr int kScanTile = 4096;

__global__ void __launch_bounds__(256) k_scan_tile(const int *__restrict__ in, int n,
                                                   int *__restrict__ out, int *__restrict__ part) {
  __shared__ int wsum[4];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const long base = (long)blockIdx.x * kScanTile + (long)t * 16;
  int v[16];
  int s = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0;
    s += v[k];
  }
  int incl = s;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int u = __shfl_up(incl, d, 64);
    if (lane >= d) incl += u;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int woff = 0;
  for (int i = 0; i < w; ++i) woff += wsum[i];
  int run = woff + incl - s;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (t == 255) part[blockIdx.x] = woff + incl;
}

__global__ void k_scan_empty(int *__restrict__ out) {
  if (threadIdx.x == 0) out[0] = 0;
}

__global__ void __launch_bounds__(256) k_scan_add(int *__restrict__ out, int n,
                                                  const int *__restrict__ offs, int tiles) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] += offs[i / kScanTile];
  if (i == 0) out[n] = offs[tiles];
}

__global__ void k_grid_scatter(const double *__restrict__ x, const double *__restrict__ y,
                               const double *__restrict__ z, int n, const int *__restrict__ cellof,
                               const int *__restrict__ start, int *__restrict__ fill,
                               int *__restrict__ ids, double *__restrict__ px,
                               double *__restrict__ py, double *__restrict__ pz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cellof[i];
  const int pos = start[c] + atomicAdd(&fill[c], 1);
  ids[pos] = i;
  px[pos] = x[i];
  py[pos] = y[i];
  pz[pos] = z[i];
}

// One thread per cell: ascending ids, then the reference's part_cell row
// [count, ids..., -1...] (cell_list.c:148-171).
__global__ void k_part_cell(const int *__restrict__ start, int C, int *__restrict__ ids,
                            int max_part, int *__restrict__ out, unsigned *__restrict__ err) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= C) return;
  const int b = start[c], n = start[c + 1] - start[c];
  int *row = out + (size_t)c * max_part;
  if (n > max_part - 1) {  // cell_list_check (cell_list.c:229-232)
    atomicOr(err, kErrOverflow);
    row[0] = n;
    return;
  }
  for (int i = 1; i < n; ++i) {  // insertion sort of a few ids
    const int v = ids[b + i];
    int j = i - 1;
    while (j >= 0 && ids[b + j] > v) {
      ids[b + j + 1] = ids[b + j];
      --j;
    }
    ids[b + j + 1] = v;
  }
  row[0] = n;
  for (int j = 1; j < max_part; ++j) row[j] = (j <= n) ? ids[b + j - 1] : -1;
}

// Iterate the distinct cells of the 27-stencil around (cx,cy,cz).
template <class F>
__device__ __forceinline__ bool for_stencil(const Grid &g, int cx, int cy, int cz, F &&f) {
  for (int dx = -1; dx <= 1; ++dx) {
    if (!stencil_axis_ok(dx, g.nx)) continue;
    const int ax = wrap_index(cx + dx, g.nx);
    for (int dy = -1; dy <= 1; ++dy) {
      if (!stencil_axis_ok(dy, g.ny)) continue;
      const int ay = wrap_index(cy + dy, g.ny);
      for (int dz = -1; dz <= 1; ++dz) {
        if (!stencil_axis_ok(dz, g.nz)) continue;
        const int c = (ax * g.ny + ay) * g.nz + wrap_index(cz + dz, g.nz);
        if (f(g.start[c], g.start[c + 1])) return true;
      }
    }
  }
  return false;
}

// Any sphere other than skip_id closer than 1 to q, in the 27 cells around
// q (check_overlap, moves.c:157-212, and widom_check_overlap,
// compute_widom_chem_pot.c:82-120).  The answer does not depend on the
// order of the cells, so they are visited nearest first (the point's own
// cell, then the 6 sharing a face, then the rest), with the ranges of the
// first 7 loaded together: a point that overlaps usually does so there, and
// leaves after one round trip for the ranges and one for the spheres.
__device__ __forceinline__ bool point_overlaps(const Grid &g, double qx, double qy, double qz,
                                               int skip_id) {
  const int cx = plain_cell(qx, g.sx, g.nx), cy = plain_cell(qy, g.sy, g.ny),
            cz = plain_cell(qz, g.sz, g.nz);
  auto cell_of = [&](int j, bool &ok) {
    int dx, dy, dz;
    stencil_by_distance(j, dx, dy, dz);
    ok = stencil_axis_ok(dx, g.nx) && stencil_axis_ok(dy, g.ny) && stencil_axis_ok(dz, g.nz);
    return ok ? (wrap_index(cx + dx, g.nx) * g.ny + wrap_index(cy + dy, g.ny)) * g.nz +
                    wrap_index(cz + dz, g.nz)
              : 0;
  };
  auto scan = [&](int b, int e) {
    for (int p = b; p < e; ++p) {
      // the id and the position load together; the skip is tested after
      const int id = g.ids[p];
      const double d2 =
          dist2(qx, qy, qz, g.px[p], g.py[p], g.pz[p], g.lx, g.ly, g.lz, g.hx, g.hy, g.hz);
      if (id != skip_id && d2 < 1.0) return true;
    }
    return false;
  };
  int c7[7], b7[7], e7[7];
  bool ok7[7];
#pragma unroll
  for (int j = 0; j < 7; ++j) c7[j] = cell_of(j, ok7[j]);
#pragma unroll
  for (int j = 0; j < 7; ++j) b7[j] = g.start[c7[j]];
#pragma unroll
  for (int j = 0; j < 7; ++j) e7[j] = g.start[c7[j] + 1];
#pragma unroll
  for (int j = 0; j < 7; ++j)
    if (!ok7[j]) e7[j] = b7[j];  // a cell the stencil leaves out (tiny grids): empty
#pragma unroll
  for (int j = 0; j < 7; ++j)
    if (scan(b7[j], e7[j])) return true;
  for (int j = 7; j < 27; ++j) {
    bool ok;
    const int c = cell_of(j, ok);
    if (ok && scan(g.start[c], g.start[c + 1])) return true;
  }
  return false;
}

__global__ void k_overlap_batch(Grid g, const int *__restrict__ idx, const double *__restrict__ q,
                                int m, unsigned char *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  out[k] = point_overlaps(g, q[3 * k], q[3 * k + 1], q[3 * k + 2], idx[k]) ? 1 : 0;
}

__global__ void k_widom_points(Grid g, const double *__restrict__ q, int m,
                               unsigned char *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  out[k] = point_overlaps(g, q[3 * k], q[3 * k + 1], q[3 * k + 2], -1) ? 0 : 1;
}

__device__ __forceinline__ unsigned long long block_sum_u64(unsigned long long v) {
  __shared__ unsigned long long red[16];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  unsigned long long s = 0;
  if (threadIdx.x == 0) {
    for (int i = 0; i < (int)((blockDim.x + 63) >> 6); ++i) s += red[i];
  }
  return s;
}

// widom_insertion (compute_widom_chem_pot.c:44-71) on a Philox stream:
// point i has counter {i_lo, i_hi, sample_lo, 0x5749xxxx}, coordinates
// u/(2^32-1) * L (widom_rand_pos, :73-80).
__global__ void __launch_bounds__(256)
k_widom_uniform(Grid g, uint32_t k0, uint32_t k1, unsigned long long sample, long long first,
                long long n, unsigned long long *__restrict__ hits) {
  unsigned long long h = 0;
  const long long stride = (long long)gridDim.x * blockDim.x;
  const uint32_t c3 = 0x57490000u | (uint32_t)((sample >> 32) & 0xFFFFu);
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) {
    const unsigned long long i = (unsigned long long)(first + t);
    uint32_t u[4];
    philox4x32_10((uint32_t)i, (uint32_t)(i >> 32), (uint32_t)sample, c3, k0, k1, u);
    const double qx = u32_unit_closed(u[0]) * g.lx;
    const double qy = u32_unit_closed(u[1]) * g.ly;
    const double qz = u32_unit_closed(u[2]) * g.lz;
    if (!point_overlaps(g, qx, qy, qz, -1)) ++h;
  }
  const unsigned long long s = block_sum_u64(h);
  if (threadIdx.x == 0 && s) atomicAdd(hits, s);
}

// Cell-stratified uniform insertion: every cell of the observable grid
// (side >= 1) receives `per_cell` points uniform inside it.  The cells have
// equal volume, so hits / (per_cell * C) is an unbiased estimate of the
// insertion probability, like widom_insertion's uniform points
// (compute_widom_chem_pot.c:44-80), with a variance no larger.  Point j of
// cell c comes from Philox counter {c, j, sample_lo, 0x5743xxxx}:
// coordinate (ic + (u + 0.5) 2^-32) * side per axis.
//
// One wave per cell: lanes j < 27 read the neighbour cells' ranges, and the
// spheres of the 27 cells are staged once in LDS in the order own cell, 6
// faces, 12 edges, 8 corners (the likeliest overlaps first).  The wave then
// takes 64 points at a time, one per lane, and walks the list in lockstep
// (every lane reads entry e together: an LDS broadcast) with the
// reference's FP64 distance (moves.c:400-431 order, minimum image), leaving
// the walk as soon as a ballot shows every point overlapping.
constexpr int kWidomCellsLcap = 256;   // bound of spheres in a 27-cell stencil (host-checked)
constexpr int kWidomCellsWaves = 4;    // waves per workgroup

__global__ void __launch_bounds__(64 * kWidomCellsWaves)
k_widom_cells(Grid g, uint32_t k0, uint32_t k1, unsigned long long sample, int per_cell,
              unsigned long long *__restrict__ hits, unsigned *__restrict__ err) {
  __shared__ double sx_[kWidomCellsWaves][kWidomCellsLcap];
  __shared__ double sy_[kWidomCellsWaves][kWidomCellsLcap];
  __shared__ double sz_[kWidomCellsWaves][kWidomCellsLcap];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double *Lx = sx_[w], *Ly = sy_[w], *Lz = sz_[w];
  const long C = (long)g.nx * g.ny * g.nz;
  const uint32_t c3 = 0x57430000u | (uint32_t)((sample >> 32) & 0xFFFFu);
  unsigned long long h = 0;
  for (long c = (long)blockIdx.x * kWidomCellsWaves + w; c < C; c += (long)gridDim.x * kWidomCellsWaves) {
    const int iz = (int)(c % g.nz);
    const long t2 = c / g.nz;
    const int iy = (int)(t2 % g.ny), ix = (int)(t2 / g.ny);
    // A cell with every index in [1, g - 2] and its neighbours hold positions
    // and points within 2 sides of each other, under half the box when
    // g >= 8: compute_dist's minimum image (moves.c:414-427) changes no
    // difference there and is left out (the same FP64 operations remain).
    // The cell is the wave's, so the choice is uniform.
    const bool inner = g.nx >= 8 && g.ny >= 8 && g.nz >= 8 && ix >= 1 && ix <= g.nx - 2 &&
                       iy >= 1 && iy <= g.ny - 2 && iz >= 1 && iz <= g.nz - 2;
    // neighbour ranges (distinct cells only on axes of fewer than 3 cells)
    int b = 0, n = 0;
    if (lane < 27) {
      int dx = 0, dy = 0, dz = 0;
      stencil_by_distance(lane, dx, dy, dz);
      if (stencil_axis_ok(dx, g.nx) && stencil_axis_ok(dy, g.ny) && stencil_axis_ok(dz, g.nz)) {
        const int nc = (wrap_index(ix + dx, g.nx) * g.ny + wrap_index(iy + dy, g.ny)) * g.nz +
                       wrap_index(iz + dz, g.nz);
        b = g.start[nc];
        n = g.start[nc + 1] - b;
      }
    }
    int incl = n;
    for (int d = 1; d < 32; d <<= 1) {
      const int v = __shfl_up(incl, d, 64);
      if (lane >= d) incl += v;
    }
    // the list length, read into a scalar register so that the point loop
    // below is a uniform loop; padded to a multiple of 4 with positions no
    // point can overlap (the loop tests 4 entries between its exit checks)
    const int T = __builtin_amdgcn_readfirstlane(__shfl(incl, 26, 64));
    const int T4 = (T + 3) & ~3;
    if (T4 > kWidomCellsLcap) {  // above the packing bound the host checks: cannot happen
      if (lane == 0) atomicOr(err, kErrListFull);
      continue;
    }
    for (int q = 0, o = incl - n; q < n; ++q) {
      Lx[o + q] = g.px[b + q];
      Ly[o + q] = g.py[b + q];
      Lz[o + q] = g.pz[b + q];
    }
    if (lane < T4 - T) {
      Lx[T + lane] = 1e300;
      Ly[T + lane] = 1e300;
      Lz[T + lane] = 1e300;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // 64 points at a time: point base + lane of the cell
    for (int base = 0; base < per_cell; base += 64) {
      const int jj = base + lane;
      const bool mine = jj < per_cell;
      uint32_t u[4];
      philox4x32_10((uint32_t)c, (uint32_t)jj, (uint32_t)sample, c3, k0, k1, u);
      const double qx = ((double)ix + ((double)u[0] + 0.5) * 0x1p-32) * g.sx;
      const double qy = ((double)iy + ((double)u[1] + 0.5) * 0x1p-32) * g.sy;
      const double qz = ((double)iz + ((double)u[2] + 0.5) * 0x1p-32) * g.sz;
      // the smallest squared distance so far: a point overlaps a sphere iff
      // it is below 1 (one v_min per pair in place of a compare and a merge)
      double m = 1e300;
      if (inner) {
        for (int e = 0; e < T4; e += 4) {
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const double dx = qx - Lx[e + i], dy = qy - Ly[e + i], dz = qz - Lz[e + i];
            m = fmin(m, dx * dx + dy * dy + dz * dz);
          }
          if (__all(m < 1.0 || !mine)) break;
        }
      } else {
        for (int e = 0; e < T4; e += 4) {
#pragma unroll 1
          for (int i = 0; i < 4; ++i)
            m = fmin(m, dist2(qx, qy, qz, Lx[e + i], Ly[e + i], Lz[e + i], g.lx, g.ly, g.lz, g.hx,
                              g.hy, g.hz));
          if (__all(m < 1.0 || !mine)) break;
        }
      }
      const bool ov = m < 1.0;
      if (mine && !ov) ++h;
    }
    __builtin_amdgcn_wave_barrier();  // the list is rewritten for the next cell
  }
  for (int o = 32; o > 0; o >>= 1) h += __shfl_xor(h, o, 64);
  if (lane == 0 && h) atomicAdd(hits, h);
}

// check_overlap(idx, sf, sf, sf) (moves.c:157-212) for each query.
__global__ void k_overlap_scaled(Grid g, const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ z, const int *__restrict__ idx,
                                 const double *__restrict__ sfv, int m,
                                 unsigned char *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int i = idx[k];
  const double sf = sfv[k];
  const double ax = x[i], ay = y[i], az = z[i];
  const int cx = plain_cell(ax, g.sx, g.nx), cy = plain_cell(ay, g.sy, g.ny), cz = plain_cell(az, g.sz, g.nz);
  const bool ov = for_stencil(g, cx, cy, cz, [&](int b, int e) {
    for (int p = b; p < e; ++p) {
      if (g.ids[p] == i) continue;
      if (dist2_scaled(ax, ay, az, g.px[p], g.py[p], g.pz[p], g.lx, g.ly, g.lz, sf) < 1.0)
        return true;
    }
    return false;
  });
  out[k] = ov ? 1 : 0;
}

// Any pair overlapping after scaling by sf[k]: bit k of *mask.  Pairs whose
// unscaled squared distance exceeds cut2 cannot overlap for any sf[k].
__global__ void __launch_bounds__(256)
k_any_overlap_scaled(Grid g, int n, const double *__restrict__ sfv, int nsf, double cut2,
                     unsigned long long *__restrict__ mask) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double ax = g.px[p], ay = g.py[p], az = g.pz[p];
  const int cx = plain_cell(ax, g.sx, g.nx), cy = plain_cell(ay, g.sy, g.ny), cz = plain_cell(az, g.sz, g.nz);
  unsigned long long m = 0;
  for_stencil(g, cx, cy, cz, [&](int b, int e) {
    for (int q = b; q < e; ++q) {
      if (q <= p) continue;  // each unordered pair once
      const double s0 = dist2(ax, ay, az, g.px[q], g.py[q], g.pz[q], g.lx, g.ly, g.lz, g.hx, g.hy, g.hz);
      if (s0 >= cut2) continue;
      for (int k = 0; k < nsf; ++k)
        if (dist2_scaled(ax, ay, az, g.px[q], g.py[q], g.pz[q], g.lx, g.ly, g.lz, sfv[k]) < 1.0)
          m |= (1ull << k);
    }
    return false;
  });
  if (m) atomicOr(mask, m);
}

__global__ void k_pair_dist(const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, const int *__restrict__ ii,
                            const int *__restrict__ jj, int m, double sf, double lx, double ly,
                            double lz, double *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int i = ii[k], j = jj[k];
  out[k] = sqrt(dist2_scaled(x[i], y[i], z[i], x[j], y[j], z[j], lx, ly, lz, sf));
}

// Pair histogram: every unordered pair with dr < rmax_eff, bin
// (int)((dr - 1.0) / dr_bin) (compute_rdf.c:118-124, compute_press.c:150-155).
__global__ void __launch_bounds__(256)
k_pair_hist(Grid g, int n, double dr_bin, int nbins, double rmax_eff, double cut2,
            unsigned long long *__restrict__ counts, unsigned long long *__restrict__ oor) {
  extern __shared__ unsigned int hist[];
  for (int b = threadIdx.x; b < nbins; b += blockDim.x) hist[b] = 0;
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long bad = 0;
  if (p < n) {
    const double ax = g.px[p], ay = g.py[p], az = g.pz[p];
    const int cx = plain_cell(ax, g.sx, g.nx), cy = plain_cell(ay, g.sy, g.ny), cz = plain_cell(az, g.sz, g.nz);
    for_stencil(g, cx, cy, cz, [&](int b, int e) {
      for (int q = max(b, p + 1); q < e; ++q) {
        const double s = dist2(ax, ay, az, g.px[q], g.py[q], g.pz[q], g.lx, g.ly, g.lz, g.hx, g.hy, g.hz);
        if (s >= cut2) continue;
        const double dr = sqrt(s);
        if (dr < rmax_eff) {
          const int bin = (int)((dr - 1.0) / dr_bin);
          if (bin >= 0 && bin < nbins)
            atomicAdd(&hist[bin], 1u);
          else
            ++bad;
        }
      }
      return false;
    });
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbins; b += blockDim.x)
    if (hist[b]) atomicAdd(&counts[b], (unsigned long long)hist[b]);
  const unsigned long long s = block_sum_u64(bad);
  if (threadIdx.x == 0 && s) atomicAdd(oor, s);
}

// The same pair histogram on cell pairs, for grids whose cells hold many
// spheres (the g(r) cutoff of several diameters), where one thread per
// sphere leaves most of the GPU without waves.  One wave per (cell A,
// stencil cell B) with B at or after A in cell order: B's sorted positions
// are staged in LDS a tile at a time and each lane tests its sphere of A
// against the tile.  Positions are sorted by cell, so "q after p in sorted
// order" is the per-sphere kernel's rule and every pair is counted once,
// with the same arithmetic (reference sphere first).
constexpr int kPairTile = 256;

size_t pair_tiled_lds(int nbins) {
  return ((sizeof(unsigned int) * (size_t)nbins + 7) & ~(size_t)7) + 3 * sizeof(double) * kPairTile;
}

// One lane's sphere (px, py, pz) against tile entries [j0, nq): the per-pair
// arithmetic of k_pair_hist; MI = false leaves out the minimum image where
// it cannot change a difference.
template <bool MI>
__device__ __forceinline__ void pair_tile(const Grid &g, double px, double py, double pz,
                                          const double *tx, const double *ty, const double *tz,
                                          int j0, int nq, double cut2, double rmax_eff,
                                          double dr_bin, int nbins, unsigned int *hist,
                                          unsigned long long &bad) {
  for (int j = j0; j < nq; ++j) {
    double s;
    if constexpr (MI) {
      s = dist2(px, py, pz, tx[j], ty[j], tz[j], g.lx, g.ly, g.lz, g.hx, g.hy, g.hz);
    } else {
      const double dx = px - tx[j], dy = py - ty[j], dz = pz - tz[j];
      s = dx * dx + dy * dy + dz * dz;
    }
    if (s >= cut2) continue;
    const double dr = sqrt(s);
    if (dr < rmax_eff) {
      const int bin = (int)((dr - 1.0) / dr_bin);
      if (bin >= 0 && bin < nbins)
        atomicAdd(&hist[bin], 1u);
      else
        ++bad;
    }
  }
}

__global__ void __launch_bounds__(64)
k_pair_hist_tiled(Grid g, int nbins, double dr_bin, double rmax_eff, double cut2,
                  unsigned long long *__restrict__ counts, unsigned long long *__restrict__ oor) {
  extern __shared__ unsigned long long smem_u64[];
  unsigned int *hist = (unsigned int *)smem_u64;
  double *tx = (double *)((unsigned char *)smem_u64 + ((sizeof(unsigned int) * (size_t)nbins + 7) & ~(size_t)7));
  double *ty = tx + kPairTile, *tz = ty + kPairTile;
  const int lane = threadIdx.x;
  const int cA = (int)(blockIdx.x / 27u), st = (int)(blockIdx.x % 27u);
  const int dx = st / 9 - 1, dy = (st / 3) % 3 - 1, dz = st % 3 - 1;
  if (!stencil_axis_ok(dx, g.nx) || !stencil_axis_ok(dy, g.ny) || !stencil_axis_ok(dz, g.nz)) return;
  const int ax = cA / (g.ny * g.nz), ay = (cA / g.nz) % g.ny, az = cA % g.nz;
  const int cB = (wrap_index(ax + dx, g.nx) * g.ny + wrap_index(ay + dy, g.ny)) * g.nz +
                 wrap_index(az + dz, g.nz);
  if (cB < cA) return;  // counted from B's side
  const int bA = g.start[cA], eA = g.start[cA + 1], bB = g.start[cB], eB = g.start[cB + 1];
  if (bA == eA || bB == eB) return;
  // No wrap between A and B on any axis, and at least 5 cells per axis: every
  // coordinate difference is within 2 sides (plus rounding of the cell
  // index) <= 0.4 L, below half a box, where the minimum image returns it
  // unchanged, so the tile skips it (the same doubles either way).
  const bool direct = g.nx >= 5 && g.ny >= 5 && g.nz >= 5 && ax + dx >= 0 && ax + dx < g.nx &&
                      ay + dy >= 0 && ay + dy < g.ny && az + dz >= 0 && az + dz < g.nz;
  for (int b = lane; b < nbins; b += 64) hist[b] = 0;
  unsigned long long bad = 0;
  for (int p0 = bA; p0 < eA; p0 += 64) {
    const int p = p0 + lane;
    const bool act = p < eA;
    const double px = act ? g.px[p] : 0.0, py = act ? g.py[p] : 0.0, pz = act ? g.pz[p] : 0.0;
    for (int q0 = bB; q0 < eB; q0 += kPairTile) {
      const int nq = min(kPairTile, eB - q0);
      __syncthreads();  // the previous tile is no longer read
      for (int j = lane; j < nq; j += 64) {
        tx[j] = g.px[q0 + j];
        ty[j] = g.py[q0 + j];
        tz[j] = g.pz[q0 + j];
      }
      __syncthreads();
      if (!act) continue;
      // in A's own cell only the spheres after p; in a later cell all
      const int j0 = max(0, p + 1 - q0);
      if (direct)
        pair_tile<false>(g, px, py, pz, tx, ty, tz, j0, nq, cut2, rmax_eff, dr_bin, nbins, hist, bad);
      else
        pair_tile<true>(g, px, py, pz, tx, ty, tz, j0, nq, cut2, rmax_eff, dr_bin, nbins, hist, bad);
    }
  }
  __syncthreads();
  for (int b = lane; b < nbins; b += 64)
    if (hist[b]) atomicAdd(&counts[b], (unsigned long long)hist[b]);
  for (int o = 32; o > 0; o >>= 1) bad += __shfl_down(bad, o, 64);
  if (lane == 0 && bad) atomicAdd(oor, bad);
}

}  // namespace

namespace hsmc {

long scan_workspace_ints(long n) {
  const long tiles = (n + kScanTile - 1) / kScanTile;
  return 2 * tiles + 1;
}

void exclusive_scan(hipStream_t s, const int *in, int n, int *out, int *part) {
  const int tiles = (int)((n + kScanTile - 1) / kScanTile);
  if (tiles == 0) {
    hipLaunchKernelGGL(k_scan_empty, dim3(1), dim3(64), 0, s, out);
    return;
  }
  int *offs = part + tiles;  // tiles + 1 entries: exclusive offsets and the total
  hipLaunchKernelGGL(k_scan_tile, dim3((unsigned)tiles), dim3(256), 0, s, in, n, out, part);
  hipLaunchKernelGGL(k_scan1, dim3(1), dim3(1024), 0, s, part, tiles, offs);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, n, offs, tiles);
}

}  // namespace hsmc

namespace {

int grid_cells_for(double L, double cut) {
  int n = (int)floor(L / cut);
  if (n < 1) n = 1;
  while (n > 1 && L / n < cut) --n;
  return n;
}

// Build (or reuse) the observable grid of `replica` for cutoff `cut`.
int build_grid(hsmc_gpu_ctx *ctx, int replica, double cut, Grid &g) {
  const RepHost &h = ctx->reps[replica];
  const int N = h.p.n_part;
  const int nx = grid_cells_for(h.p.lx, cut), ny = grid_cells_for(h.p.ly, cut),
            nz = grid_cells_for(h.p.lz, cut);
  const long C = (long)nx * ny * nz;
  if (C > (1L << 28)) return HSMC_ERR_NOT_SUPPORTED;
  int *cnt = (int *)scratch(ctx, S_CNT, sizeof(int) * C);
  int *fill = (int *)scratch(ctx, S_FILL, sizeof(int) * C);
  int *start = (int *)scratch(ctx, S_START, sizeof(int) * (C + 1));
  int *cellof = (int *)scratch(ctx, S_CELLOF, sizeof(int) * N);
  int *ids = (int *)scratch(ctx, S_IDS, sizeof(int) * N);
  double *px = (double *)scratch(ctx, S_PX, sizeof(double) * N);
  double *py = (double *)scratch(ctx, S_PY, sizeof(double) * N);
  double *pz = (double *)scratch(ctx, S_PZ, sizeof(double) * N);
  int *part = (int *)scratch(ctx, S_PART, sizeof(int) * scan_workspace_ints(C));
  if (!cnt || !fill || !start || !cellof || !ids || !px || !py || !pz || !part)
    return HSMC_ERR_NO_MEMORY;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.lx = h.p.lx; g.ly = h.p.ly; g.lz = h.p.lz;
  g.hx = h.p.lx / 2.0; g.hy = h.p.ly / 2.0; g.hz = h.p.lz / 2.0;
  g.sx = h.p.lx / nx; g.sy = h.p.ly / ny; g.sz = h.p.lz / nz;
  g.start = start; g.ids = ids; g.px = px; g.py = py; g.pz = pz;
  if (ctx->aux_replica == replica && ctx->aux_version == ctx->version && ctx->aux_n[0] == nx &&
      ctx->aux_n[1] == ny && ctx->aux_n[2] == nz)
    return HSMC_OK;
  const double *x = ctx->d_x + h.poff, *y = ctx->d_y + h.poff, *z = ctx->d_z + h.poff;
  HSMC_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int) * C, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(fill, 0, sizeof(int) * C, ctx->stream));
  const unsigned blocks = (unsigned)((N + 255) / 256);
  hipLaunchKernelGGL(k_grid_count, dim3(blocks), dim3(256), 0, ctx->stream, x, y, z, N, nx, ny, nz,
                     g.sx, g.sy, g.sz, cnt, cellof);
  exclusive_scan(ctx->stream, cnt, (int)C, start, part);
  hipLaunchKernelGGL(k_grid_scatter, dim3(blocks), dim3(256), 0, ctx->stream, x, y, z, N, cellof,
                     start, fill, ids, px, py, pz);
  HSMC_HIP(ctx, hipGetLastError());
  ctx->aux_replica = replica;
  ctx->aux_version = ctx->version;
  ctx->aux_n[0] = nx; ctx->aux_n[1] = ny; ctx->aux_n[2] = nz;
  ctx->aux_cut = cut;
  return HSMC_OK;
}

bool bad_replica(const hsmc_gpu_ctx *ctx, int r) { return !ctx || r < 0 || r >= ctx->R; }

// Scale factors of one scan, passed by value (no host-to-device copy).
struct SfTable {
  double v[64];
};

__global__ void k_store_sf(SfTable t, int n, double *__restrict__ dst) {
  const int k = threadIdx.x;
  if (k < n) dst[k] = t.v[k];
}

__global__ void k_zero_words(unsigned int *__restrict__ p, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Workspace layout of the scan: counts, fill counters and starts of the
// grid's C cells, cell and id of the N spheres, their sorted positions, the
// scale table and the result mask, each region 256-B aligned.
struct ScanLayout {
  size_t cnt, fill, start, cellof, ids, px, py, pz, sf, mask, part, bytes;
};

ScanLayout scan_layout(long N, long C) {
  ScanLayout l{};
  size_t o = 0;
  auto take = [&](size_t b) {
    const size_t at = o;
    o += align256(b);
    return at;
  };
  l.cnt = take(sizeof(int) * C);
  l.fill = take(sizeof(int) * C);
  l.start = take(sizeof(int) * (C + 1));
  l.cellof = take(sizeof(int) * N);
  l.ids = take(sizeof(int) * N);
  l.px = take(sizeof(double) * N);
  l.py = take(sizeof(double) * N);
  l.pz = take(sizeof(double) * N);
  l.sf = take(sizeof(double) * 64);
  l.mask = take(sizeof(unsigned long long));
  l.part = take(sizeof(int) * scan_workspace_ints(C));
  l.bytes = o;
  return l;
}

int scan_cut(const double *sf, int n_sf, double *cut) {
  double smin = 1.0;
  for (int k = 0; k < n_sf; ++k) {
    if (!(sf[k] > 0.0)) return HSMC_ERR_BAD_ARG;
    smin = std::min(smin, sf[k]);
  }
  *cut = (1.0 / smin) * (1.0 + 1e-6);
  return HSMC_OK;
}

}  // namespace

namespace hsmc {

size_t overlap_scan_workspace_bytes(const hsmc_gpu_ctx *ctx, int replica, const double *sf, int n_sf) {
  if (bad_replica(ctx, replica) || n_sf < 1 || n_sf > 64 || !sf) return 0;
  double cut = 0.0;
  if (scan_cut(sf, n_sf, &cut)) return 0;
  const hsmc_gpu_replica &p = ctx->reps[replica].p;
  const long C = (long)grid_cells_for(p.lx, cut) * grid_cells_for(p.ly, cut) * grid_cells_for(p.lz, cut);
  return scan_layout(p.n_part, C).bytes;
}

// The scan of hsmc_gpu_any_overlap_scaled with every buffer carved from the
// device workspace ws (ws_bytes long): the same kernels, the grid rebuilt from
// the replica's positions, counters cleared by kernels, the scale factors
// passed by value.  Nothing is allocated.
int overlap_scan_in_workspace(hsmc_gpu_ctx *ctx, int replica, const double *sf, int n_sf,
                              unsigned char *ws, size_t ws_bytes, uint8_t *no_overlap) {
  if (bad_replica(ctx, replica) || n_sf < 1 || n_sf > 64 || !sf || !no_overlap || !ws)
    return HSMC_ERR_BAD_ARG;
  double cut = 0.0;
  int st = scan_cut(sf, n_sf, &cut);
  if (st) return st;
  const RepHost &h = ctx->reps[replica];
  const int N = h.p.n_part;
  const int nx = grid_cells_for(h.p.lx, cut), ny = grid_cells_for(h.p.ly, cut),
            nz = grid_cells_for(h.p.lz, cut);
  const long C = (long)nx * ny * nz;
  const ScanLayout l = scan_layout(N, C);
  if (l.bytes > ws_bytes) {
    ctx->last_error = "the scan needs " + std::to_string(l.bytes) + " bytes of workspace, " +
                      std::to_string(ws_bytes) + " available";
    return HSMC_ERR_NOT_SUPPORTED;
  }
  int *cnt = (int *)(ws + l.cnt), *fill = (int *)(ws + l.fill), *start = (int *)(ws + l.start);
  int *cellof = (int *)(ws + l.cellof), *ids = (int *)(ws + l.ids);
  double *px = (double *)(ws + l.px), *py = (double *)(ws + l.py), *pz = (double *)(ws + l.pz);
  double *dsf = (double *)(ws + l.sf);
  unsigned long long *dmask = (unsigned long long *)(ws + l.mask);
  Grid g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.lx = h.p.lx; g.ly = h.p.ly; g.lz = h.p.lz;
  g.hx = h.p.lx / 2.0; g.hy = h.p.ly / 2.0; g.hz = h.p.lz / 2.0;
  g.sx = h.p.lx / nx; g.sy = h.p.ly / ny; g.sz = h.p.lz / nz;
  g.start = start; g.ids = ids; g.px = px; g.py = py; g.pz = pz;
  // counts and fill counters are adjacent regions: one clearing kernel
  const long zero_words = (long)((l.start - l.cnt) / sizeof(unsigned int));
  hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((zero_words + 255) / 256)), dim3(256), 0, ctx->stream,
                     (unsigned int *)cnt, zero_words);
  hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(64), 0, ctx->stream, (unsigned int *)dmask, 2L);
  SfTable t{};
  for (int k = 0; k < n_sf; ++k) t.v[k] = sf[k];
  hipLaunchKernelGGL(k_store_sf, dim3(1), dim3(64), 0, ctx->stream, t, n_sf, dsf);
  const double *x = ctx->d_x + h.poff, *y = ctx->d_y + h.poff, *z = ctx->d_z + h.poff;
  const unsigned blocks = (unsigned)((N + 255) / 256);
  hipLaunchKernelGGL(k_grid_count, dim3(blocks), dim3(256), 0, ctx->stream, x, y, z, N, nx, ny, nz,
                     g.sx, g.sy, g.sz, cnt, cellof);
  exclusive_scan(ctx->stream, cnt, (int)C, start, (int *)(ws + l.part));
  hipLaunchKernelGGL(k_grid_scatter, dim3(blocks), dim3(256), 0, ctx->stream, x, y, z, N, cellof,
                     start, fill, ids, px, py, pz);
  hipLaunchKernelGGL(k_any_overlap_scaled, dim3(blocks), dim3(256), 0, ctx->stream, g, N, dsf, n_sf,
                     cut * cut, dmask);
  HSMC_HIP(ctx, hipGetLastError());
  unsigned long long mask = 0;
  HSMC_HIP(ctx, hipMemcpyAsync(&mask, dmask, sizeof(mask), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < n_sf; ++k) no_overlap[k] = ((mask >> k) & 1ull) ? 0 : 1;
  return HSMC_OK;
}

}  // namespace hsmc

extern "C" {

int hsmc_gpu_cell_list(hsmc_gpu_ctx *ctx, int replica, int nx, int ny, int nz, int max_part,
                       int32_t *part_cell, int32_t *neigh_mat) {
  if (bad_replica(ctx, replica) || nx < 1 || ny < 1 || nz < 1 || max_part < 2)
    return HSMC_ERR_BAD_ARG;
  const RepHost &h = ctx->reps[replica];
  const double sx = h.p.lx / nx, sy = h.p.ly / ny, sz = h.p.lz / nz;
  if (sx < 1.0 || sy < 1.0 || sz < 1.0) return HSMC_ERR_CELL_TOO_SMALL;  // cell_list.c:126-129
  const long C = (long)nx * ny * nz;
  const int N = h.p.n_part;
  if (part_cell) {
    int *cnt = (int *)scratch(ctx, S_CNT, sizeof(int) * C);
    int *fill = (int *)scratch(ctx, S_FILL, sizeof(int) * C);
    int *start = (int *)scratch(ctx, S_START, sizeof(int) * (C + 1));
    int *cellof = (int *)scratch(ctx, S_CELLOF, sizeof(int) * N);
    int *ids = (int *)scratch(ctx, S_IDS, sizeof(int) * N);
    double *px = (double *)scratch(ctx, S_PX, sizeof(double) * N);
    double *py = (double *)scratch(ctx, S_PY, sizeof(double) * N);
    double *pz = (double *)scratch(ctx, S_PZ, sizeof(double) * N);
    int *rows = (int *)scratch(ctx, S_OUT0, sizeof(int) * C * max_part);
    int *part = (int *)scratch(ctx, S_PART, sizeof(int) * scan_workspace_ints(C));
    if (!cnt || !fill || !start || !cellof || !ids || !px || !py || !pz || !rows || !part)
      return HSMC_ERR_NO_MEMORY;
    ctx->aux_replica = -1;  // the CSR scratch now holds this grid
    const double *x = ctx->d_x + h.poff, *y = ctx->d_y + h.poff, *z = ctx->d_z + h.poff;
    HSMC_HIP(ctx, hipMemsetAsync(cnt, 0, sizeof(int) * C, ctx->stream));
    HSMC_HIP(ctx, hipMemsetAsync(fill, 0, sizeof(int) * C, ctx->stream));
    const unsigned blocks = (unsigned)((N + 255) / 256);
    hipLaunchKernelGGL(k_grid_count, dim3(blocks), dim3(256), 0, ctx->stream, x, y, z, N, nx, ny,
                       nz, sx, sy, sz, cnt, cellof);
    exclusive_scan(ctx->stream, cnt, (int)C, start, part);
    hipLaunchKernelGGL(k_grid_scatter, dim3(blocks), dim3(256), 0, ctx->stream, x, y, z, N, cellof,
                       start, fill, ids, px, py, pz);
    hipLaunchKernelGGL(k_part_cell, dim3((unsigned)((C + 255) / 256)), dim3(256), 0, ctx->stream,
                       start, (int)C, ids, max_part, rows, ctx->d_err);
    HSMC_HIP(ctx, hipGetLastError());
    HSMC_HIP(ctx, hipMemcpyAsync(part_cell, rows, sizeof(int) * C * max_part,
                                 hipMemcpyDeviceToHost, ctx->stream));
    HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int st = check_device_errors(ctx);
    if (st) return st;
  }
  if (neigh_mat) {  // neigh_id (cell_list.c:266-301), host-side table
    for (int a = 0; a < nx; ++a)
      for (int b = 0; b < ny; ++b)
        for (int c = 0; c < nz; ++c) {
          const long ref = ((long)a * ny + b) * nz + c;
          int counter = 0;
          for (int i = -1; i < 2; ++i)
            for (int j = -1; j < 2; ++j)
              for (int k = -1; k < 2; ++k) {
                int ix = a + i, iy = b + j, iz = c + k;
                if (ix > nx - 1) ix -= nx; else if (ix < 0) ix += nx;
                if (iy > ny - 1) iy -= ny; else if (iy < 0) iy += ny;
                if (iz > nz - 1) iz -= nz; else if (iz < 0) iz += nz;
                neigh_mat[ref * 27 + counter] = (ix * ny + iy) * nz + iz;
                ++counter;
              }
        }
  }
  return HSMC_OK;
}

int hsmc_gpu_overlap_batch(hsmc_gpu_ctx *ctx, int replica, const int32_t *idx, const double *xyz,
                           int m, uint8_t *out) {
  if (bad_replica(ctx, replica) || m < 0 || (m > 0 && (!idx || !xyz || !out))) return HSMC_ERR_BAD_ARG;
  if (m == 0) return HSMC_OK;
  for (int k = 0; k < m; ++k)
    if (idx[k] < 0 || idx[k] >= ctx->reps[replica].p.n_part) return HSMC_ERR_BAD_ARG;
  Grid g;
  int st = build_grid(ctx, replica, 1.0 + 1e-9, g);
  if (st) return st;
  int *didx = (int *)scratch(ctx, S_Q0, sizeof(int) * m);
  double *dq = (double *)scratch(ctx, S_Q1, sizeof(double) * 3 * m);
  unsigned char *dout = (unsigned char *)scratch(ctx, S_Q2, m);
  if (!didx || !dq || !dout) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemcpyAsync(didx, idx, sizeof(int) * m, hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(dq, xyz, sizeof(double) * 3 * m, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_overlap_batch, dim3((m + 255) / 256), dim3(256), 0, ctx->stream, g, didx, dq,
                     m, dout);
  HSMC_HIP(ctx, hipGetLastError());
  HSMC_HIP(ctx, hipMemcpyAsync(out, dout, m, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HSMC_OK;
}

int hsmc_gpu_overlap_scaled(hsmc_gpu_ctx *ctx, int replica, const int32_t *idx, const double *sf,
                            int m, uint8_t *out) {
  if (bad_replica(ctx, replica) || m < 0 || (m > 0 && (!idx || !sf || !out))) return HSMC_ERR_BAD_ARG;
  if (m == 0) return HSMC_OK;
  double smin = 1.0;
  for (int k = 0; k < m; ++k) {
    if (idx[k] < 0 || idx[k] >= ctx->reps[replica].p.n_part || !(sf[k] > 0.0)) return HSMC_ERR_BAD_ARG;
    smin = std::min(smin, sf[k]);
  }
  Grid g;
  int st = build_grid(ctx, replica, (1.0 / smin) * (1.0 + 1e-9), g);
  if (st) return st;
  const RepHost &h = ctx->reps[replica];
  int *didx = (int *)scratch(ctx, S_Q0, sizeof(int) * m);
  double *dsf = (double *)scratch(ctx, S_Q1, sizeof(double) * m);
  unsigned char *dout = (unsigned char *)scratch(ctx, S_Q2, m);
  if (!didx || !dsf || !dout) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemcpyAsync(didx, idx, sizeof(int) * m, hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(dsf, sf, sizeof(double) * m, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_overlap_scaled, dim3((m + 255) / 256), dim3(256), 0, ctx->stream, g,
                     ctx->d_x + h.poff, ctx->d_y + h.poff, ctx->d_z + h.poff, didx, dsf, m, dout);
  HSMC_HIP(ctx, hipGetLastError());
  HSMC_HIP(ctx, hipMemcpyAsync(out, dout, m, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HSMC_OK;
}

int hsmc_gpu_any_overlap_scaled(hsmc_gpu_ctx *ctx, int replica, const double *sf, int n_sf,
                                uint8_t *no_overlap) {
  if (bad_replica(ctx, replica) || n_sf < 1 || n_sf > 64 || !sf || !no_overlap) return HSMC_ERR_BAD_ARG;
  double smin = 1.0;
  for (int k = 0; k < n_sf; ++k) {
    if (!(sf[k] > 0.0)) return HSMC_ERR_BAD_ARG;
    smin = std::min(smin, sf[k]);
  }
  const double cut = (1.0 / smin) * (1.0 + 1e-6);
  Grid g;
  int st = build_grid(ctx, replica, cut, g);
  if (st) return st;
  const int N = ctx->reps[replica].p.n_part;
  double *dsf = (double *)scratch(ctx, S_Q1, sizeof(double) * n_sf);
  unsigned long long *dmask = (unsigned long long *)scratch(ctx, S_OUT1, sizeof(unsigned long long));
  if (!dsf || !dmask) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemcpyAsync(dsf, sf, sizeof(double) * n_sf, hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(dmask, 0, sizeof(unsigned long long), ctx->stream));
  hipLaunchKernelGGL(k_any_overlap_scaled, dim3((N + 255) / 256), dim3(256), 0, ctx->stream, g, N,
                     dsf, n_sf, cut * cut, dmask);
  HSMC_HIP(ctx, hipGetLastError());
  unsigned long long mask = 0;
  HSMC_HIP(ctx, hipMemcpyAsync(&mask, dmask, sizeof(mask), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int k = 0; k < n_sf; ++k) no_overlap[k] = ((mask >> k) & 1ull) ? 0 : 1;
  return HSMC_OK;
}

int hsmc_gpu_pair_dist(hsmc_gpu_ctx *ctx, int replica, const int32_t *i, const int32_t *j, int m,
                       double sf, double *out) {
  if (bad_replica(ctx, replica) || m < 0 || (m > 0 && (!i || !j || !out)) || !(sf > 0.0))
    return HSMC_ERR_BAD_ARG;
  if (m == 0) return HSMC_OK;
  const RepHost &h = ctx->reps[replica];
  for (int k = 0; k < m; ++k)
    if (i[k] < 0 || i[k] >= h.p.n_part || j[k] < 0 || j[k] >= h.p.n_part) return HSMC_ERR_BAD_ARG;
  int *di = (int *)scratch(ctx, S_Q0, sizeof(int) * m);
  int *dj = (int *)scratch(ctx, S_Q2, sizeof(int) * m);
  double *dout = (double *)scratch(ctx, S_Q1, sizeof(double) * m);
  if (!di || !dj || !dout) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemcpyAsync(di, i, sizeof(int) * m, hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(dj, j, sizeof(int) * m, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_pair_dist, dim3((m + 255) / 256), dim3(256), 0, ctx->stream,
                     ctx->d_x + h.poff, ctx->d_y + h.poff, ctx->d_z + h.poff, di, dj, m, sf,
                     h.p.lx, h.p.ly, h.p.lz, dout);
  HSMC_HIP(ctx, hipGetLastError());
  HSMC_HIP(ctx, hipMemcpyAsync(out, dout, sizeof(double) * m, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HSMC_OK;
}

static int pair_hist(hsmc_gpu_ctx *ctx, int replica, double dr_bin, int nbins, double rmax_eff,
                     int64_t *counts, int64_t *oor) {
  if (bad_replica(ctx, replica) || !(dr_bin > 0.0) || nbins < 1 || nbins > 8192 ||
      !(rmax_eff > 1.0) || !counts)
    return HSMC_ERR_BAD_ARG;
  const RepHost &h = ctx->reps[replica];
  const double half = std::min(h.p.lx, std::min(h.p.ly, h.p.lz)) / 2.0;
  if (rmax_eff > half * (1.0 + 1e-12)) {
    ctx->last_error = "histogram cutoff beyond half the box (minimum image)";
    return HSMC_ERR_BAD_ARG;
  }
  const double cut = rmax_eff * (1.0 + 1e-9);
  Grid g;
  int st = build_grid(ctx, replica, cut, g);
  if (st) return st;
  const int N = h.p.n_part;
  unsigned long long *dcounts = (unsigned long long *)scratch(ctx, S_OUT0, sizeof(unsigned long long) * (nbins + 1));
  if (!dcounts) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemsetAsync(dcounts, 0, sizeof(unsigned long long) * (nbins + 1), ctx->stream));
  const double cut2 = (rmax_eff * (1.0 + 1e-6)) * (rmax_eff * (1.0 + 1e-6));
  const long C = (long)g.nx * g.ny * g.nz;
  if (N >= 16 * C && pair_tiled_lds(nbins) <= 64 * 1024 && 27 * C < (1L << 31)) {
    // cells of 16 spheres or more on average: one wave per cell pair
    hipLaunchKernelGGL(k_pair_hist_tiled, dim3((unsigned)(27 * C)), dim3(64), pair_tiled_lds(nbins),
                       ctx->stream, g, nbins, dr_bin, rmax_eff, cut2, dcounts, dcounts + nbins);
  } else {
    hipLaunchKernelGGL(k_pair_hist, dim3((N + 255) / 256), dim3(256), sizeof(unsigned int) * nbins,
                       ctx->stream, g, N, dr_bin, nbins, rmax_eff, cut2, dcounts, dcounts + nbins);
  }
  HSMC_HIP(ctx, hipGetLastError());
  std::vector<unsigned long long> hc(nbins + 1);
  HSMC_HIP(ctx, hipMemcpyAsync(hc.data(), dcounts, sizeof(unsigned long long) * (nbins + 1),
                               hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int b = 0; b < nbins; ++b) counts[b] = (int64_t)hc[b];
  if (oor) *oor = (int64_t)hc[nbins];
  return HSMC_OK;
}

int hsmc_gpu_rdf_counts(hsmc_gpu_ctx *ctx, int replica, double dr_bin, int nbins, double rmax_eff,
                        int64_t *counts, int64_t *out_of_range) {
  return pair_hist(ctx, replica, dr_bin, nbins, rmax_eff, counts, out_of_range);
}

int hsmc_gpu_contact_counts(hsmc_gpu_ctx *ctx, int replica, double dr_bin, int nbins,
                            double rmax_eff, int64_t *counts, int64_t *out_of_range) {
  return pair_hist(ctx, replica, dr_bin, nbins, rmax_eff, counts, out_of_range);
}

int hsmc_gpu_widom_points(hsmc_gpu_ctx *ctx, int replica, const double *xyz, int m, uint8_t *hit,
                          int64_t *hits) {
  if (bad_replica(ctx, replica) || m < 0 || (m > 0 && !xyz)) return HSMC_ERR_BAD_ARG;
  if (m == 0) {
    if (hits) *hits = 0;
    return HSMC_OK;
  }
  Grid g;
  int st = build_grid(ctx, replica, 1.0 + 1e-9, g);
  if (st) return st;
  double *dq = (double *)scratch(ctx, S_Q1, sizeof(double) * 3 * m);
  unsigned char *dout = (unsigned char *)scratch(ctx, S_Q2, m);
  if (!dq || !dout) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemcpyAsync(dq, xyz, sizeof(double) * 3 * m, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_widom_points, dim3((m + 255) / 256), dim3(256), 0, ctx->stream, g, dq, m, dout);
  HSMC_HIP(ctx, hipGetLastError());
  std::vector<uint8_t> tmp;
  uint8_t *dst = hit;
  if (!dst) {
    tmp.resize(m);
    dst = tmp.data();
  }
  HSMC_HIP(ctx, hipMemcpyAsync(dst, dout, m, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (hits) {
    int64_t s = 0;
    for (int k = 0; k < m; ++k) s += dst[k];
    *hits = s;
  }
  return HSMC_OK;
}

int hsmc_gpu_widom_uniform(hsmc_gpu_ctx *ctx, int replica, uint64_t seed, uint64_t sample,
                           int64_t first, int64_t n, int64_t *hits) {
  if (bad_replica(ctx, replica) || n < 0 || first < 0 || !hits) return HSMC_ERR_BAD_ARG;
  *hits = 0;
  if (n == 0) return HSMC_OK;
  Grid g;
  int st = build_grid(ctx, replica, 1.0 + 1e-9, g);
  if (st) return st;
  unsigned long long *dh = (unsigned long long *)scratch(ctx, S_OUT1, sizeof(unsigned long long));
  if (!dh) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemsetAsync(dh, 0, sizeof(unsigned long long), ctx->stream));
  const long long per_block = 256LL * 16;
  long long blocks = (n + per_block - 1) / per_block;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(k_widom_uniform, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, g,
                     (uint32_t)seed, (uint32_t)(seed >> 32), (unsigned long long)sample,
                     (long long)first, (long long)n, dh);
  HSMC_HIP(ctx, hipGetLastError());
  unsigned long long hv = 0;
  HSMC_HIP(ctx, hipMemcpyAsync(&hv, dh, sizeof(hv), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *hits = (int64_t)hv;
  return HSMC_OK;
}

int hsmc_gpu_widom_cells(hsmc_gpu_ctx *ctx, int replica, uint64_t seed, uint64_t sample,
                         int32_t per_cell, int64_t *hits, int64_t *n_points, int32_t out_grid[3]) {
  if (bad_replica(ctx, replica) || per_cell < 0 || !hits) return HSMC_ERR_BAD_ARG;
  *hits = 0;
  Grid g;
  int st = build_grid(ctx, replica, 1.0 + 1e-9, g);
  if (st) return st;
  const long C = (long)g.nx * g.ny * g.nz;
  if (n_points) *n_points = (int64_t)per_cell * C;
  if (out_grid) {
    out_grid[0] = g.nx;
    out_grid[1] = g.ny;
    out_grid[2] = g.nz;
  }
  if (per_cell == 0) return HSMC_OK;
  // spheres of unit diameter centred in a 27-cell stencil fit in a box one
  // diameter wider: at most the close-packing fraction (0.7405, taken as
  // 0.78) of its volume over the sphere volume
  const double span = (std::min(3, g.nx) * g.sx + 1.0) * (std::min(3, g.ny) * g.sy + 1.0) *
                      (std::min(3, g.nz) * g.sz + 1.0);
  const long bound = std::min((long)ctx->reps[replica].p.n_part, (long)ceil(0.78 * span / (M_PI / 6.0)));
  if (bound > kWidomCellsLcap) {
    ctx->last_error = "observable-grid cells too wide for the stratified insertion kernel";
    return HSMC_ERR_NOT_SUPPORTED;
  }
  unsigned long long *dh = (unsigned long long *)scratch(ctx, S_OUT1, sizeof(unsigned long long));
  if (!dh) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemsetAsync(dh, 0, sizeof(unsigned long long), ctx->stream));
  const long blocks = std::min((C + kWidomCellsWaves - 1) / kWidomCellsWaves, 1L << 20);
  hipLaunchKernelGGL(k_widom_cells, dim3((unsigned)blocks), dim3(64 * kWidomCellsWaves), 0, ctx->stream, g,
                     (uint32_t)seed, (uint32_t)(seed >> 32), (unsigned long long)sample, (int)per_cell,
                     dh, ctx->d_err);
  HSMC_HIP(ctx, hipGetLastError());
  unsigned long long hv = 0;
  HSMC_HIP(ctx, hipMemcpyAsync(&hv, dh, sizeof(hv), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *hits = (int64_t)hv;
  return check_device_errors(ctx);
}

}  // extern "C"
