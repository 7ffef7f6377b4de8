// device_common.h -- FP64 distance arithmetic, cell indexing and the
// Philox4x32-10 counter RNG shared by every kernel of libhsmc_gpu.
//
// All FP64 expressions here are evaluated without contraction (the TU is
// built with -ffp-contract=off and the pragma below), in the reference's
// order (moves.c:400-431), so every overlap decision and histogram bin is
// rounded exactly as the reference rounds it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace hsmc {

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): 10 rounds of the
// multiply-xor bijection with the Weyl key schedule.
__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2,
                                                       uint32_t c3, uint32_t k0, uint32_t k1,
                                                       uint32_t out[4]) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0;
    c1 = (uint32_t)p1;
    c2 = n2;
    c3 = (uint32_t)p0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

// rng_get_double (rng.c:29-31): a 32-bit draw divided by 2^32-1, in [0,1].
__host__ __device__ __forceinline__ double u32_unit_closed(uint32_t u) {
  return (double)u / 4294967295.0;
}

// Occupant pick of the checkerboard sweep: rng_get_int's rule (rng.c:34-36,
// gsl_rng_uniform_int) applied to the sweep's Philox stream.  `first` is word
// 0 of attempt t's block (counter cell, t, sweep); k = first / scale with
// scale = (2^32 - 1) / n, and a k >= n is drawn again from the block with
// counter word 1 = t | (retry << 16), retry = 1, 2, ...  Uniform on [0, n)
// for any n in [1, 2^32 - 1] and t < 2^16.
__host__ __device__ __forceinline__ uint32_t occupant_pick(uint32_t first, uint32_t n,
                                                           uint32_t cell, uint32_t t,
                                                           unsigned long long sweep, uint32_t k0,
                                                           uint32_t k1) {
  const uint32_t scale = 0xFFFFFFFFu / n;
  uint32_t k = first / scale;
  for (uint32_t rr = 1; k >= n; ++rr) {
    uint32_t w[4];
    philox4x32_10(cell, t | (rr << 16), (uint32_t)sweep, (uint32_t)(sweep >> 32), k0, k1, w);
    k = w[0] / scale;
  }
  return k;
}

// Minimum image of one component (moves.c:419-424): d - L above half a box,
// d + L below minus half a box.  Both candidates are computed and one is
// selected, so the compiler emits selects rather than branches.
__host__ __device__ __forceinline__ double min_image(double d, double L, double half) {
  const double dm = d - L, dp = d + L;
  return d > half ? dm : (d < -half ? dp : d);
}

// Squared distance of the reference's compute_dist with unit scale
// (moves.c:414-427): a - b per axis, minimum image, dx*dx+dy*dy+dz*dz.
// dr < 1 <=> dr^2 < 1 for correctly rounded sqrt, so overlap tests use it.
__host__ __device__ __forceinline__ double dist2(double ax, double ay, double az, double bx,
                                                 double by, double bz, double lx, double ly,
                                                 double lz, double hx, double hy, double hz) {
  const double dx = min_image(ax - bx, lx, hx);
  const double dy = min_image(ay - by, ly, hy);
  const double dz = min_image(az - bz, lz, hz);
  return dx * dx + dy * dy + dz * dz;
}

// compute_dist with a scale factor (moves.c:405-427): dx = (a-b)*sf and a
// box of L*sf, half box (L*sf)/2.
__host__ __device__ __forceinline__ double dist2_scaled(double ax, double ay, double az,
                                                        double bx, double by, double bz,
                                                        double lx, double ly, double lz,
                                                        double sf) {
  const double Lx = lx * sf, Ly = ly * sf, Lz = lz * sf;
  const double hx = Lx / 2.0, hy = Ly / 2.0, hz = Lz / 2.0;
  const double dx = min_image((ax - bx) * sf, Lx, hx);
  const double dy = min_image((ay - by) * sf, Ly, hy);
  const double dz = min_image((az - bz) * sf, Lz, hz);
  return dx * dx + dy * dy + dz * dz;
}

// apply_pbc for one coordinate (moves.c:215-226): a single wrap (selected,
// not branched, as in min_image).
__host__ __device__ __forceinline__ double wrap_once(double x, double L) {
  const double xm = x - L, xp = x + L;
  return x > L ? xm : (x < 0.0 ? xp : x);
}

// Cell of coordinate x on a grid of g cells of side `side` whose origin is
// shifted by `shift` in [0, side): t = x - shift (wrapped), (int)(t * (1/side)),
// clamped to g-1 at t == L.  Multiplying by the rounded reciprocal moves each
// cell edge by a few ulp at most, so every cell stays wider than the grid's
// minimum side of 1 + 1e-9 diameters.
__host__ __device__ __forceinline__ int shifted_cell(double x, double shift, double L,
                                                     double inv_side, int g) {
  const double t0 = x - shift, t1 = t0 + L;
  const double t = t0 < 0.0 ? t1 : t0;
  const int c = (int)(t * inv_side);
  return c < g ? c : g - 1;
}

// Cell of coordinate x on an unshifted grid (cell_part_idx, cell_list.c:244-246),
// clamped into [0, g-1].
__host__ __device__ __forceinline__ int plain_cell(double x, double side, int g) {
  int c = (int)(x / side);
  if (c >= g) c = g - 1;
  if (c < 0) c = 0;
  return c;
}

// Per-axis neighbour offsets of the 27-cell stencil that are distinct on a
// periodic grid of g cells: g >= 3 -> {-1,0,1}; g == 2 -> {0,1}; g == 1 -> {0}.
__host__ __device__ __forceinline__ bool stencil_axis_ok(int d, int g) {
  if (g >= 3) return true;
  if (g == 2) return d != -1;
  return d == 0;
}

// The 27-cell stencil in order of distance class |dx| + |dy| + |dz| (own
// cell, faces, edges, corners; lexicographic within a class), for scans
// that can stop at the first sphere found near a point.  Entry k of the
// order is the lexicographic index t = (dx+1)*9 + (dy+1)*3 + (dz+1), kept
// in 5 bits at position k % 9 of word k / 9; the words are built at compile
// time, so a lane finds its offset with a few shifts.
namespace detail {
constexpr uint64_t stencil_order_word(int w) {
  uint64_t word = 0;
  int k = 0;
  for (int cls = 0; cls <= 3; ++cls)
    for (int t = 0; t < 27; ++t) {
      const int ox = t / 9 - 1, oy = (t / 3) % 3 - 1, oz = t % 3 - 1;
      const int c = (ox < 0 ? -ox : ox) + (oy < 0 ? -oy : oy) + (oz < 0 ? -oz : oz);
      if (c != cls) continue;
      if (k / 9 == w) word |= (uint64_t)t << (5 * (k % 9));
      ++k;
    }
  return word;
}
}  // namespace detail

// Offset j in [0, 27) of that order; returns its distance class.
__host__ __device__ __forceinline__ int stencil_by_distance(int j, int &dx, int &dy, int &dz) {
  constexpr uint64_t w0 = detail::stencil_order_word(0), w1 = detail::stencil_order_word(1),
                     w2 = detail::stencil_order_word(2);
  const uint64_t word = j < 9 ? w0 : (j < 18 ? w1 : w2);
  const int t = (int)((word >> (5 * (j % 9))) & 31u);
  dx = t / 9 - 1;
  dy = (t / 3) % 3 - 1;
  dz = t % 3 - 1;
  return (dx < 0 ? -dx : dx) + (dy < 0 ? -dy : dy) + (dz < 0 ? -dz : dz);
}

__host__ __device__ __forceinline__ int wrap_index(int i, int g) {
  if (i < 0) i += g;
  else if (i >= g) i -= g;
  return i;
}

// Inclusive sum of v over each group of LPC (32 or 64) consecutive lanes of
// a full wave, in DPP moves with no LDS round trip: shifts by 1, 2, 4 and 8
// within each row of 16 lanes, then lane 15 of each even row added to the
// row above it (the two halves of a 32-lane group) and, for 64-lane groups,
// lane 31 added to rows 2 and 3.  A shifted-in lane outside the row reads 0.
template <int LPC>
__device__ __forceinline__ int group_inclusive_sum(int v) {
  v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xf, 0xf, true);   // row_shr:1
  v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xf, 0xf, true);   // row_shr:2
  v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xf, 0xf, true);   // row_shr:4
  v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xf, 0xf, true);   // row_shr:8
  v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xa, 0xf, false);  // row_bcast:15 into rows 1, 3
  if constexpr (LPC == 64)
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xc, 0xf, false);  // row_bcast:31 into rows 2, 3
  return v;
}

}  // namespace hsmc
