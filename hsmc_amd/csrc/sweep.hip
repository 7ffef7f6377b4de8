// sweep.hip -- checkerboard cell-parallel Metropolis sweep (the hot path).
//
// One sweep of every replica (the work of sweep_nvt, nvt.c:201-209, which
// makes N part_move attempts, moves.c:27-80):
//   1. k_rebin: the sweep grid is shifted by a random vector in [0, side)^3
//      and every particle is filed into its cell (slot order by atomics);
//   2. k_phase x 8: the 8 sublattices (cells whose three indices have a
//      fixed parity) are visited in a random order.  Cells of one sublattice
//      are at least one cell apart, cells are wider than a diameter, and a
//      move that leaves its cell is rejected, so the active cells of a phase
//      never interact and run concurrently.  An active cell makes as many
//      attempts as it has occupants, each on an occupant drawn uniformly
//      (rng_get_int semantics) and displaced by (r - 0.5) * dr_max per axis
//      (moves.c:52-54), wrapped by apply_pbc (moves.c:215-226) and tested
//      for overlap with every other sphere in the 27 neighbouring cells with
//      the reference's FP64 arithmetic (moves.c:400-431).
// With one replica, the sweeps of a batch after the first are chained: the
// last phase of each sweep files every sphere into the next sweep's shifted
// grid (at its final position), so those sweeps need no k_rebin, and the
// batch's first rebin draws the shift and order of all its sweeps.  The cell
// list a sweep reads is the same either way, up to the order of a cell's
// slots, which no result depends on (occupants are ranked by id).
// The random numbers come from Philox4x32-10 keyed by the replica seed with
// counters (cell, attempt, sweep), so a run is reproducible bit for bit and
// equals the CPU restatement in the oracle.
#include <stdio.h>

#include <algorithm>
#include <vector>

#include "ctx.h"
#include "device_common.h"

using namespace hsmc;

namespace {

// Vector store at agent scope: written through the XCD's L2 (the next
// phase's workgroups read it from other XCDs).
__device__ __forceinline__ void store_through(double *p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Shift of the sweep grid for one sweep: 3 uniform draws in [0, side).
__device__ __forceinline__ void sweep_shift(const RepDev &rp, unsigned long long sweep,
                                            double &shx, double &shy, double &shz) {
  uint32_t u[4];
  philox4x32_10(0xFFFFFFFFu, 0u, (uint32_t)sweep, (uint32_t)(sweep >> 32), rp.k0, rp.k1, u);
  shx = ((double)u[0] * 0x1p-32) * rp.sx;
  shy = ((double)u[1] * 0x1p-32) * rp.sy;
  shz = ((double)u[2] * 0x1p-32) * rp.sz;
}

// Order of the 8 sublattices in one sweep: a Fisher-Yates order of 0..7
// drawn from one Philox block, kept as 8 nibbles (nibble p = sublattice of
// phase p).
__device__ __forceinline__ uint32_t sweep_order(const RepDev &rp, unsigned long long sweep) {
  uint32_t u[4];
  philox4x32_10(0xFFFFFFFEu, 0u, (uint32_t)sweep, (uint32_t)(sweep >> 32), rp.k0, rp.k1, u);
  uint32_t pk = 0x76543210u;
#pragma unroll
  for (int h = 0; h < 7; ++h) {
    const int i = 7 - h;
    const uint32_t half = (u[h >> 1] >> (16 * (h & 1))) & 0xFFFFu;
    const int j = (int)((half * (uint32_t)(i + 1)) >> 16);
    const uint32_t vi = (pk >> (4 * i)) & 0xFu, vj = (pk >> (4 * j)) & 0xFu;
    pk &= ~((0xFu << (4 * i)) | (0xFu << (4 * j)));
    pk |= (vi << (4 * j)) | (vj << (4 * i));
  }
  return pk;
}

// Work item i of a sweep's rebin: clear cell i of the next sweep's count
// buffer, file particle i into its cell of the shifted grid, and draw
// replica i's shift and sublattice order.  Counts have one buffer per sweep
// parity; with one replica the slot records do too, so that a sweep's last
// phase can file the next sweep's list while it reads its own.  The
// particle's position (and, for several replicas, its replica index) is
// loaded first, so the shift's draw runs while the loads are in flight; one
// replica's record is the kernel argument (no per-lane loads of it); the
// table entries are drawn last, while the slot index's atomic is in flight.
template <bool ONE>
__device__ __forceinline__ void rebin_one(int i, const RepDev *__restrict__ reps, const RepDev &rp0,
                                          SweepDev *__restrict__ swp, int R, int kidx, int nsw,
                                          const int *__restrict__ prep, int P, int C,
                                          const double *__restrict__ x, const double *__restrict__ y,
                                          const double *__restrict__ z, int *__restrict__ cnt,
                                          SlotRec *__restrict__ slot, int cap,
                                          unsigned long long sweep, unsigned *__restrict__ err) {
  const int buf = (int)(sweep & 1ull);
  const int ip = i < P ? i : 0;  // a work item past the particles reads particle 0 and drops it
  const double px = x[ip], py = y[ip], pz = z[ip];
  const int rr = ONE ? 0 : prep[ip];
  if (i < C) cnt[(size_t)(buf ^ 1) * C + i] = 0;  // next sweep's buffer
  if (i < P) {
    const RepDev rp = ONE ? rp0 : reps[rr];
    double shx, shy, shz;
    sweep_shift(rp, sweep, shx, shy, shz);
    const int cx = shifted_cell(px, shx, rp.lx, rp.ix, rp.gx);
    const int cy = shifted_cell(py, shy, rp.ly, rp.iy, rp.gy);
    const int cz = shifted_cell(pz, shz, rp.lz, rp.iz, rp.gz);
    const int c = rp.coff + (cx * rp.gy + cy) * rp.gz + cz;
    const int sl = atomicAdd(&cnt[(size_t)buf * C + c], 1);
    if (sl < cap) {
      SlotRec rec;
      rec.x = px;
      rec.y = py;
      rec.z = pz;
      rec.id = i - rp.poff;
      rec.pad = 0;
      // slot-major: slot q of every cell is one array; with one replica,
      // one record buffer per sweep parity, as for the counts
      slot[((size_t)(ONE ? buf : 0) * cap + sl) * C + c] = rec;
    } else {
      atomicOr(err, kErrOverflow);
    }
  }
  // the shift and sublattice order read by the phases: entry i of replica i
  // for several replicas; with one replica, entries kidx .. kidx + nsw - 1
  // for this sweep and the nsw - 1 after it (a batch of chained sweeps)
  if (i < (ONE ? nsw : R)) {
    const RepDev ri = ONE ? rp0 : reps[i];
    const unsigned long long si = ONE ? sweep + (unsigned long long)i : sweep;
    SweepDev d;
    sweep_shift(ri, si, d.shx, d.shy, d.shz);
    d.order = sweep_order(ri, si);
    d.pad = 0u;
    swp[ONE ? kidx + i : i] = d;
  }
}

template <bool ONE>
__global__ void __launch_bounds__(256)
k_rebin(const RepDev *__restrict__ reps, const RepDev rp0, SweepDev *__restrict__ swp, int R,
        int kidx, int nsw,
        const int *__restrict__ prep, int P, int C,
        const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
        int *__restrict__ cnt, SlotRec *__restrict__ slot, int cap,
        const unsigned long long *__restrict__ sweep_base, int k, unsigned *__restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned long long sweep = *sweep_base + (unsigned long long)(long long)k;
  rebin_one<ONE>(i, reps, rp0, swp, R, kidx, nsw, prep, P, C, x, y, z, cnt, slot, cap, sweep, err);
}

// Value of v in lane t of the wave, for a t that is the same in every lane:
// v_readlane, an SGPR read with no LDS round trip.  Used when a lane group is
// the whole wave; with two 32-lane groups per wave the halves need different
// lanes, and ds_bpermute (__shfl) is faster than two readlanes and a select.
__device__ __forceinline__ double wave_lane(double v, int t) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, t);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), t);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// A sphere that one lane of a last phase files into the next sweep's list:
// an occupant of the lane's active cell at its final position.
struct FileReq {
  double x, y, z;
  int id;
  bool has;
};

// A lane's copies of its share of the stencil list (entries lane,
// lane + LPC, ...): a vector value, so variable indexing stays in registers.
template <int M>
using dvec = double __attribute__((ext_vector_type(M)));

// One active cell's attempts in one sublattice phase: the work of one group
// of LPC lanes.  a is the active-cell index, a0 the first active cell of the
// group's workgroup (of cells_per_block consecutive cells), Lm the group's
// LDS region.  One replica: sw0 is the sweep's table entry, and with FILE
// set (the last phase of a sweep whose successor has no rebin) each lane
// returns in req the occupant of the cell it files into the next sweep's
// cell list, at its final position.
template <int CAP, int LPC, int LCAP, bool ONE, bool FILE>
__device__ __forceinline__ void phase_cell(
    int a, int a0, int lane, int *Lm, const RepDev *__restrict__ reps, const RepDev &rp0,
    const SweepDev *__restrict__ swp, const SweepDev &sw0, const int *__restrict__ arep, int R,
    int A, int phase, unsigned long long sweep,
    double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
    int *__restrict__ cnt, int C, SlotRec *__restrict__ slot_all,
    unsigned int *__restrict__ cacc, unsigned int *__restrict__ cleft,
    unsigned int *__restrict__ err, FileReq &req) {
  // The stencil list (the spheres of the 27 neighbour cells, at most LCAP, a
  // bound the host derives from the cell side) lives in registers: entry e
  // is held by lane e % LPC as its copy e / LPC.  LDS holds the slot of each
  // entry, the particle ids, id order and positions of the active cell's
  // occupants, and (32-lane groups) each attempt's displacement.
  constexpr int M = (LCAP + LPC - 1) / LPC;
  int *Sp = Lm + LCAP;                                 // ids of the active cell's occupants
  int *Lo = Sp + CAP;                                  // occupants in ascending id order
  double *Op = reinterpret_cast<double *>(Lo + CAP);   // occupants' positions (x, y, z)
  double *Od = Op + 3 * CAP;                           // each attempt's displacement

  const int par = (int)(sweep & 1ull);
  const int *cntb = cnt + (size_t)par * C;
  SlotRec *slot = slot_all + (size_t)(ONE ? par : 0) * CAP * C;
  // the replica's record (a kernel argument when there is one replica) and
  // the sweep's shift and sublattice order, which the rebin drew (loaded by
  // the caller when there is one replica).  Each replica's active cells are
  // padded to a multiple of 8 and cells_per_block divides 8, so a workgroup
  // lies in one replica: r is uniform and the record and table are scalar
  // loads (a0 < A always).
  const int r = ONE ? 0 : arep[a0];
  const RepDev rp = ONE ? rp0 : reps[r];
  const SweepDev sw = ONE ? sw0 : swp[r];
  const int colour = (int)((sw.order >> (4 * phase)) & 0xFu);
  const double shx = sw.shx, shy = sw.shy, shz = sw.shz;

  // active cell of this work item.  No exit before the count loads: a
  // branch there would hold the kernel-argument, sweep-counter and table
  // loads behind it, one round trip each.  A lane group past the last
  // active cell or in the padding of its replica's range loads no count
  // and leaves with them (nc = 0).
  const int hz2 = rp.gz >> 1, hy2 = rp.gy >> 1;
  const int la = a - rp.aoff;
  const bool live = a < A && la < hz2 * hy2 * (rp.gx >> 1);
  const int iz2 = la % hz2;
  const int t2 = la / hz2;
  const int iy2 = t2 % hy2;
  const int ix2 = t2 / hy2;
  const int cx = 2 * ix2 + (colour & 1);
  const int cy = 2 * iy2 + ((colour >> 1) & 1);
  const int cz = 2 * iz2 + ((colour >> 2) & 1);
  const int cid = (cx * rp.gy + cy) * rp.gz + cz;

  // Loads are issued ahead of any work that does not need their values, and
  // without a branch around them: the compiler keeps the order of these
  // lane-masked blocks, and a value loaded inside one is waited for before
  // the block ends.  Lane j < 27 owns neighbour cell (j/9-1, (j/3)%3-1,
  // j%3-1), 13 being the active cell itself; the other lanes read the count
  // of cell 0 and discard it.
  const int sdx = lane / 9 - 1, sdy = (lane / 3) % 3 - 1, sdz = lane % 3 - 1;
  const bool owner = lane < 27 && live && stencil_axis_ok(sdx, rp.gx) &&
                     stencil_axis_ok(sdy, rp.gy) && stencil_axis_ok(sdz, rp.gz);
  const int ncell =
      owner ? rp.coff + (wrap_index(cx + sdx, rp.gx) * rp.gy + wrap_index(cy + sdy, rp.gy)) * rp.gz +
                  wrap_index(cz + sdz, rp.gz)
            : 0;
  const int n_loaded = cntb[ncell];

  // draw block of attempt t = lane for every attempt a cell can make (t <
  // CAP), computed while the counts are in flight
  uint32_t w[4] = {0u, 0u, 0u, 0u};
  if (lane < CAP)
    philox4x32_10((uint32_t)cid, (uint32_t)lane, (uint32_t)sweep, (uint32_t)(sweep >> 32), rp.k0,
                  rp.k1, w);

  const int nj = owner ? (n_loaded < CAP ? n_loaded : CAP) : 0;
  const int incl = group_inclusive_sum<LPC>(nj);
  const int off = incl - nj;
  const int T = __shfl(incl, LPC - 1, LPC);
  const int nc = __shfl(nj, 13, LPC);
  const int oc = __shfl(off, 13, LPC);
  if (nc == 0) return;  // empty active cell: nothing to move
  if (T > LCAP) {        // cannot happen below the host's packing bound; reported
    if (lane == 0) atomicOr(err, kErrListFull);
    return;
  }

  // the active cell's occupants (slot t of the active cell for lane t < nc,
  // slot 0 for the other lanes: no slot map needed), so their ranking and the
  // attempt preparation below overlap the neighbour loads
  const SlotRec *orec = slot + (size_t)(lane < nc ? lane : 0) * C + rp.coff + cid;
  const int my_id = orec->id;
  const double o_x = orec->x, o_y = orec->y, o_z = orec->z;
  // stage the stencil list: the slot of every entry into LDS, then each
  // lane loads its entries' FP64 positions into registers, all loads of the
  // work item in flight together.  Entries past T read entry 0's record (T
  // >= nc >= 1 here); copies past T are never tested.
  for (int q = 0; q < nj; ++q) Lm[off + q] = q * C + ncell;
  wave_lds_sync();
  dvec<M> px, py, pz;
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const int q = lane + j * LPC;
    const SlotRec *rec = slot + Lm[q < T ? q : 0];
    px[j] = rec->x;
    py[j] = rec->y;
    pz[j] = rec->z;
  }

  // while the records are in flight, lane t < nc draws attempt t: occupant
  // rank (rng_get_int rule) and the displacement (r - 0.5) * dr_max per axis
  uint32_t my_k = 0;
  double my_dx = 0.0, my_dy = 0.0, my_dz = 0.0;
  if (lane < nc) {
    my_k = occupant_pick(w[0], (uint32_t)nc, (uint32_t)cid, (uint32_t)lane, sweep, rp.k0, rp.k1);
    my_dx = (u32_unit_closed(w[1]) - 0.5) * rp.dr_max;
    my_dy = (u32_unit_closed(w[2]) - 0.5) * rp.dr_max;
    my_dz = (u32_unit_closed(w[3]) - 0.5) * rp.dr_max;
  }
  if (lane < nc) {
    Sp[lane] = my_id;
    if constexpr (LPC != 64) {
      Op[3 * lane] = o_x;
      Op[3 * lane + 1] = o_y;
      Op[3 * lane + 2] = o_z;
    }
  }
  wave_lds_sync();
  // occupants of the active cell in ascending particle id; attempt t moves
  // list entry oc + Lo[k_t]
  if (lane < nc) {
    const int me = Sp[lane];
    int rank = 0;
    for (int m = 0; m < nc; ++m) rank += (Sp[m] < me) ? 1 : 0;
    Lo[rank] = lane;
  }
  wave_lds_sync();
  // 32-lane groups: lane t < nc prepares attempt t before the serial loop
  // (its sphere's position, the trial position, whether it stays in the
  // cell), so each attempt is one broadcast.  Whole-wave groups broadcast
  // with scalar reads, cheap enough to fetch the sphere at each attempt.
  int my_e = 0, my_stay = 0;
  double my_nx = 0.0, my_ny = 0.0, my_nz = 0.0;
  auto trial = [&](double bx, double by, double bz, double dx, double dy, double dz) {
    my_nx = wrap_once(bx + dx, rp.lx);
    my_ny = wrap_once(by + dy, rp.ly);
    my_nz = wrap_once(bz + dz, rp.lz);
    // all three axes are evaluated (no short circuit)
    my_stay = (int)(shifted_cell(my_nx, shx, rp.lx, rp.ix, rp.gx) == cx) &
              (int)(shifted_cell(my_ny, shy, rp.ly, rp.iy, rp.gy) == cy) &
              (int)(shifted_cell(my_nz, shz, rp.lz, rp.iz, rp.gz) == cz);
  };
  // one replica (few cells in flight, registers to spare): lane t also
  // stores its sphere when attempt t is accepted, its slot record (slot o
  // of the active cell) and by-id position, instead of a write-back after
  // the loop
  int my_slot = 0, my_g = 0;
  if (lane < nc) {
    const int o = Lo[my_k];
    my_e = oc + o;
    if constexpr (ONE) {
      my_slot = o * C + rp.coff + cid;  // below cap * C < 2^31
      my_g = rp.poff + Sp[o];
    }
    if constexpr (LPC != 64) {
      Od[3 * lane] = my_dx;  // kept for a repeated pick, off the registers
      Od[3 * lane + 1] = my_dy;
      Od[3 * lane + 2] = my_dz;
      trial(Op[3 * o], Op[3 * o + 1], Op[3 * o + 2], my_dx, my_dy, my_dz);
    }
  }

  // Several replicas: whether every pair this wave tests is free of the box
  // edge.  A cell index k <= g - 2 of the shifted grid holds positions
  // x = t + shift with no wrap (t < (g - 1) * side <= L - shift), so an
  // active cell with every index in [1, g - 3] and its 26 neighbours hold
  // positions within 3 sides of each other, less than half the box when
  // g >= 8: there compute_dist's minimum image (moves.c:414-427) changes no
  // difference.  Trial positions are tested only when they stay in the
  // active cell.  The multi-replica launches keep 6-7 waves per SIMD and
  // gain from the shorter test; one replica's 2 waves per SIMD do not.
  const bool inner_cell = rp.gx >= 8 && rp.gy >= 8 && rp.gz >= 8 && cx >= 1 && cx <= rp.gx - 3 &&
                          cy >= 1 && cy <= rp.gy - 3 && cz >= 1 && cz <= rp.gz - 3;
  const bool inner = !ONE && __ballot(!inner_cell) == 0ull;  // the lanes still here
  unsigned int n_acc = 0, n_left = 0;
  unsigned int moved = 0u;  // several replicas: bit j, this lane's copy j moved
  for (int t = 0; t < nc; ++t) {
    unsigned e;
    double nx, ny, nz;
    int stays;
    if constexpr (LPC == 64) {
      // attempt t: its draw, then the moving sphere's current position
      // from the lane holding it
      e = (unsigned)__builtin_amdgcn_readlane(my_e, t);
      const double ddx = wave_lane(my_dx, t), ddy = wave_lane(my_dy, t), ddz = wave_lane(my_dz, t);
      const int o0 = (int)(e % LPC), j0 = (int)(e / LPC);
      nx = wrap_once(wave_lane((double)px[j0], o0) + ddx, rp.lx);
      ny = wrap_once(wave_lane((double)py[j0], o0) + ddy, rp.ly);
      nz = wrap_once(wave_lane((double)pz[j0], o0) + ddz, rp.lz);
      stays = (int)(shifted_cell(nx, shx, rp.lx, rp.ix, rp.gx) == cx) &
              (int)(shifted_cell(ny, shy, rp.ly, rp.iy, rp.gy) == cy) &
              (int)(shifted_cell(nz, shz, rp.lz, rp.iz, rp.gz) == cz);
    } else {
      // attempt t as lane t prepared it
      e = (unsigned)__shfl(my_e, t, LPC);
      stays = __shfl(my_stay, t, LPC);
      nx = __shfl(my_nx, t, LPC);
      ny = __shfl(my_ny, t, LPC);
      nz = __shfl(my_nz, t, LPC);
    }
    if (!stays) {
      ++n_left;
      continue;
    }
    const int owner = (int)(e % LPC), je = (int)(e / LPC);
    bool ov = false;
    if (inner) {
      // no pair of this wave crosses the box edge: the minimum image leaves
      // every difference as it is, so it is left out (same FP64 operations)
#pragma unroll
      for (int j = 0; j < M; ++j) {
        const int q = lane + j * LPC;
        if (q < T) {
          const double dx = nx - px[j], dy = ny - py[j], dz = nz - pz[j];
          const bool hit = dx * dx + dy * dy + dz * dz < 1.0;
          ov |= hit & (q != (int)e);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < M; ++j) {
        const int q = lane + j * LPC;
        if (q < T) {
          const bool hit =
              dist2(nx, ny, nz, px[j], py[j], pz[j], rp.lx, rp.ly, rp.lz, rp.hx, rp.hy, rp.hz) < 1.0;
          ov |= hit & (q != (int)e);
        }
      }
    }
    const unsigned long long bal = __ballot(ov);
    const unsigned long long mine =
        (LPC == 64) ? bal : ((bal >> (threadIdx.x & 32)) & 0xFFFFFFFFull);
    if (mine == 0ull) {
      if (lane == owner) {
        px[je] = nx;
        py[je] = ny;
        pz[je] = nz;
        if constexpr (!ONE) moved |= 1u << je;
      }
      // write the move through to the device-coherent level at once (no
      // dirty L2 lines are left for the end-of-kernel write-back); a later
      // move of the same sphere stores after it
      if (ONE && lane == t) {
        SlotRec *rec = slot + my_slot;
        store_through(&rec->x, nx);
        store_through(&rec->y, ny);
        store_through(&rec->z, nz);
        store_through(&x[my_g], nx);
        store_through(&y[my_g], ny);
        store_through(&z[my_g], nz);
      }
      ++n_acc;
      if constexpr (LPC != 64) {
        // a later attempt on the same sphere starts from its new position
        if (lane > t && lane < nc && my_e == (int)e)
          trial(nx, ny, nz, Od[3 * lane], Od[3 * lane + 1], Od[3 * lane + 2]);
      }
    }
  }
  if constexpr (!ONE) {
    // write the spheres that moved back (slot copy and by-id arrays),
    // through to the device-coherent level
#pragma unroll
    for (int j = 0; j < M; ++j) {
      if ((moved >> j) & 1u) {
        const int q = lane + j * LPC;
        SlotRec *rec = slot + Lm[q];
        store_through(&rec->x, (double)px[j]);
        store_through(&rec->y, (double)py[j]);
        store_through(&rec->z, (double)pz[j]);
        const int g = rp.poff + Sp[q - oc];
        store_through(&x[g], (double)px[j]);
        store_through(&y[g], (double)py[j]);
        store_through(&z[g], (double)pz[j]);
      }
    }
  }
  if constexpr (ONE && FILE) {
    // the occupants (list entries oc .. oc + nc - 1) at their final
    // positions: the nc <= CAP <= LPC entries lie in distinct lanes, so a
    // lane has at most one, picked by selects
    int jsel = -1;
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const int q = lane + j * LPC;
      if (q >= oc && q < oc + nc) jsel = j;
    }
    double fx = px[0], fy = py[0], fz = pz[0];
#pragma unroll
    for (int j = 1; j < M; ++j) {
      fx = j == jsel ? (double)px[j] : fx;
      fy = j == jsel ? (double)py[j] : fy;
      fz = j == jsel ? (double)pz[j] : fz;
    }
    req.x = fx;
    req.y = fy;
    req.z = fz;
    req.has = jsel >= 0;
    req.id = jsel >= 0 ? Sp[lane + jsel * LPC - oc] : 0;
  }
  // move counters per active-cell index (the cells of a workgroup share a
  // cache line); the adds never contend and their results are not waited for
  if (lane == 13) {
    if (n_acc) atomicAdd(&cacc[a], n_acc);
    if (n_left) atomicAdd(&cleft[a], n_left);
  }
}

// The filing of a one-replica last phase (MODE 2: the next sweep has no
// rebin).  Lane g of the launch files particle g, unless it lies in one of
// this phase's active cells, and the occupant its cell gave it (req).  An
// active cell's spheres never leave it (a move out is rejected, axis by
// axis), so particle g's position, read whether or not its cell has moved it
// yet, gives the same answer to that test.  Both atomics of a lane are
// issued before either result is used, so a lane waits one round trip.
template <int CAP>
__device__ __forceinline__ void file_next_sweep(int g, int P, int C, int phase, const RepDev &rp,
                                                const SweepDev &cur, const SweepDev &nxt,
                                                double ax, double ay, double az,
                                                const FileReq &req, unsigned long long sweep,
                                                int *__restrict__ cnt,
                                                SlotRec *__restrict__ slot_all,
                                                unsigned int *__restrict__ err) {
  const int par = (int)(sweep & 1ull);
  int *cntn = cnt + (size_t)(par ^ 1) * C;
  SlotRec *slotn = slot_all + (size_t)(par ^ 1) * CAP * C;
  const int colour = (int)((cur.order >> (4 * phase)) & 0xFu);
  const int cx = shifted_cell(ax, cur.shx, rp.lx, rp.ix, rp.gx);
  const int cy = shifted_cell(ay, cur.shy, rp.ly, rp.iy, rp.gy);
  const int cz = shifted_cell(az, cur.shz, rp.lz, rp.iz, rp.gz);
  const bool active = (((cx ^ colour) | (cy ^ (colour >> 1)) | (cz ^ (colour >> 2))) & 1) == 0;
  const bool af = g < P && !active;
  auto next_cell = [&](double px, double py, double pz) {
    return rp.coff +
           (shifted_cell(px, nxt.shx, rp.lx, rp.ix, rp.gx) * rp.gy +
            shifted_cell(py, nxt.shy, rp.ly, rp.iy, rp.gy)) * rp.gz +
           shifted_cell(pz, nxt.shz, rp.lz, rp.iz, rp.gz);
  };
  const int ca = next_cell(ax, ay, az), ct = next_cell(req.x, req.y, req.z);
  int s1 = 0, s2 = 0;
  if (af || req.has) {
    s1 = atomicAdd(&cntn[af ? ca : ct], 1);
    if (af && req.has) s2 = atomicAdd(&cntn[ct], 1);
  }
  const int sa = s1, st = af ? s2 : s1;
  auto put = [&](int sl, int c, double px, double py, double pz, int id) {
    if (sl < CAP) {
      SlotRec rec;
      rec.x = px;
      rec.y = py;
      rec.z = pz;
      rec.id = id;
      rec.pad = 0;
      slotn[(size_t)sl * C + c] = rec;
    } else {
      atomicOr(err, kErrOverflow);
    }
  };
  if (af) put(sa, ca, ax, ay, az, g - rp.poff);
  if (req.has) put(st, ct, req.x, req.y, req.z, req.id);
}

// One sublattice phase: each workgroup takes cells_per_block active cells.
// One replica: the sweep's table entry is entry kt (its offset in a chained
// batch, 0 for a sweep with its own rebin) and the next sweep's entry
// kt + 1.  MODE 1 (the first phase of a sweep with
// no rebin): lane g also clears the next sweep's count of cell g, the
// rebin's clear.  MODE 2 (the last phase of a sweep whose successor has no
// rebin): lane g files particle g and its cell's occupant into the next
// sweep's list.  The host checks that the lanes cover C and P.  Each mode is
// its own kernel, so the plain phases carry none of this code.
// Occupancy: at least 6 waves per SIMD (at most 80 VGPRs) for stencil lists
// of up to 192 entries, 4 for the longest lists and for one replica (whose
// few active cells never fill the device).
template <int CAP, int LPC, int LCAP, bool ONE, int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(!ONE && LCAP <= 192 ? 6 : 4)))
k_phase(const unsigned long long *__restrict__ sweep_base, const SweepDev *__restrict__ swp,
        const RepDev *__restrict__ reps, const RepDev rp0, const int *__restrict__ arep, int R,
        int P, int A, int phase, int k, int kt,
        double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
        int *__restrict__ cnt, int C, SlotRec *__restrict__ slot,
        unsigned int *__restrict__ cacc, unsigned int *__restrict__ cleft,
        unsigned int *__restrict__ err) {
  extern __shared__ __attribute__((aligned(16))) int smem_i[];
  SweepDev sw0{}, swn0{};
  if constexpr (ONE) sw0 = swp[kt];
  if constexpr (ONE && MODE == 2) swn0 = swp[kt + 1];
  const unsigned long long sweep = *sweep_base + (unsigned long long)k;
  const int cpb = blockDim.x / LPC;
  const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  double ax = 0.0, ay = 0.0, az = 0.0;
  if constexpr (ONE && MODE == 2) {  // particle g's position, loaded first
    const int gp = g < P ? g : 0;
    ax = x[gp];
    ay = y[gp];
    az = z[gp];
  }
  const int lane = threadIdx.x & (LPC - 1);
  const int sub = threadIdx.x / LPC;
  FileReq req{};
  phase_cell<CAP, LPC, LCAP, ONE, MODE == 2>(blockIdx.x * cpb + sub, blockIdx.x * cpb, lane,
                                             smem_i + (size_t)sub * (LCAP + 14 * CAP), reps, rp0,
                                             swp, sw0, arep, R, A, phase, sweep, x, y, z, cnt, C,
                                             slot, cacc, cleft, err, req);
  if constexpr (ONE && MODE == 1) {
    if (g < C) cnt[(size_t)((sweep & 1ull) ^ 1ull) * C + g] = 0;
  }
  if constexpr (ONE && MODE == 2)
    file_next_sweep<CAP>(g, P, C, phase, rp0, sw0, swn0, ax, ay, az, req, sweep, cnt, slot, err);
}

// Barrier of a cooperative launch (every workgroup resident): an arrival
// counter and a generation word.  The wait is bounded: a workgroup that
// waits too long, or sees another's stall, sets kErrStall and the barrier
// returns false, after which every workgroup leaves the kernel, so a launch
// always drains.
struct GridBarrier {
  unsigned count, gen;
};

__device__ bool grid_barrier(GridBarrier *__restrict__ b, unsigned *__restrict__ err) {
  __shared__ int ok;
  __syncthreads();
  if (threadIdx.x == 0) {
    ok = 1;
    __threadfence();  // this workgroup's stores before its arrival
    const unsigned g = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned arrived =
        __hip_atomic_fetch_add(&b->count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    if (arrived == gridDim.x - 1) {
      __hip_atomic_store(&b->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(&b->gen, g + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      unsigned spins = 0;
      while (__hip_atomic_load(&b->gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g) {
        if (++spins > (1u << 22) ||
            (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kErrStall)) {
          atomicOr(err, kErrStall);
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __threadfence();  // the other workgroups' stores before this one goes on
  }
  __syncthreads();
  return ok != 0;
}

// Persistent sweeps: one cooperative launch runs n_sweeps whole sweeps, each
// the rebin's work (every thread of the grid strides over particles and
// cells) and the 8 sublattice phases (each workgroup strides over its share
// of the active cells), with a grid barrier after each of the 9 steps in
// place of a kernel boundary.  The result is the same as the launch-per-step
// sequence: the same work items, the same order of steps, the same Philox
// streams.
template <int CAP, int LPC, int LCAP, bool ONE>
__global__ void __launch_bounds__(256)
k_sweeps_persistent(const RepDev *__restrict__ reps, const RepDev rp0, SweepDev *__restrict__ swp,
                    const int *__restrict__ arep, const int *__restrict__ prep, int R, int P, int C,
                    int A, double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
                    int *__restrict__ cnt, SlotRec *__restrict__ slot, int cap,
                    const unsigned long long *__restrict__ sweep_base, int n_sweeps,
                    unsigned int *__restrict__ cacc, unsigned int *__restrict__ cleft,
                    unsigned int *__restrict__ err, GridBarrier *__restrict__ bar) {
  extern __shared__ __attribute__((aligned(16))) int smem_i[];
  const int cpb = blockDim.x / LPC;
  const int lane = threadIdx.x & (LPC - 1);
  const int sub = threadIdx.x / LPC;
  int *Lm = smem_i + (size_t)sub * (LCAP + 14 * CAP);
  const int work = P > C ? P : C;
  const int gtid = blockIdx.x * blockDim.x + threadIdx.x, gstride = gridDim.x * blockDim.x;
  const unsigned long long base = *sweep_base;
  for (int s = 0; s < n_sweeps; ++s) {
    const unsigned long long sweep = base + (unsigned long long)s;
    for (int i = gtid; i < work; i += gstride)
      rebin_one<ONE>(i, reps, rp0, swp, R, s, 1, prep, P, C, x, y, z, cnt, slot, cap, sweep, err);
    if (!grid_barrier(bar, err)) return;
    for (int p = 0; p < 8; ++p) {
      for (int a0 = blockIdx.x * cpb; a0 < A; a0 += gridDim.x * cpb) {
        wave_lds_sync();  // the group's LDS region is reused cell after cell
        const SweepDev sw0 = ONE ? swp[s] : SweepDev{};
        FileReq req{};
        phase_cell<CAP, LPC, LCAP, ONE, false>(a0 + sub, a0, lane, Lm, reps, rp0, swp, sw0, arep,
                                               R, A, p, sweep, x, y, z, cnt, C, slot, cacc, cleft,
                                               err, req);
      }
      if (!grid_barrier(bar, err)) return;
    }
  }
}

__global__ void k_advance(unsigned long long *sweep_base, int n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *sweep_base += (unsigned long long)n;
}

// The phase kernel's occupant pick for attempts t < count of one cell, any n.
__global__ void __launch_bounds__(256)
k_occupant_picks(uint32_t k0, uint32_t k1, uint32_t cell, unsigned long long sweep, uint32_t n,
                 int count, uint32_t *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  uint32_t w[4];
  philox4x32_10(cell, (uint32_t)t, (uint32_t)sweep, (uint32_t)(sweep >> 32), k0, k1, w);
  out[t] = occupant_pick(w[0], n, cell, (uint32_t)t, sweep, k0, k1);
}

using PhaseFn = void (*)(const unsigned long long *, const SweepDev *, const RepDev *, const RepDev,
                         const int *, int, int, int, int, int, int, double *,
                         double *, double *, int *, int, SlotRec *, unsigned int *, unsigned int *,
                         unsigned int *);

template <int CAP, int LCAP, bool ONE, int MODE>
PhaseFn phase_fn_l(int lpc) {
  return lpc == 64 ? k_phase<CAP, 64, LCAP, ONE, MODE> : k_phase<CAP, 32, LCAP, ONE, MODE>;
}

template <bool ONE, int MODE>
PhaseFn phase_fn_o(int cap, int lpc, int lcap) {
  if (cap == 8) {
    if (lcap <= 128) return phase_fn_l<8, 128, ONE, MODE>(lpc);
    return lcap <= 192 ? phase_fn_l<8, 192, ONE, MODE>(lpc) : phase_fn_l<8, 256, ONE, MODE>(lpc);
  }
  if (lcap <= 128) return phase_fn_l<16, 128, ONE, MODE>(lpc);
  if (lcap <= 192) return phase_fn_l<16, 192, ONE, MODE>(lpc);
  return phase_fn_l<16, 256, ONE, MODE>(lpc);
}

// mode 1 and 2 (the first and last phases of chained sweeps) exist for one
// replica only
PhaseFn phase_fn(int cap, int lpc, int lcap, bool one, int mode = 0) {
  if (!one) return phase_fn_o<false, 0>(cap, lpc, lcap);
  return mode == 1   ? phase_fn_o<true, 1>(cap, lpc, lcap)
         : mode == 2 ? phase_fn_o<true, 2>(cap, lpc, lcap)
                     : phase_fn_o<true, 0>(cap, lpc, lcap);
}

using PersistentFn = void (*)(const RepDev *, const RepDev, SweepDev *, const int *, const int *,
                              int, int, int, int, double *, double *, double *, int *, SlotRec *,
                              int, const unsigned long long *, int, unsigned int *,
                              unsigned int *, unsigned int *, GridBarrier *);

template <int CAP, int LCAP, bool ONE>
PersistentFn persistent_fn_l(int lpc) {
  return lpc == 64 ? k_sweeps_persistent<CAP, 64, LCAP, ONE> : k_sweeps_persistent<CAP, 32, LCAP, ONE>;
}

template <bool ONE>
PersistentFn persistent_fn_o(int cap, int lpc, int lcap) {
  if (cap == 8) {
    if (lcap <= 128) return persistent_fn_l<8, 128, ONE>(lpc);
    return lcap <= 192 ? persistent_fn_l<8, 192, ONE>(lpc) : persistent_fn_l<8, 256, ONE>(lpc);
  }
  if (lcap <= 128) return persistent_fn_l<16, 128, ONE>(lpc);
  if (lcap <= 192) return persistent_fn_l<16, 192, ONE>(lpc);
  return persistent_fn_l<16, 256, ONE>(lpc);
}

PersistentFn persistent_fn(int cap, int lpc, int lcap, bool one) {
  return one ? persistent_fn_o<true>(cap, lpc, lcap) : persistent_fn_o<false>(cap, lpc, lcap);
}

__global__ void k_zero_barrier(GridBarrier *b) {
  if (threadIdx.x == 0) {
    b->count = 0u;
    b->gen = 0u;
  }
}

// One cooperative launch of n sweeps: as many workgroups as the device
// holds at once (the launch fails rather than oversubscribe), each striding
// over its share of particles and active cells; then the counter advance.
int launch_persistent(hsmc_gpu_ctx *ctx, int n) {
  const int lpc = ctx->tune.lanes_per_cell, cpb = ctx->tune.cells_per_block;
  const size_t lds = phase_lds_bytes(ctx->cap, ctx->lcap, cpb);
  PersistentFn fn = persistent_fn(ctx->cap, lpc, ctx->lcap, ctx->R == 1);
  int per_cu = 0;
  HSMC_HIP(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn, lpc * cpb, lds));
  const long need = (ctx->Atot + cpb - 1) / cpb;
  const long grid = std::min(need, (long)per_cu * ctx->n_cu);
  if (grid < 1) {
    ctx->last_error = "the persistent sweep kernel does not fit on a compute unit";
    return HSMC_ERR_NOT_SUPPORTED;
  }
  // the barrier starts empty (a launch that stopped at a stalled barrier
  // leaves it partly filled); cleared by a kernel, not a memset
  hipLaunchKernelGGL(k_zero_barrier, dim3(1), dim3(64), 0, ctx->stream,
                     reinterpret_cast<GridBarrier *>(ctx->d_bar));
  const RepDev *reps = ctx->d_reps;
  RepDev rp0 = ctx->rep0;
  SweepDev *swp = ctx->d_swp;
  const int *arep = ctx->d_arep, *prep = ctx->d_prep;
  int R = ctx->R, P = (int)ctx->Ptot, C = (int)ctx->Ctot, A = (int)ctx->Atot, cap = ctx->cap;
  double *x = ctx->d_x, *y = ctx->d_y, *z = ctx->d_z;
  int *cnt = ctx->d_cnt;
  SlotRec *slot = ctx->d_slot;
  const unsigned long long *sweep = ctx->d_sweep;
  int ns = n;
  unsigned int *cacc = ctx->d_cacc, *cleft = ctx->d_cleft, *err = ctx->d_err;
  GridBarrier *bar = reinterpret_cast<GridBarrier *>(ctx->d_bar);
  void *args[] = {&reps, &rp0, &swp, &arep, &prep, &R, &P, &C, &A, &x, &y, &z, &cnt, &slot,
                  &cap, &sweep, &ns, &cacc, &cleft, &err, &bar};
  HSMC_HIP(ctx, hipLaunchCooperativeKernel((const void *)fn, dim3((unsigned)grid), dim3(lpc * cpb),
                                           args, (unsigned)lds, ctx->stream));
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, ctx->stream, ctx->d_sweep, n);
  HSMC_HIP(ctx, hipGetLastError());
  return HSMC_OK;
}

bool tuning_ok(const hsmc_gpu_ctx *ctx, const hsmc_gpu_tuning &t) {
  if (t.lanes_per_cell != 32 && t.lanes_per_cell != 64) return false;
  // 1, 2, 4 or 8 cells per workgroup: a divisor of the padding of each
  // replica's active-cell range
  if ((t.cells_per_block & (t.cells_per_block - 1)) != 0 || t.cells_per_block < 1 ||
      t.cells_per_block > 8 || t.lanes_per_cell * t.cells_per_block > 256)
    return false;
  if (phase_lds_bytes(ctx->cap, ctx->lcap, t.cells_per_block) > 64 * 1024) return false;
  if (t.sweeps_per_graph < 0 || t.sweeps_per_graph > 512) return false;
  if (t.persistent != 0 && t.persistent != 1) return false;
  return true;
}

// Enqueue the kernels of sweep k of the current batch on stream s.  With
// `chained` set, the sweep's rebin is left out (the previous sweep's last
// phase filed its cell list) and its first phase clears the next sweep's
// counts in the rebin's place; with `file_next` set, its last phase files
// the next sweep's cell list.  A one-replica rebin draws the table entries
// of nsw sweeps from this one (k .. k + nsw - 1: the chained sweeps after
// it read theirs there); a sweep outside a chained run uses entry 0.
void enqueue_sweep(hsmc_gpu_ctx *ctx, int k, hipStream_t s, hipEvent_t *phase_events,
                   bool chained = false, bool file_next = false, int nsw = 1) {
  const long work = std::max(std::max(ctx->Ptot, ctx->Ctot), (long)nsw);
  const int kt = chained || file_next || nsw > 1 ? k : 0;  // below kSweepTable - 1
  if (!chained)
    hipLaunchKernelGGL(ctx->R == 1 ? k_rebin<true> : k_rebin<false>,
                       dim3((unsigned)((work + 255) / 256)), dim3(256), 0, s, ctx->d_reps,
                       ctx->rep0, ctx->d_swp, ctx->R, kt, nsw, ctx->d_prep, (int)ctx->Ptot,
                       (int)ctx->Ctot, ctx->d_x, ctx->d_y, ctx->d_z, ctx->d_cnt, ctx->d_slot,
                       ctx->cap, ctx->d_sweep, k, ctx->d_err);
  const int lpc = ctx->tune.lanes_per_cell, cpb = ctx->tune.cells_per_block;
  const int bs = lpc * cpb;
  const unsigned blocks = (unsigned)((ctx->Atot + cpb - 1) / cpb);
  const size_t lds = phase_lds_bytes(ctx->cap, ctx->lcap, cpb);
  for (int p = 0; p < 8; ++p) {
    const int mode = p == 0 && chained ? 1 : p == 7 && file_next ? 2 : 0;
    PhaseFn fn = phase_fn(ctx->cap, lpc, ctx->lcap, ctx->R == 1, mode);
    if (phase_events && p == 0) (void)hipEventRecord(phase_events[0], s);
    hipLaunchKernelGGL(fn, dim3(blocks), dim3(bs), lds, s, ctx->d_sweep, ctx->d_swp,
                       ctx->d_reps, ctx->rep0, ctx->d_arep, ctx->R, (int)ctx->Ptot,
                       (int)ctx->Atot, p, k, kt, ctx->d_x, ctx->d_y, ctx->d_z,
                       ctx->d_cnt, (int)ctx->Ctot, ctx->d_slot, ctx->d_cacc,
                       ctx->d_cleft, ctx->d_err);
    if (phase_events && p == 7) (void)hipEventRecord(phase_events[15], s);
  }
}

// Whether sweeps can run chained (no rebin; the previous sweep's last phase
// files the list): one replica, and at least as many phase lanes as cells
// (whose next counts they clear) and particles (which they file).
bool can_chain(const hsmc_gpu_ctx *ctx) {
  const int lpc = ctx->tune.lanes_per_cell, cpb = ctx->tune.cells_per_block;
  const long lanes = (ctx->Atot + cpb - 1) / cpb * lpc * cpb;
  return ctx->R == 1 && lanes >= ctx->Ctot && lanes >= ctx->Ptot;
}

// Capture K sweeps (and the counter advance) into one graph: the main graph
// (a batch of sweeps_per_graph sweeps) or, with remainder set, the graph of
// a call's remaining n mod sweeps_per_graph sweeps.
int capture_graph(hsmc_gpu_ctx *ctx, int K, bool remainder = false) {
  hipGraph_t &graph = remainder ? ctx->graph_r : ctx->graph;
  hipGraphExec_t &gexec = remainder ? ctx->gexec_r : ctx->gexec;
  if (gexec) (void)hipGraphExecDestroy(gexec);
  if (graph) (void)hipGraphDestroy(graph);
  gexec = nullptr;
  graph = nullptr;
  (remainder ? ctx->graph_r_sweeps : ctx->graph_sweeps) = 0;
  HSMC_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  // one replica: the batch's first sweep rebins and each later one takes
  // the cell list its predecessor's last phase filed (with several replicas
  // the filing's registers would lower the phase kernel's occupancy)
  const bool chain = can_chain(ctx);
  for (int k = 0; k < K; ++k)
    enqueue_sweep(ctx, k, ctx->stream, nullptr, chain && k > 0, chain && k + 1 < K,
                  chain && k == 0 ? K : 1);
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, ctx->stream, ctx->d_sweep, K);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(ctx->stream, &g);
  if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
  graph = g;
  e = hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(ctx);
    return hip_fail(ctx, e, "hipGraphInstantiate");
  }
  if (remainder) {
    ctx->graph_r_sweeps = K;
    ctx->graph_r_tune = ctx->tune;
  } else {
    ctx->graph_sweeps = K;
    ctx->graph_tune = ctx->tune;
  }
  return HSMC_OK;
}

bool same_launch_shape(const hsmc_gpu_tuning &a, const hsmc_gpu_tuning &b) {
  return a.lanes_per_cell == b.lanes_per_cell && a.cells_per_block == b.cells_per_block;
}


}  // namespace

namespace hsmc {
// After each batch of sweeps the device's error bits are copied to pinned
// host memory without waiting; a batch is launched only while the copies
// that have landed show no error (a cell overflow or a full candidate list
// leaves sweeps running on an incomplete cell list).  The next
// synchronising call reports the error.
int run_sweeps(hsmc_gpu_ctx *ctx, int n) {
  if (n <= 0) return HSMC_OK;
  if (*ctx->h_err) return check_device_errors(ctx);
  ctx->version++;
  if (ctx->tune.persistent) {  // cooperative launches of whole sweeps
    const int per = ctx->tune.sweeps_per_graph > 0 ? ctx->tune.sweeps_per_graph : 64;
    for (int done = 0; done < n; done += per) {
      if (*ctx->h_err) return check_device_errors(ctx);
      int st = launch_persistent(ctx, std::min(per, n - done));
      if (st) return st;
      ctx->swp_last = std::min(per, n - done) - 1;
      HSMC_HIP(ctx, hipMemcpyAsync((void *)ctx->h_err, ctx->d_err, sizeof(unsigned int),
                                   hipMemcpyDeviceToHost, ctx->stream));
    }
    ctx->cells_version = ctx->version;
    return HSMC_OK;
  }
  const int K = ctx->tune.sweeps_per_graph;
  int done = 0;
  if (K > 0 && n >= K) {
    if (!ctx->gexec || ctx->graph_sweeps != K || !same_launch_shape(ctx->graph_tune, ctx->tune)) {
      int st = capture_graph(ctx, K);
      if (st) return st;
    }
    for (; done + K <= n; done += K) {
      if (*ctx->h_err) return check_device_errors(ctx);
      HSMC_HIP(ctx, hipGraphLaunch(ctx->gexec, ctx->stream));
      HSMC_HIP(ctx, hipMemcpyAsync((void *)ctx->h_err, ctx->d_err, sizeof(unsigned int),
                                   hipMemcpyDeviceToHost, ctx->stream));
      ctx->swp_last = can_chain(ctx) ? K - 1 : 0;  // the last sweep's table entry
    }
  }
  const int rem = n - done;
  if (rem > 0) {
    if (*ctx->h_err) return check_device_errors(ctx);
    if (K > 0 && ctx->gexec_r && ctx->graph_r_sweeps == rem &&
        same_launch_shape(ctx->graph_r_tune, ctx->tune)) {
      HSMC_HIP(ctx, hipGraphLaunch(ctx->gexec_r, ctx->stream));  // captured by a timing call
    } else {
      const bool chain = can_chain(ctx) && rem < kSweepTable;
      for (int k = 0; k < rem; ++k)
        enqueue_sweep(ctx, k, ctx->stream, nullptr, chain && k > 0, chain && k + 1 < rem,
                      chain && k == 0 ? rem : 1);
      hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, ctx->stream, ctx->d_sweep, rem);
      HSMC_HIP(ctx, hipGetLastError());
    }
    HSMC_HIP(ctx, hipMemcpyAsync((void *)ctx->h_err, ctx->d_err, sizeof(unsigned int),
                                 hipMemcpyDeviceToHost, ctx->stream));
    ctx->swp_last = can_chain(ctx) && rem < kSweepTable ? rem - 1 : 0;
  }
  ctx->cells_version = ctx->version;  // the last sweep's cells hold every sphere
  return HSMC_OK;
}

int rebin_last_sweep(hsmc_gpu_ctx *ctx) {
  // the last sweep done is *d_sweep - 1; its count buffer is refilled from
  // zero and the rebin clears the other buffer as a sweep's rebin does
  unsigned long long next = 0;
  HSMC_HIP(ctx, hipMemcpyAsync(&next, ctx->d_sweep, sizeof(next), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const int buf = (int)((next - 1ull) & 1ull);
  HSMC_HIP(ctx, hipMemsetAsync(ctx->d_cnt + (size_t)buf * ctx->Ctot, 0, sizeof(int) * ctx->Ctot,
                               ctx->stream));
  const long work = std::max(ctx->Ptot, ctx->Ctot);
  hipLaunchKernelGGL(ctx->R == 1 ? k_rebin<true> : k_rebin<false>,
                     dim3((unsigned)((work + 255) / 256)), dim3(256), 0, ctx->stream,
                     ctx->d_reps, ctx->rep0, ctx->d_swp, ctx->R, 0, 1, ctx->d_prep, (int)ctx->Ptot,
                     (int)ctx->Ctot, ctx->d_x, ctx->d_y, ctx->d_z, ctx->d_cnt, ctx->d_slot,
                     ctx->cap, ctx->d_sweep, -1, ctx->d_err);
  HSMC_HIP(ctx, hipGetLastError());
  ctx->swp_last = 0;
  ctx->cells_version = ctx->version;
  return HSMC_OK;
}

// per cell: the slot map of up to lcap stencil spheres and the id and id
// order of the cell's own occupants
size_t phase_lds_bytes(int cap, int lcap, int cpb) {
  // per cell: slot map and occupant ids and order (ints); occupant positions
  // and attempt displacements (doubles)
  return (size_t)cpb * (((size_t)lcap + 2 * (size_t)cap) * sizeof(int) + 6 * (size_t)cap * sizeof(double));
}
}  // namespace hsmc

extern "C" {

int hsmc_gpu_sweeps(hsmc_gpu_ctx *ctx, int n_sweeps) {
  if (!ctx || n_sweeps < 0) return HSMC_ERR_BAD_ARG;
  return run_sweeps(ctx, n_sweeps);
}

int hsmc_gpu_time_sweeps(hsmc_gpu_ctx *ctx, int n_sweeps, float *ms) {
  if (!ctx || n_sweeps < 0 || !ms) return HSMC_ERR_BAD_ARG;
  // capture outside the timed region: the batch graph, and a graph of
  // exactly the remaining n mod sweeps_per_graph sweeps (the whole call when
  // it is shorter than a batch), so every timed sweep is a graph replay
  // (persistent launches capture nothing)
  const int K = ctx->tune.persistent ? 0 : ctx->tune.sweeps_per_graph;
  if (K > 0 && n_sweeps >= K &&
      (!ctx->gexec || ctx->graph_sweeps != K || !same_launch_shape(ctx->graph_tune, ctx->tune))) {
    int st = capture_graph(ctx, K);
    if (st) return st;
  }
  const int rem = K > 0 ? n_sweeps % K : 0;
  if (rem > 0 && (!ctx->gexec_r || ctx->graph_r_sweeps != rem ||
                  !same_launch_shape(ctx->graph_r_tune, ctx->tune))) {
    int st = capture_graph(ctx, rem, true);
    if (st) return st;
  }
  HSMC_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  int st = run_sweeps(ctx, n_sweeps);
  if (st) return st;
  HSMC_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HSMC_HIP(ctx, hipEventSynchronize(ctx->ev1));
  HSMC_HIP(ctx, hipEventElapsedTime(ms, ctx->ev0, ctx->ev1));
  return check_device_errors(ctx);
}

int hsmc_gpu_time_phase_kernel(hsmc_gpu_ctx *ctx, int n_sweeps, float *avg_ms, int32_t *launches) {
  if (!ctx || n_sweeps <= 0 || !avg_ms) return HSMC_ERR_BAD_ARG;
  // events bracket the 8 back-to-back phase launches of each sweep (the
  // rebin kernel runs outside the bracket), so per-event overhead is paid
  // once per 8 launches
  hipEvent_t ev[16];
  for (int i = 0; i < 16; ++i) HSMC_HIP(ctx, hipEventCreate(&ev[i]));
  double total = 0.0;
  int count = 0;
  for (int s = 0; s < n_sweeps; ++s) {
    enqueue_sweep(ctx, 0, ctx->stream, ev);
    hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, ctx->stream, ctx->d_sweep, 1);
    HSMC_HIP(ctx, hipEventSynchronize(ev[15]));
    float ms = 0.f;
    HSMC_HIP(ctx, hipEventElapsedTime(&ms, ev[0], ev[15]));
    total += ms;
    count += 8;
  }
  for (int i = 0; i < 16; ++i) (void)hipEventDestroy(ev[i]);
  ctx->version++;
  ctx->swp_last = 0;
  ctx->cells_version = ctx->version;
  *avg_ms = (float)(total / count);
  if (launches) *launches = count;
  return check_device_errors(ctx);
}

int hsmc_gpu_occupant_picks(int device, uint64_t seed, uint32_t cell, uint64_t sweep, uint32_t n,
                            int32_t count, uint32_t *out) {
  if (n == 0 || count < 0 || count > 65536 || (count > 0 && !out)) return HSMC_ERR_BAD_ARG;
  if (count == 0) return HSMC_OK;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return HSMC_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return HSMC_ERR_BAD_ARG;
  if (hipSetDevice(device) != hipSuccess) return HSMC_ERR_HIP;
  uint32_t *d = nullptr;
  if (hipMalloc(&d, sizeof(uint32_t) * (size_t)count) != hipSuccess) return HSMC_ERR_NO_MEMORY;
  hipLaunchKernelGGL(k_occupant_picks, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, 0,
                     (uint32_t)seed, (uint32_t)(seed >> 32), cell, (unsigned long long)sweep, n,
                     (int)count, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpy(out, d, sizeof(uint32_t) * (size_t)count, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? HSMC_OK : HSMC_ERR_HIP;
}

int hsmc_gpu_counters(hsmc_gpu_ctx *ctx, int replica, uint64_t out[4]) {
  if (!ctx || replica < 0 || replica >= ctx->R || !out) return HSMC_ERR_BAD_ARG;
  int st = fold_counters(ctx);
  if (st) return st;
  unsigned long long sw = 0;
  HSMC_HIP(ctx, hipMemcpyAsync(&sw, ctx->d_sweep, sizeof(sw), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // every occupant of every cell makes one attempt per sweep: N per sweep
  out[0] = (uint64_t)ctx->reps[replica].p.n_part * (sw - ctx->sweep_at_reset);
  out[1] = ctx->tot_acc[replica];
  out[2] = ctx->tot_left[replica];
  out[3] = sw;
  return check_device_errors(ctx);
}

int hsmc_gpu_reset_counters(hsmc_gpu_ctx *ctx) {
  if (!ctx) return HSMC_ERR_BAD_ARG;
  int st = fold_counters(ctx);
  if (st) return st;
  std::fill(ctx->tot_acc.begin(), ctx->tot_acc.end(), 0ull);
  std::fill(ctx->tot_left.begin(), ctx->tot_left.end(), 0ull);
  HSMC_HIP(ctx, hipMemcpyAsync(&ctx->sweep_at_reset, ctx->d_sweep, sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HSMC_OK;
}

int hsmc_gpu_set_sweep_counter(hsmc_gpu_ctx *ctx, uint64_t sweep) {
  if (!ctx) return HSMC_ERR_BAD_ARG;
  unsigned long long now = 0;
  HSMC_HIP(ctx, hipMemcpyAsync(&now, ctx->d_sweep, sizeof(now), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const unsigned long long want = (unsigned long long)sweep;
  HSMC_HIP(ctx, hipMemcpyAsync(ctx->d_sweep, &want, sizeof(want), hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  // attempts are N * (sweep - sweep_at_reset): move the origin with the counter
  ctx->sweep_at_reset += want - now;
  // the last sweep's cell list is found by the counter's parity: refile it
  // before anything reads it
  ctx->cells_version = 0;
  return HSMC_OK;
}

int hsmc_gpu_set_tuning(hsmc_gpu_ctx *ctx, const hsmc_gpu_tuning *t) {
  if (!ctx || !t || !tuning_ok(ctx, *t)) return HSMC_ERR_BAD_ARG;
  ctx->tune = *t;
  drop_graph(ctx);
  return HSMC_OK;
}

int hsmc_gpu_get_tuning(const hsmc_gpu_ctx *ctx, hsmc_gpu_tuning *t) {
  if (!ctx || !t) return HSMC_ERR_BAD_ARG;
  *t = ctx->tune;
  return HSMC_OK;
}

int hsmc_gpu_autotune(hsmc_gpu_ctx *ctx, int sweeps_per_trial, hsmc_gpu_tuning *best) {
  if (!ctx || sweeps_per_trial <= 0) return HSMC_ERR_BAD_ARG;
  const size_t P = (size_t)ctx->Ptot;
  // save positions, sweep counter and move counters; restore afterwards
  double *save = (double *)scratch(ctx, 23, sizeof(double) * 3 * P);
  if (!save) return HSMC_ERR_NO_MEMORY;
  unsigned long long sweep0 = 0;
  {
    int st0 = fold_counters(ctx);
    if (st0) return st0;
  }
  const std::vector<unsigned long long> acc0 = ctx->tot_acc, left0 = ctx->tot_left;
  HSMC_HIP(ctx, hipMemcpyAsync(save, ctx->d_x, sizeof(double) * P, hipMemcpyDeviceToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(save + P, ctx->d_y, sizeof(double) * P, hipMemcpyDeviceToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(save + 2 * P, ctx->d_z, sizeof(double) * P, hipMemcpyDeviceToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(&sweep0, ctx->d_sweep, sizeof(sweep0), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const hsmc_gpu_tuning start = ctx->tune;
  hsmc_gpu_tuning winner = start;
  float best_ms = 1e30f;
  const int lpcs[2] = {32, 64};
  const int cpbs[4] = {1, 2, 4, 8};
  const int graphs[3] = {0, 16, 64};
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 4; ++b)
      for (int g = 0; g < 3; ++g) {
        hsmc_gpu_tuning t{lpcs[a], cpbs[b], graphs[g], 0};
        if (!tuning_ok(ctx, t)) continue;
        if (t.sweeps_per_graph > 0 && t.sweeps_per_graph > sweeps_per_trial) continue;
        ctx->tune = t;
        drop_graph(ctx);
        float ms = 0.f;
        int st = hsmc_gpu_time_sweeps(ctx, std::max(2, sweeps_per_trial / 4), &ms);  // warm-up, capture
        if (st == HSMC_OK) st = hsmc_gpu_time_sweeps(ctx, sweeps_per_trial, &ms);
        if (st != HSMC_OK) {
          ctx->tune = start;
          return st;
        }
        if (ms < best_ms) {
          best_ms = ms;
          winner = t;
        }
      }
  ctx->tune = winner;
  drop_graph(ctx);
  HSMC_HIP(ctx, hipMemcpyAsync(ctx->d_x, save, sizeof(double) * P, hipMemcpyDeviceToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(ctx->d_y, save + P, sizeof(double) * P, hipMemcpyDeviceToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(ctx->d_z, save + 2 * P, sizeof(double) * P, hipMemcpyDeviceToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(ctx->d_sweep, &sweep0, sizeof(sweep0), hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(ctx->d_cacc, 0, sizeof(unsigned int) * ctx->cell_alloc, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(ctx->d_cleft, 0, sizeof(unsigned int) * ctx->cell_alloc, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(ctx->d_cnt, 0, sizeof(int) * 2 * ctx->cell_alloc, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->version++;  // positions restored, cell lists cleared
  ctx->tot_acc = acc0;
  ctx->tot_left = left0;
  if (best) *best = winner;
  return HSMC_OK;
}

}  // extern "C"
