// npt.hip -- isobaric (NpT) sweeps of many replicas at once.
//
// A sweep of the reference's NpT chain (sweep_npt, npt.c:177-196) makes N
// steps, each a volume move with probability 1/(N+1) (npt.c:185).  Here a
// sweep is one checkerboard particle sweep of every replica followed by each
// replica's Binomial(N, 1/(N+1)) volume moves (vol_move, moves.c:83-153):
//   ln V' = ln V + (r - 0.5) dv_max,  sf = (V'/V)^(1/3);
//   rejected if any pair overlaps after scaling every coordinate and the box
//   by sf (moves.c:104-112); else accepted when
//   r' < exp(-beta P (V' - V) + (N + 1) ln(V'/V)) (moves.c:117-121), and then
//   positions and box are scaled by sf (moves.c:129-142).
// Both kernels leave the NpT distribution invariant, so their composition
// samples it.  Volume moves are done in rounds: round t holds the t-th move of
// every replica that draws more than t, and one kernel checks every
// compression of the round at once.
//
// The scan reads the cell list the last sweep left (the sweep grid: cells of
// side > 1 holding every sphere, filed by the sweep's rebin and kept by its
// moves, none of which leaves its cell).  A pair closer than 1/sf before
// scaling lies in neighbouring cells whenever the cell side is at least 1/sf,
// which holds for the small volume steps NpT runs use; a replica whose cells
// are narrower falls back to the single-replica scan on a grid of side 1/sf.
//
// Each replica's volume draws come from a host Philox4x32-10 stream keyed by
// its seed (counter: sweep, round, 0, "NPT\0"), so its run does not depend
// on the other replicas of the batch.
#include <math.h>

#include <algorithm>
#include <vector>

#include "ctx.h"
#include "device_common.h"

using namespace hsmc;

namespace {

constexpr int S_SF = 40, S_BOX = 41, S_FLAG = 42, S_BLK = 43;
constexpr uint32_t kNptTag = 0x4E505400u;  // "NPT\0"

// Any pair of a replica with a proposed compression (sf[r] > 0) closer than 1
// after scaling: thread per sphere, neighbours from the last sweep's cells.
// Blocks cover only the replicas that propose a compression: block b takes
// the 256 spheres from blk[b] on, all of one replica, so the replica (and
// its record) is uniform and the block stops as a whole once the replica
// has an overlap.
// Each pair is tested once: from the sphere's own cell towards the 13
// lexicographically forward neighbour cells, and within the cell towards
// higher ids (the scaled distance is the same from either end, as
// fl(a - b) = -fl(b - a)).  A thread stops once its replica has an overlap.
__global__ void __launch_bounds__(256)
k_vol_overlap(const RepDev *__restrict__ reps, const SweepDev *__restrict__ swp, int R,
              const int *__restrict__ prep, const int *__restrict__ blk, int C,
              const int *__restrict__ cnt, const SlotRec *__restrict__ slot_all, int cap,
              const unsigned long long *__restrict__ sweep_next, const double *__restrict__ x,
              const double *__restrict__ y, const double *__restrict__ z,
              const double *__restrict__ sfr, int *__restrict__ flags) {
  const int i0 = blk[blockIdx.x];
  const int r = prep[i0];
  const RepDev rp = reps[r];
  const int i = i0 + (int)threadIdx.x;
  if (i >= rp.poff + rp.N) return;
  const double sf = sfr[r];
  if (__hip_atomic_load(&flags[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  const unsigned long long sweep = *sweep_next - 1ull;
  const int *cntb = cnt + (size_t)(sweep & 1ull) * C;
  // one replica keeps slot records and table entries per sweep parity
  const int sb = R == 1 ? (int)(sweep & 1ull) : 0;
  const SlotRec *slot = slot_all + (size_t)sb * cap * C;
  // the last sweep's grid shift, as its rebin drew it (swp points at the
  // last sweep's entry when there is one replica)
  const double shx = swp[r].shx, shy = swp[r].shy, shz = swp[r].shz;
  const double ax = x[i], ay = y[i], az = z[i];
  const int cx = shifted_cell(ax, shx, rp.lx, rp.ix, rp.gx);
  const int cy = shifted_cell(ay, shy, rp.ly, rp.iy, rp.gy);
  const int cz = shifted_cell(az, shz, rp.lz, rp.iz, rp.gz);
  const int me = i - rp.poff;
  // d = 13 is the own cell; d = 14..26 are the forward offsets.  The
  // answer does not depend on the order of the tests, so the 14 cells'
  // counts load together, and each cell's first two slots load with the
  // poll of the replica's flag (an overlap found elsewhere ends the scan).
  int cell[14], n[14];
#pragma unroll
  for (int j = 0; j < 14; ++j) {
    const int d = 13 + j;
    const int dx = d / 9 - 1, dy = (d / 3) % 3 - 1, dz = d % 3 - 1;
    const bool ok = stencil_axis_ok(dx, rp.gx) && stencil_axis_ok(dy, rp.gy) && stencil_axis_ok(dz, rp.gz);
    cell[j] = rp.coff + (ok ? (wrap_index(cx + dx, rp.gx) * rp.gy + wrap_index(cy + dy, rp.gy)) * rp.gz +
                                  wrap_index(cz + dz, rp.gz)
                            : 0);
    n[j] = ok ? 1 : 0;  // a cell the stencil leaves out (tiny grids) is skipped
  }
  int m[14];
#pragma unroll
  for (int j = 0; j < 14; ++j) m[j] = cntb[cell[j]];  // no branch around the loads
#pragma unroll
  for (int j = 0; j < 14; ++j) n[j] = n[j] ? min(m[j], cap) : 0;
  auto hits = [&](const SlotRec &rec, bool own) {
    // compute_dist with sf (moves.c:400-431), reference particle first;
    // in the own cell each pair once, self excluded
    return !(own && rec.id <= me) &&
           dist2_scaled(ax, ay, az, rec.x, rec.y, rec.z, rp.lx, rp.ly, rp.lz, sf) < 1.0;
  };
  bool ov = false;
  for (int j = 0; j < 14 && !ov; ++j) {
    const int c = cell[j];
    const SlotRec r0 = slot[c], r1 = slot[(size_t)(cap > 1 ? 1 : 0) * C + c];
    const int stop = j > 0 ? __hip_atomic_load(&flags[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
    if (stop) break;
    ov = (n[j] > 0 && hits(r0, j == 0)) || (n[j] > 1 && hits(r1, j == 0));
    for (int q = 2; q < n[j] && !ov; ++q) ov = hits(slot[(size_t)q * C + c], j == 0);
  }
  if (ov) atomicOr(&flags[r], 1);
}

// Scale the spheres of replicas with sf[r] > 0 into their new boxes
// (moves.c:134-140: coordinates times sf, then apply_pbc in the new box).
__global__ void k_scale_batch(const int *__restrict__ prep, int P, double *__restrict__ x,
                              double *__restrict__ y, double *__restrict__ z,
                              const double *__restrict__ sfr, const double *__restrict__ box) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const int r = prep[i];
  const double sf = sfr[r];
  if (!(sf > 0.0)) return;
  x[i] = wrap_once(x[i] * sf, box[3 * r]);
  y[i] = wrap_once(y[i] * sf, box[3 * r + 1]);
  z[i] = wrap_once(z[i] * sf, box[3 * r + 2]);
}

__global__ void k_zero_flags(int *__restrict__ f, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = 0;
}

// Uniform in [0, 1] from one Philox output word (rng_get_double's range).
double unit_closed(uint32_t u) { return (double)u / 4294967295.0; }

// Number of volume moves among the N steps of one sweep: Binomial(N, p),
// p = 1/(N+1), by inversion of u in [0, 1).
int volume_steps(int N, uint32_t word) {
  const double u = (double)word * 0x1p-32;
  const double p = 1.0 / (N + 1.0);
  double pk = pow(1.0 - p, N), cdf = pk;
  int k = 0;
  while (u > cdf && k < N) {
    pk *= (double)(N - k) / (double)(k + 1) * p / (1.0 - p);
    ++k;
    cdf += pk;
  }
  return k;
}

void draws(const RepHost &h, unsigned long long sweep, uint32_t round, uint32_t out[4]) {
  philox4x32_10((uint32_t)sweep, (uint32_t)(sweep >> 32), round, kNptTag, (uint32_t)h.p.seed,
                (uint32_t)(h.p.seed >> 32), out);
}

}  // namespace

extern "C" {

int hsmc_gpu_box(const hsmc_gpu_ctx *ctx, int replica, double out[3]) {
  if (!ctx || replica < 0 || replica >= ctx->R || !out) return HSMC_ERR_BAD_ARG;
  out[0] = ctx->reps[replica].p.lx;
  out[1] = ctx->reps[replica].p.ly;
  out[2] = ctx->reps[replica].p.lz;
  return HSMC_OK;
}

int hsmc_gpu_any_overlap_scaled_batch(hsmc_gpu_ctx *ctx, const double *sf, uint8_t *no_overlap) {
  if (!ctx || !sf || !no_overlap) return HSMC_ERR_BAD_ARG;
  const int R = ctx->R;
  double *d_sf = (double *)scratch(ctx, S_SF, sizeof(double) * R);
  int *d_flag = (int *)scratch(ctx, S_FLAG, sizeof(int) * R);
  if (!d_sf || !d_flag) return HSMC_ERR_NO_MEMORY;
  // replicas whose sweep-grid cells cover the scaled cutoff 1/sf are scanned
  // together on the sweep grid; the others on their own grid of side 1/sf
  std::vector<double> sf_dev(R, 0.0);
  bool any = false;
  for (int r = 0; r < R; ++r) {
    no_overlap[r] = 1;
    if (!(sf[r] > 0.0)) continue;
    const RepHost &h = ctx->reps[r];
    const double side = std::min(h.p.lx / h.gx, std::min(h.p.ly / h.gy, h.p.lz / h.gz));
    if (side >= (1.0 / sf[r]) * (1.0 + 1e-9)) {
      sf_dev[r] = sf[r];
      any = true;
    } else {
      int st = hsmc_gpu_any_overlap_scaled(ctx, r, &sf[r], 1, &no_overlap[r]);
      if (st) return st;
    }
  }
  if (!any) return HSMC_OK;
  if (ctx->cells_version != ctx->version) {  // positions changed since the last sweep
    int st = rebin_last_sweep(ctx);
    if (st) return st;
  }
  // first sphere of each 256-sphere block of the replicas scanned here
  std::vector<int> blk;
  for (int r = 0; r < R; ++r) {
    if (!(sf_dev[r] > 0.0)) continue;
    const RepHost &h = ctx->reps[r];
    for (long b = h.poff; b < h.poff + h.p.n_part; b += 256) blk.push_back((int)b);
  }
  int *d_blk = (int *)scratch(ctx, S_BLK, sizeof(int) * blk.size());
  if (!d_blk) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemcpyAsync(d_sf, sf_dev.data(), sizeof(double) * R, hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(d_blk, blk.data(), sizeof(int) * blk.size(), hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_zero_flags, dim3((R + 255) / 256), dim3(256), 0, ctx->stream, d_flag, R);
  hipLaunchKernelGGL(k_vol_overlap, dim3((unsigned)blk.size()), dim3(256), 0, ctx->stream, ctx->d_reps,
                     ctx->d_swp + (ctx->R == 1 ? ctx->swp_last : 0), ctx->R, ctx->d_prep, d_blk, (int)ctx->Ctot, ctx->d_cnt, ctx->d_slot,
                     ctx->cap, ctx->d_sweep, ctx->d_x, ctx->d_y, ctx->d_z, d_sf, d_flag);
  HSMC_HIP(ctx, hipGetLastError());
  std::vector<int> hf(R);
  HSMC_HIP(ctx, hipMemcpyAsync(hf.data(), d_flag, sizeof(int) * R, hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int r = 0; r < R; ++r)
    if (sf_dev[r] > 0.0) no_overlap[r] = hf[r] ? 0 : 1;
  return check_device_errors(ctx);
}

int hsmc_gpu_npt_sweeps(hsmc_gpu_ctx *ctx, int n_sweeps, const double *press,
                        const double *dv_max, uint64_t *vol_counts) {
  if (!ctx || n_sweeps < 0 || !press || !dv_max) return HSMC_ERR_BAD_ARG;
  const int R = ctx->R;
  for (int r = 0; r < R; ++r)
    if (!(press[r] >= 0.0) || !(dv_max[r] > 0.0) || !std::isfinite(press[r])) return HSMC_ERR_BAD_ARG;
  if (n_sweeps == 0) return HSMC_OK;
  double *d_sf = (double *)scratch(ctx, S_SF, sizeof(double) * R);
  double *d_box = (double *)scratch(ctx, S_BOX, sizeof(double) * 3 * R);
  int *d_flag = (int *)scratch(ctx, S_FLAG, sizeof(int) * R);
  if (!d_sf || !d_box || !d_flag) return HSMC_ERR_NO_MEMORY;
  unsigned long long sweep0 = 0;
  HSMC_HIP(ctx, hipMemcpyAsync(&sweep0, ctx->d_sweep, sizeof(sweep0), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));

  std::vector<int> steps(R), flag(R);
  std::vector<uint8_t> no_ov(R);
  std::vector<double> sf(R), vnew(R), sf_dev(R), box(3 * R);
  std::vector<uint32_t> acc_word(R);
  const unsigned blocks = (unsigned)((ctx->Ptot + 255) / 256);
  for (int sw = 0; sw < n_sweeps; ++sw) {
    int st = run_sweeps(ctx, 1);
    if (st) return st;
    const unsigned long long s = sweep0 + (unsigned long long)sw;  // the sweep just done
    int rounds = 0;
    for (int r = 0; r < R; ++r) {
      uint32_t u[4];
      draws(ctx->reps[r], s, 0u, u);
      steps[r] = volume_steps(ctx->reps[r].p.n_part, u[0]);
      rounds = std::max(rounds, steps[r]);
    }
    for (int t = 0; t < rounds; ++t) {
      // proposals (moves.c:91-102)
      bool any_compression = false;
      for (int r = 0; r < R; ++r) {
        sf[r] = 0.0;
        if (steps[r] <= t) continue;
        const hsmc_gpu_replica &p = ctx->reps[r].p;
        uint32_t u[4];
        draws(ctx->reps[r], s, (uint32_t)t + 1u, u);
        const double vol = p.lx * p.ly * p.lz;
        const double log_vol_new = log(vol) + (unit_closed(u[0]) - 0.5) * dv_max[r];
        vnew[r] = exp(log_vol_new);
        sf[r] = pow(vnew[r] / vol, 1. / 3.);
        acc_word[r] = u[1];
        if (sf[r] < 1.0) any_compression = true;
      }
      // overlap scan of the compressions: an expansion cannot create an
      // overlap in a configuration that has none
      std::fill(flag.begin(), flag.end(), 0);
      if (any_compression) {
        for (int r = 0; r < R; ++r) sf_dev[r] = (sf[r] > 0.0 && sf[r] < 1.0) ? sf[r] : 0.0;
        st = hsmc_gpu_any_overlap_scaled_batch(ctx, sf_dev.data(), no_ov.data());
        if (st) return st;
        for (int r = 0; r < R; ++r)
          if (sf_dev[r] > 0.0) flag[r] = no_ov[r] ? 0 : 1;
      }
      // acceptance (moves.c:114-127)
      bool any_accepted = false;
      for (int r = 0; r < R; ++r) {
        sf_dev[r] = 0.0;
        if (!(sf[r] > 0.0)) continue;
        hsmc_gpu_replica &p = ctx->reps[r].p;
        uint64_t *vc = vol_counts ? vol_counts + 4 * (size_t)r : nullptr;
        if (vc) vc[0] += 1;
        if (flag[r]) {
          if (vc) vc[2] += 1;
          continue;
        }
        const double vol = p.lx * p.ly * p.lz;
        const double boltz = exp(press[r] * (vol - vnew[r]) + (p.n_part + 1) * log(vnew[r] / vol));
        if (unit_closed(acc_word[r]) >= boltz) {
          if (vc) vc[3] += 1;
          continue;
        }
        if (vc) vc[1] += 1;
        p.lx *= sf[r];
        p.ly *= sf[r];
        p.lz *= sf[r];
        sf_dev[r] = sf[r];
        box[3 * r] = p.lx;
        box[3 * r + 1] = p.ly;
        box[3 * r + 2] = p.lz;
        any_accepted = true;
      }
      if (any_accepted) {
        HSMC_HIP(ctx, hipMemcpyAsync(d_sf, sf_dev.data(), sizeof(double) * R, hipMemcpyHostToDevice, ctx->stream));
        HSMC_HIP(ctx, hipMemcpyAsync(d_box, box.data(), sizeof(double) * 3 * R, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_scale_batch, dim3(blocks), dim3(256), 0, ctx->stream, ctx->d_prep,
                           (int)ctx->Ptot, ctx->d_x, ctx->d_y, ctx->d_z, d_sf, d_box);
        HSMC_HIP(ctx, hipGetLastError());
        // the host copies of sf/box are reused next round: wait for the copies
        HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
        st = layout(ctx);
        if (st) return st;
      }
    }
  }
  return check_device_errors(ctx);
}

}  // extern "C"
