// ctx.h -- internal layout of an hsmc_gpu_ctx (not part of the C ABI).
//
// Device memory of a context (R replicas, P = sum N_r particles, C = sum of
// sweep-grid cells, A = C/8 active cells per sublattice phase):
//   x, y, z          FP64[P]     positions by (replica offset + particle id)
//   prep             int32[P]    replica of each particle
//   cnt              int32[2*C]  occupancy of each sweep cell (two buffers,
//                                alternating by sweep parity)
//   slot             SlotRec[2*cap*C] position copy and particle id in each
//                                cell slot (32 B, one cache-line request each),
//                                slot-major: slot q of cell c is record q*C + c,
//                                so the first occupants of z-adjacent cells
//                                share cache lines (one replica: two buffers,
//                                by sweep parity, as for the counts)
//   reps             RepDev[R]   box, grid and RNG key of each replica
//   swp              SweepDev[max(R, 520)] the sweep's grid shift and
//                                sublattice order per replica (one replica:
//                                entry k for the k-th sweep of a batch)
//   aoff             int32[R+1]  prefix of active cells per phase, each
//                                replica's range padded to a multiple of 8
//   arep             int32[A]    replica owning each active cell
//   sweep            uint64      sweeps done (read by every sweep kernel)
//   cacc, cleft      uint32[C]   accepted moves / moves rejected for leaving
//                                the cell, per sweep cell; each cell is
//                                updated by the one work item that owns it in
//                                its phase, so plain stores suffice
//   err              uint32      sticky error bits set by kernels
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "hsmc_gpu.h"

namespace hsmc {

constexpr unsigned kErrOverflow = 1u;  // a cell received more than cap occupants
constexpr unsigned kErrListFull = 2u;  // an LDS candidate list overflowed
constexpr unsigned kErrStall = 4u;     // a persistent launch's grid barrier timed out

struct RepDev {
  double lx, ly, lz;     // box
  double hx, hy, hz;     // half box (lx/2.0 ...)
  double sx, sy, sz;     // sweep-grid cell side (lx/gx ...)
  double ix, iy, iz;     // 1.0 / sx ...: the sweep grid's cell index is (int)(t * ix)
  double dr_max;         // displacement width
  int N, gx, gy, gz;     // particles, sweep grid
  int poff, coff, aoff;  // offsets into particle, cell and active-cell arrays
  uint32_t k0, k1;       // Philox key
};

// Entries of the one-replica sweep table: one per sweep of a batch (at most
// 512 sweeps per graph) and one past the last.
constexpr int kSweepTable = 520;

// What one sweep of a replica draws before any move: the grid shift and the
// order of the 8 sublattices (nibble p = sublattice of phase p).  k_rebin
// writes it; the 8 phase launches of the sweep read it.
struct SweepDev {
  double shx, shy, shz;
  uint32_t order;
  uint32_t pad;
};

// One cell slot: a sphere's position and its id in its replica.  32 B and
// 32-B aligned, so one sphere is one cache-line request however many of its
// fields a kernel reads or writes.
struct __attribute__((aligned(32))) SlotRec {
  double x, y, z;
  int id;
  int pad;
};

struct RepHost {
  hsmc_gpu_replica p;
  int gx, gy, gz;
  long poff, coff, aoff;
};

}  // namespace hsmc

struct hsmc_gpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int R = 0;
  std::vector<hsmc::RepHost> reps;
  long Ptot = 0, Ctot = 0, Atot = 0;
  int cap = 8;
  size_t cell_alloc = 0;  // cells allocated (>= Ctot)
  hsmc::RepDev *d_reps = nullptr;
  hsmc::RepDev rep0{};  // host copy of replica 0's record (a kernel argument when R == 1)
  // shift and order: entry r of replica r; one replica: entry k of the k-th
  // sweep of the last batch (the last sweep's entry is swp_last)
  hsmc::SweepDev *d_swp = nullptr;
  int swp_last = 0;
  int *d_aoff = nullptr;
  int *d_arep = nullptr;  // replica of each active cell (Atot entries)
  size_t arep_alloc = 0;
  int lcap = 128;         // LDS candidate-list capacity of the phase kernel
  int *d_prep = nullptr;
  double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
  int *d_cnt = nullptr;
  hsmc::SlotRec *d_slot = nullptr;
  unsigned long long *d_sweep = nullptr;
  unsigned int *d_cacc = nullptr, *d_cleft = nullptr;
  // host totals per replica: accepted, left-cell (folded from the per-cell
  // device counters), and the sweep counter value at the last reset
  std::vector<unsigned long long> tot_acc, tot_left;
  unsigned long long sweep_at_reset = 0;
  unsigned int *d_err = nullptr;
  // arrival counter and generation of the persistent sweep's grid barrier
  unsigned int *d_bar = nullptr;
  int n_cu = 0;  // compute units of the device (co-resident workgroup bound)
  // pinned host copy of the error bits, refreshed by an asynchronous copy
  // after each batch of sweeps, so later batches are not launched onto a
  // cell list that has already overflowed
  volatile unsigned int *h_err = nullptr;
  hsmc_gpu_tuning tune{32, 8, 32, 0};
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int graph_sweeps = 0;
  hsmc_gpu_tuning graph_tune{0, 0, 0, 0};
  // a second captured graph for a call's remainder (n mod sweeps_per_graph
  // sweeps), captured by the timing call outside its timed region
  hipGraph_t graph_r = nullptr;
  hipGraphExec_t gexec_r = nullptr;
  int graph_r_sweeps = 0;
  hsmc_gpu_tuning graph_r_tune{0, 0, 0, 0};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string last_error;
  // configuration version (bumped by anything that moves particles) and the
  // cached observable grid built for (replica, cutoff, version)
  unsigned long long version = 1;
  // configuration version the sweep grid's cell list (the last sweep's count
  // buffer and slots) describes; 0 when it describes none
  unsigned long long cells_version = 0;
  int aux_replica = -1;
  double aux_cut = 0.0;
  unsigned long long aux_version = 0;
  int aux_n[3] = {0, 0, 0};
  // owner stamp of the learned proposal's voxel-score and cumulative-mass
  // scratch buffers: every call that writes them takes a new value, and a
  // network's cached scores are valid only while the stamp is the one it took
  unsigned long long score_stamp = 0;
  // neighbour-grid cache of the learned proposal: entry 0 the descriptor
  // grid alone, entry 1 the descriptor and overlap grids built together (its
  // cut is the sum of the two cuts)
  int ngrid_replica[2] = {-1, -1};
  unsigned long long ngrid_version[2] = {0, 0};
  double ngrid_cut[2] = {0.0, 0.0};
  // scratch buffers of the observables, grown on demand
  static constexpr int kScratch = 44;
  void *scratch[kScratch] = {};
  size_t scratch_bytes[kScratch] = {};
};

namespace hsmc {

// Record a HIP failure and return HSMC_ERR_HIP.
int hip_fail(hsmc_gpu_ctx *ctx, hipError_t e, const char *what);
// Device buffer `slot` of at least `bytes` bytes.
void *scratch(hsmc_gpu_ctx *ctx, int slot, size_t bytes);
// (Re)build replica offsets, the sweep grids, and the device cell arrays.
int layout(hsmc_gpu_ctx *ctx);
// LDS bytes of one sublattice-phase workgroup of cpb cells
size_t phase_lds_bytes(int cap, int lcap, int cpb);
// n checkerboard sweeps of every replica (enqueued on the context stream)
int run_sweeps(hsmc_gpu_ctx *ctx, int n);
// File the current by-id positions into the cell slots of the last sweep's
// grid (the cell list the volume-move scan reads after positions changed).
int rebin_last_sweep(hsmc_gpu_ctx *ctx);
// Sweep-grid cells per axis for a box edge: largest even g with L/g > 1.
int sweep_grid_cells(double L);
// Read and clear the kernels' sticky error bits; translate to a status.
int check_device_errors(hsmc_gpu_ctx *ctx);
// Invalidate any captured sweep graph.
void drop_graph(hsmc_gpu_ctx *ctx);
// Add the per-cell device counters into the host totals and clear them.
int fold_counters(hsmc_gpu_ctx *ctx);
// Workspace bytes the scaled-overlap scan of a replica needs (0 for bad
// arguments), and the scan itself with every buffer carved from a device
// workspace of ws_bytes (nothing allocated).
size_t overlap_scan_workspace_bytes(const hsmc_gpu_ctx *ctx, int replica, const double *sf, int n_sf);
int overlap_scan_in_workspace(hsmc_gpu_ctx *ctx, int replica, const double *sf, int n_sf,
                              unsigned char *ws, size_t ws_bytes, uint8_t *no_overlap);
// Look up the allocation-free scratch array's address on `device` (current).
int noalloc_scratch_init(int device);
// Exclusive prefix sum of in[0..n) into out[0..n] (out[n] = total), three
// launches on stream s; part is a device workspace of scan_workspace_ints(n)
// ints.
long scan_workspace_ints(long n);
void exclusive_scan(hipStream_t s, const int *in, int n, int *out, int *part);

}  // namespace hsmc

#define HSMC_HIP(ctx, call)                                   \
  do {                                                        \
    hipError_t e_ = (call);                                   \
    if (e_ != hipSuccess) return hsmc::hip_fail((ctx), e_, #call); \
  } while (0)
