// ctx.hip -- life cycle, layout and configuration transfer of a context.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ctx.h"
#include "device_common.h"

using namespace hsmc;

namespace hsmc {

int hip_fail(hsmc_gpu_ctx *ctx, hipError_t e, const char *what) {
  if (ctx) {
    ctx->last_error = std::string(what) + ": " + hipGetErrorString(e);
  }
  return HSMC_ERR_HIP;
}

void *scratch(hsmc_gpu_ctx *ctx, int slot, size_t bytes) {
  if (slot < 0 || slot >= hsmc_gpu_ctx::kScratch) return nullptr;
  if (bytes == 0) bytes = 16;
  if (ctx->scratch_bytes[slot] >= bytes) return ctx->scratch[slot];
  if (ctx->scratch[slot]) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->scratch[slot]);
    ctx->scratch[slot] = nullptr;
    ctx->scratch_bytes[slot] = 0;
  }
  size_t want = bytes + bytes / 4;
  if (hipMalloc(&ctx->scratch[slot], want) != hipSuccess) {
    ctx->last_error = "scratch allocation failed";
    return nullptr;
  }
  ctx->scratch_bytes[slot] = want;
  return ctx->scratch[slot];
}

int sweep_grid_cells(double L) {
  const double minside = 1.0 + 1e-9;
  int g = (int)floor(L / minside);
  if (g & 1) g -= 1;
  while (g >= 2 && !(L / g > minside)) g -= 2;
  return g;
}

void drop_graph(hsmc_gpu_ctx *ctx) {
  if (ctx->gexec) (void)hipGraphExecDestroy(ctx->gexec);
  if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
  ctx->gexec = nullptr;
  ctx->graph = nullptr;
  ctx->graph_sweeps = 0;
  if (ctx->gexec_r) (void)hipGraphExecDestroy(ctx->gexec_r);
  if (ctx->graph_r) (void)hipGraphDestroy(ctx->graph_r);
  ctx->gexec_r = nullptr;
  ctx->graph_r = nullptr;
  ctx->graph_r_sweeps = 0;
}

int fold_counters(hsmc_gpu_ctx *ctx) {
  // the counters are indexed by active-cell index (Atot entries, replica r
  // owning [aoff, aoff + cells/8) and padding up to a multiple of 8 that no
  // kernel writes)
  if (!ctx->d_cacc || ctx->Atot == 0) return HSMC_OK;
  std::vector<unsigned int> a(ctx->Atot), l(ctx->Atot);
  HSMC_HIP(ctx, hipMemcpyAsync(a.data(), ctx->d_cacc, sizeof(unsigned int) * ctx->Atot,
                               hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(l.data(), ctx->d_cleft, sizeof(unsigned int) * ctx->Atot,
                               hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(ctx->d_cacc, 0, sizeof(unsigned int) * ctx->Atot, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(ctx->d_cleft, 0, sizeof(unsigned int) * ctx->Atot, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int r = 0; r < ctx->R; ++r) {
    const RepHost &h = ctx->reps[r];
    const long n = (long)h.gx * h.gy * h.gz / 8;
    unsigned long long sa = 0, sl = 0;
    for (long c = h.aoff; c < h.aoff + n; ++c) {
      sa += a[c];
      sl += l[c];
    }
    ctx->tot_acc[r] += sa;
    ctx->tot_left[r] += sl;
  }
  return HSMC_OK;
}

// Capacity of the phase kernel's stencil list: spheres of unit diameter whose
// centres lie in the 27-cell stencil fit in a box one diameter wider; their
// count is bounded by the close-packing fraction pi/(3 sqrt 2) = 0.7405
// (taken as 0.78 for margin) times that volume over the sphere volume, and
// by N.
long stencil_list_bound(const RepDev &d) {
  const double span_x = std::min(3, d.gx) * d.sx + 1.0;
  const double span_y = std::min(3, d.gy) * d.sy + 1.0;
  const double span_z = std::min(3, d.gz) * d.sz + 1.0;
  long b = (long)ceil(0.78 * span_x * span_y * span_z / (M_PI / 6.0));
  return b > d.N ? d.N : b;
}

int lcap_for(long bound) { return bound <= 128 ? 128 : (bound <= 192 ? 192 : 256); }

// Device parameters of one replica from its box, grid and offsets.
void fill_repdev(const RepHost &h, RepDev &d) {
  const hsmc_gpu_replica &p = h.p;
  d.lx = p.lx; d.ly = p.ly; d.lz = p.lz;
  d.hx = p.lx / 2.0; d.hy = p.ly / 2.0; d.hz = p.lz / 2.0;
  d.sx = p.lx / h.gx; d.sy = p.ly / h.gy; d.sz = p.lz / h.gz;
  d.ix = 1.0 / d.sx; d.iy = 1.0 / d.sy; d.iz = 1.0 / d.sz;
  d.dr_max = p.dr_max;
  d.N = p.n_part;
  d.gx = h.gx; d.gy = h.gy; d.gz = h.gz;
  d.poff = (int)h.poff; d.coff = (int)h.coff; d.aoff = (int)h.aoff;
  d.k0 = (uint32_t)p.seed;
  d.k1 = (uint32_t)(p.seed >> 32);
}

// A cube of side < 2/sqrt(3) holds at most 8 points 1 apart; wider cells get
// 16 slots, which bound the occupancy of cubes narrower than 1.5 (an fcc
// cell of side sqrt(2) holds 14 points 1 apart).  Wider cells only occur on
// 2-cell axes of boxes under 4 diameters and are refused.
bool needs_big_cells(const RepDev &d) { return d.sx >= 1.1547 || d.sy >= 1.1547 || d.sz >= 1.1547; }
bool cells_too_wide(const RepDev &d) { return d.sx >= 1.5 || d.sy >= 1.5 || d.sz >= 1.5; }

// Box or step-size change that keeps every replica's sweep grid, the slot
// capacity and the stencil-list bound: only the parameter table changes.
// The cell arrays and counters stay valid (the next sweep re-bins every
// sphere), and so does a captured graph of a batch (its kernels read the
// table at run time); a one-replica graph carries replica 0's record as a
// kernel argument and is dropped when that record changes.
bool layout_in_place(hsmc_gpu_ctx *ctx) {
  if (!ctx->d_cnt || !ctx->d_reps) return false;
  std::vector<RepDev> rd(ctx->R);
  bool big = false;
  long tmax = 0;
  for (int r = 0; r < ctx->R; ++r) {
    const RepHost &h = ctx->reps[r];
    const hsmc_gpu_replica &p = h.p;
    if (sweep_grid_cells(p.lx) != h.gx || sweep_grid_cells(p.ly) != h.gy ||
        sweep_grid_cells(p.lz) != h.gz)
      return false;
    RepDev &d = rd[r];
    fill_repdev(h, d);
    if (cells_too_wide(d)) return false;  // refused by the full layout
    big = big || needs_big_cells(d);
    tmax = std::max(tmax, stencil_list_bound(d));
  }
  if ((big ? 16 : 8) != ctx->cap || lcap_for(tmax) != ctx->lcap) return false;
  if (ctx->R == 1 && memcmp(&ctx->rep0, &rd[0], sizeof(RepDev)) != 0) drop_graph(ctx);
  ctx->rep0 = rd[0];
  if (hipMemcpyAsync(ctx->d_reps, rd.data(), sizeof(RepDev) * ctx->R, hipMemcpyHostToDevice,
                     ctx->stream) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess)  // rd is freed on return
    return false;
  ctx->version++;
  return true;
}

int layout(hsmc_gpu_ctx *ctx) {
  if (layout_in_place(ctx)) return HSMC_OK;
  {
    int st = fold_counters(ctx);  // the cell arrays may move
    if (st) return st;
  }
  long coff = 0, aoff = 0;
  bool big = false;
  std::vector<RepDev> rd(ctx->R);
  std::vector<int> ao(ctx->R + 1);
  for (int r = 0; r < ctx->R; ++r) {
    RepHost &h = ctx->reps[r];
    const hsmc_gpu_replica &p = h.p;
    h.gx = sweep_grid_cells(p.lx);
    h.gy = sweep_grid_cells(p.ly);
    h.gz = sweep_grid_cells(p.lz);
    if (h.gx < 2 || h.gy < 2 || h.gz < 2) {
      ctx->last_error = "box edge must exceed 2 diameters for the checkerboard sweep";
      return HSMC_ERR_NOT_SUPPORTED;
    }
    h.coff = coff;
    h.aoff = aoff;
    const long cells = (long)h.gx * h.gy * h.gz;
    coff += cells;
    // a workgroup of the phase kernel takes 1, 2, 4 or 8 consecutive active
    // cells: padding each replica's range to a multiple of 8 keeps every
    // workgroup inside one replica (the padding cells exit at once)
    aoff += (cells / 8 + 7) / 8 * 8;
    RepDev &d = rd[r];
    fill_repdev(h, d);
    if (cells_too_wide(d)) {
      ctx->last_error = "sweep-grid cells of 1.5 diameters or more (a box edge in [3, 4)) are not supported";
      return HSMC_ERR_NOT_SUPPORTED;
    }
    big = big || needs_big_cells(d);
    ao[r] = (int)h.aoff;
  }
  ao[ctx->R] = (int)aoff;
  if (coff > (1L << 30)) return HSMC_ERR_NOT_SUPPORTED;
  // capacity of the phase kernel's stencil list
  {
    long tmax = 0;
    for (int r = 0; r < ctx->R; ++r) tmax = std::max(tmax, stencil_list_bound(rd[r]));
    ctx->lcap = lcap_for(tmax);
  }
  if ((size_t)aoff > ctx->arep_alloc) {
    (void)hipFree(ctx->d_arep);
    ctx->d_arep = nullptr;
    HSMC_HIP(ctx, hipMalloc(&ctx->d_arep, sizeof(int) * (aoff + aoff / 8 + 64)));
    ctx->arep_alloc = aoff + aoff / 8 + 64;
  }
  {
    std::vector<int> arep(aoff);
    for (int r = 0; r < ctx->R; ++r)
      for (long a = ao[r]; a < ao[r + 1]; ++a) arep[a] = r;
    HSMC_HIP(ctx, hipMemcpy(ctx->d_arep, arep.data(), sizeof(int) * aoff, hipMemcpyHostToDevice));
  }
  ctx->Ctot = coff;
  ctx->Atot = aoff;
  const int cap = big ? 16 : 8;
  if (coff * cap >= (1L << 31)) return HSMC_ERR_NOT_SUPPORTED;  // slot indices are int
  const size_t need = (size_t)coff;
  if (need > ctx->cell_alloc || cap != ctx->cap || !ctx->d_cnt) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(ctx->d_cnt); (void)hipFree(ctx->d_slot);
    (void)hipFree(ctx->d_cacc); (void)hipFree(ctx->d_cleft);
    ctx->d_cnt = nullptr; ctx->d_slot = nullptr;
    ctx->d_cacc = ctx->d_cleft = nullptr;
    size_t alloc = need + need / 8 + 64;
    HSMC_HIP(ctx, hipMalloc(&ctx->d_cnt, sizeof(int) * 2 * alloc));
    HSMC_HIP(ctx, hipMalloc(&ctx->d_slot, sizeof(SlotRec) * 2 * alloc * cap));  // per sweep parity
    HSMC_HIP(ctx, hipMalloc(&ctx->d_cacc, sizeof(unsigned int) * alloc));
    HSMC_HIP(ctx, hipMalloc(&ctx->d_cleft, sizeof(unsigned int) * alloc));
    ctx->cell_alloc = alloc;
    ctx->cap = cap;
  }
  // keep the phase kernel's LDS staging within 64 KiB per workgroup
  {
    while (ctx->tune.cells_per_block > 1 &&
           phase_lds_bytes(ctx->cap, ctx->lcap, ctx->tune.cells_per_block) > 64 * 1024)
      ctx->tune.cells_per_block >>= 1;
  }
  HSMC_HIP(ctx, hipMemsetAsync(ctx->d_cnt, 0, sizeof(int) * 2 * ctx->cell_alloc, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(ctx->d_cacc, 0, sizeof(unsigned int) * ctx->cell_alloc, ctx->stream));
  HSMC_HIP(ctx, hipMemsetAsync(ctx->d_cleft, 0, sizeof(unsigned int) * ctx->cell_alloc, ctx->stream));
  ctx->rep0 = rd[0];
  HSMC_HIP(ctx, hipMemcpyAsync(ctx->d_reps, rd.data(), sizeof(RepDev) * ctx->R,
                               hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(ctx->d_aoff, ao.data(), sizeof(int) * (ctx->R + 1),
                               hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  drop_graph(ctx);
  ctx->version++;
  return HSMC_OK;
}

int check_device_errors(hsmc_gpu_ctx *ctx) {
  unsigned int e = 0;
  HSMC_HIP(ctx, hipMemcpyAsync(&e, ctx->d_err, sizeof(e), hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->h_err) *ctx->h_err = 0u;
  if (e) {
    unsigned int zero = 0;
    HSMC_HIP(ctx, hipMemcpy(ctx->d_err, &zero, sizeof(zero), hipMemcpyHostToDevice));
    if (e & kErrStall) {
      ctx->last_error = "a persistent sweep launch stopped at a grid barrier that did not fill";
      return HSMC_ERR_HIP;
    }
    ctx->last_error = (e & kErrOverflow) ? "a cell received more occupants than it holds"
                                         : "an overlap candidate list overflowed";
    return HSMC_ERR_CELL_OVERFLOW;
  }
  return HSMC_OK;
}

}  // namespace hsmc

namespace {

__global__ void k_scale(double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
                        int n, double sf, double lx, double ly, double lz) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // moves.c:135-140: part_conf *= sf, then apply_pbc with the new box
  x[i] = wrap_once(x[i] * sf, lx);
  y[i] = wrap_once(y[i] * sf, ly);
  z[i] = wrap_once(z[i] * sf, lz);
}

bool valid_box(double L) { return L > 2.0 && std::isfinite(L); }

__global__ void k_set_particles(double *__restrict__ x, double *__restrict__ y,
                                double *__restrict__ z, const int *__restrict__ idx,
                                const double *__restrict__ q, int m) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int i = idx[k];
  x[i] = q[3 * k];
  y[i] = q[3 * k + 1];
  z[i] = q[3 * k + 2];
}

}  // namespace

extern "C" {

const char *hsmc_gpu_strerror(int code) {
  switch (code) {
    case HSMC_OK: return "ok";
    case HSMC_ERR_CELL_OVERFLOW: return "more particles in one cell than it holds";
    case HSMC_ERR_CELL_TOO_SMALL: return "size of cells in cell list is smaller than the size of the particles";
    case HSMC_ERR_INIT_OVERLAP: return "overlap in the initial configuration";
    case HSMC_ERR_HIP: return "HIP runtime error";
    case HSMC_ERR_BAD_ARG: return "argument out of range";
    case HSMC_ERR_NO_MEMORY: return "allocation failed";
    case HSMC_ERR_NOT_SUPPORTED: return "not supported";
    case HSMC_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

const char *hsmc_gpu_last_error(const hsmc_gpu_ctx *ctx) {
  return ctx ? ctx->last_error.c_str() : "";
}

int hsmc_gpu_create(hsmc_gpu_ctx **out, int device, int n_replicas,
                    const hsmc_gpu_replica *replicas) {
  if (!out || n_replicas <= 0 || !replicas) return HSMC_ERR_BAD_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return HSMC_ERR_NO_DEVICE;
  if (device < 0 || device >= ndev) return HSMC_ERR_BAD_ARG;
  long ptot = 0;
  for (int r = 0; r < n_replicas; ++r) {
    const hsmc_gpu_replica &p = replicas[r];
    if (p.n_part <= 0 || !valid_box(p.lx) || !valid_box(p.ly) || !valid_box(p.lz) ||
        !(p.dr_max >= 0.0))
      return HSMC_ERR_BAD_ARG;
    ptot += p.n_part;
  }
  if (ptot > (1L << 30)) return HSMC_ERR_NOT_SUPPORTED;
  hsmc_gpu_ctx *ctx = new hsmc_gpu_ctx();
  ctx->device = device;
  ctx->R = n_replicas;
  ctx->reps.resize(n_replicas);
  std::vector<int> prep(ptot);
  long poff = 0;
  for (int r = 0; r < n_replicas; ++r) {
    ctx->reps[r].p = replicas[r];
    ctx->reps[r].poff = poff;
    for (int i = 0; i < replicas[r].n_part; ++i) prep[poff + i] = r;
    poff += replicas[r].n_part;
  }
  ctx->Ptot = ptot;
  ctx->tot_acc.assign(n_replicas, 0);
  ctx->tot_left.assign(n_replicas, 0);
  int st = HSMC_OK;
  auto fail = [&](int code) {
    hsmc_gpu_destroy(ctx);
    return code;
  };
  if (hipSetDevice(device) != hipSuccess) return fail(HSMC_ERR_HIP);
  // the allocation-free entry point's scratch array, resolved once per device
  if (hsmc::noalloc_scratch_init(device) != HSMC_OK) return fail(HSMC_ERR_HIP);
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(HSMC_ERR_HIP);
  if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess)
    return fail(HSMC_ERR_HIP);
  if (hipMalloc(&ctx->d_reps, sizeof(RepDev) * n_replicas) != hipSuccess ||
      hipMalloc(&ctx->d_swp, sizeof(SweepDev) * std::max(n_replicas, kSweepTable)) != hipSuccess ||
      hipMalloc(&ctx->d_aoff, sizeof(int) * (n_replicas + 1)) != hipSuccess ||
      hipMalloc(&ctx->d_prep, sizeof(int) * ptot) != hipSuccess ||
      hipMalloc(&ctx->d_x, sizeof(double) * ptot) != hipSuccess ||
      hipMalloc(&ctx->d_y, sizeof(double) * ptot) != hipSuccess ||
      hipMalloc(&ctx->d_z, sizeof(double) * ptot) != hipSuccess ||
      hipMalloc(&ctx->d_sweep, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&ctx->d_err, sizeof(unsigned int)) != hipSuccess ||
      hipMalloc(&ctx->d_bar, 2 * sizeof(unsigned int)) != hipSuccess)
    return fail(HSMC_ERR_NO_MEMORY);
  if (hipDeviceGetAttribute(&ctx->n_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
    return fail(HSMC_ERR_HIP);
  {
    void *pinned = nullptr;
    if (hipHostMalloc(&pinned, sizeof(unsigned int), hipHostMallocDefault) != hipSuccess)
      return fail(HSMC_ERR_NO_MEMORY);
    ctx->h_err = (volatile unsigned int *)pinned;
    *ctx->h_err = 0u;
  }
  if (hipMemcpy(ctx->d_prep, prep.data(), sizeof(int) * ptot, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(ctx->d_sweep, 0, sizeof(unsigned long long)) != hipSuccess ||
      hipMemset(ctx->d_swp, 0, sizeof(SweepDev) * std::max(n_replicas, kSweepTable)) != hipSuccess ||
      hipMemset(ctx->d_err, 0, sizeof(unsigned int)) != hipSuccess ||
      hipMemset(ctx->d_bar, 0, 2 * sizeof(unsigned int)) != hipSuccess ||
      hipMemset(ctx->d_x, 0, sizeof(double) * ptot) != hipSuccess ||
      hipMemset(ctx->d_y, 0, sizeof(double) * ptot) != hipSuccess ||
      hipMemset(ctx->d_z, 0, sizeof(double) * ptot) != hipSuccess)
    return fail(HSMC_ERR_HIP);
  st = layout(ctx);
  if (st) return fail(st);
  *out = ctx;
  return HSMC_OK;
}

int hsmc_gpu_destroy(hsmc_gpu_ctx *ctx) {
  if (!ctx) return HSMC_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  drop_graph(ctx);
  void *bufs[] = {ctx->d_reps, ctx->d_swp, ctx->d_aoff, ctx->d_arep, ctx->d_prep, ctx->d_x, ctx->d_y, ctx->d_z,
                  ctx->d_cnt, ctx->d_slot, ctx->d_sweep,
                  ctx->d_cacc, ctx->d_cleft, ctx->d_err, ctx->d_bar};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (int i = 0; i < hsmc_gpu_ctx::kScratch; ++i)
    if (ctx->scratch[i]) (void)hipFree(ctx->scratch[i]);
  if (ctx->h_err) (void)hipHostFree((void *)ctx->h_err);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return HSMC_OK;
}

int hsmc_gpu_sync(hsmc_gpu_ctx *ctx) {
  if (!ctx) return HSMC_ERR_BAD_ARG;
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return check_device_errors(ctx);
}

int hsmc_gpu_upload(hsmc_gpu_ctx *ctx, int replica, const double *conf4) {
  if (!ctx || replica < 0 || replica >= ctx->R || !conf4) return HSMC_ERR_BAD_ARG;
  const RepHost &h = ctx->reps[replica];
  const int n = h.p.n_part;
  std::vector<double> x(n), y(n), z(n);
  for (int i = 0; i < n; ++i) {
    x[i] = conf4[4 * (size_t)i + 1];
    y[i] = conf4[4 * (size_t)i + 2];
    z[i] = conf4[4 * (size_t)i + 3];
    if (!(x[i] >= 0.0 && x[i] <= h.p.lx && y[i] >= 0.0 && y[i] <= h.p.ly && z[i] >= 0.0 &&
          z[i] <= h.p.lz)) {
      ctx->last_error = "particle outside [0, L] (apply_pbc range)";
      return HSMC_ERR_BAD_ARG;
    }
  }
  HSMC_HIP(ctx, hipMemcpyAsync(ctx->d_x + h.poff, x.data(), sizeof(double) * n,
                               hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(ctx->d_y + h.poff, y.data(), sizeof(double) * n,
                               hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(ctx->d_z + h.poff, z.data(), sizeof(double) * n,
                               hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->version++;
  return HSMC_OK;
}

int hsmc_gpu_download(hsmc_gpu_ctx *ctx, int replica, double *conf4) {
  if (!ctx || replica < 0 || replica >= ctx->R || !conf4) return HSMC_ERR_BAD_ARG;
  const RepHost &h = ctx->reps[replica];
  const int n = h.p.n_part;
  std::vector<double> x(n), y(n), z(n);
  HSMC_HIP(ctx, hipMemcpyAsync(x.data(), ctx->d_x + h.poff, sizeof(double) * n,
                               hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(y.data(), ctx->d_y + h.poff, sizeof(double) * n,
                               hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(z.data(), ctx->d_z + h.poff, sizeof(double) * n,
                               hipMemcpyDeviceToHost, ctx->stream));
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < n; ++i) {
    conf4[4 * (size_t)i] = (double)i;
    conf4[4 * (size_t)i + 1] = x[i];
    conf4[4 * (size_t)i + 2] = y[i];
    conf4[4 * (size_t)i + 3] = z[i];
  }
  return check_device_errors(ctx);
}

int hsmc_gpu_set_particles(hsmc_gpu_ctx *ctx, int replica, const int32_t *idx,
                           const double *xyz, int m) {
  if (!ctx || replica < 0 || replica >= ctx->R || m < 0 || (m > 0 && (!idx || !xyz)))
    return HSMC_ERR_BAD_ARG;
  if (m == 0) return HSMC_OK;
  const RepHost &h = ctx->reps[replica];
  for (int k = 0; k < m; ++k) {
    if (idx[k] < 0 || idx[k] >= h.p.n_part) return HSMC_ERR_BAD_ARG;
    const double *q = xyz + 3 * (size_t)k;
    if (!(q[0] >= 0.0 && q[0] <= h.p.lx && q[1] >= 0.0 && q[1] <= h.p.ly && q[2] >= 0.0 &&
          q[2] <= h.p.lz))
      return HSMC_ERR_BAD_ARG;
  }
  int *didx = (int *)scratch(ctx, 21, sizeof(int) * m);
  double *dq = (double *)scratch(ctx, 22, sizeof(double) * 3 * m);
  if (!didx || !dq) return HSMC_ERR_NO_MEMORY;
  HSMC_HIP(ctx, hipMemcpyAsync(didx, idx, sizeof(int) * m, hipMemcpyHostToDevice, ctx->stream));
  HSMC_HIP(ctx, hipMemcpyAsync(dq, xyz, sizeof(double) * 3 * m, hipMemcpyHostToDevice, ctx->stream));
  hipLaunchKernelGGL(k_set_particles, dim3((m + 255) / 256), dim3(256), 0, ctx->stream,
                     ctx->d_x + h.poff, ctx->d_y + h.poff, ctx->d_z + h.poff, didx, dq, m);
  HSMC_HIP(ctx, hipGetLastError());
  HSMC_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->version++;
  return HSMC_OK;
}

int hsmc_gpu_set_box(hsmc_gpu_ctx *ctx, int replica, double lx, double ly, double lz) {
  if (!ctx || replica < 0 || replica >= ctx->R) return HSMC_ERR_BAD_ARG;
  if (!valid_box(lx) || !valid_box(ly) || !valid_box(lz)) return HSMC_ERR_BAD_ARG;
  hsmc_gpu_replica &p = ctx->reps[replica].p;
  p.lx = lx; p.ly = ly; p.lz = lz;
  return layout(ctx);
}

int hsmc_gpu_scale(hsmc_gpu_ctx *ctx, int replica, double sf, double lx, double ly, double lz) {
  if (!ctx || replica < 0 || replica >= ctx->R || !(sf > 0.0)) return HSMC_ERR_BAD_ARG;
  if (!valid_box(lx) || !valid_box(ly) || !valid_box(lz)) return HSMC_ERR_BAD_ARG;
  RepHost &h = ctx->reps[replica];
  const int n = h.p.n_part;
  hipLaunchKernelGGL(k_scale, dim3((n + 255) / 256), dim3(256), 0, ctx->stream,
                     ctx->d_x + h.poff, ctx->d_y + h.poff, ctx->d_z + h.poff, n, sf, lx, ly, lz);
  HSMC_HIP(ctx, hipGetLastError());
  h.p.lx = lx; h.p.ly = ly; h.p.lz = lz;
  return layout(ctx);
}

int hsmc_gpu_set_dr_max(hsmc_gpu_ctx *ctx, int replica, double dr_max) {
  if (!ctx || replica < 0 || replica >= ctx->R || !(dr_max >= 0.0)) return HSMC_ERR_BAD_ARG;
  ctx->reps[replica].p.dr_max = dr_max;
  return layout(ctx);
}

int hsmc_gpu_grid(const hsmc_gpu_ctx *ctx, int replica, int32_t out_g[3]) {
  if (!ctx || replica < 0 || replica >= ctx->R || !out_g) return HSMC_ERR_BAD_ARG;
  out_g[0] = ctx->reps[replica].gx;
  out_g[1] = ctx->reps[replica].gy;
  out_g[2] = ctx->reps[replica].gz;
  return HSMC_OK;
}

}  // extern "C"
