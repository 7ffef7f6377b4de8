/*
 * state.c -- box, particles, RNG, cell list and moves of the host layer.
 *
 * The configuration lives on the GPU (libhsmc_gpu.so); this module keeps the
 * reference's host-side view of it (part_config_get returns a
 * double(*)[4], sim_info.c:188) and moves data only when one side is behind:
 *   dev_behind  - the host copy was written (lattice, restart, part_move)
 *   host_behind - the GPU moved particles (sweeps, volume moves)
 *   cells_behind - the host cell list no longer matches the configuration
 * Every overlap decision, distance and cell list is computed on the GPU.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "hsmc.h"
#include "hsmc_gpu.h"
#include "host_internal.h"

static uint64_t pending_sweep_counter = 0;

/* ---- box and particles (sim_info.c) ---- */
static box_info sim_box;
static p_info part_inf;
static config part_conf = NULL;
static int dev_behind = 1, host_behind = 0, cells_behind = 1;

static hsmc_gpu_ctx *gpu = NULL;

hsmc_gpu_ctx *hsmc_host_gpu(void) { return gpu; }

void host_gpu_check(int st, const char *where) {
  if (st != HSMC_OK) {
    printf("ERROR: %s: %s (%s)\n", where, hsmc_gpu_strerror(st),
           gpu ? hsmc_gpu_last_error(gpu) : "");
    exit(EXIT_FAILURE);
  }
}

void *host_alloc(size_t bytes) {
  void *p = malloc(bytes > 0 ? bytes : 1);
  if (!p) {
    printf("ERROR: out of host memory (%zu bytes)\n", bytes);
    exit(EXIT_FAILURE);
  }
  return p;
}

/* sim_box_init (sim_info.c:32-71) */
void sim_box_init(int cell_type, int nx, int ny, int nz, double rho) {
  int part_cell;
  if (cell_type == 1) part_cell = 1;
  else if (cell_type == 2) part_cell = 4;
  else {
    printf("Unknown lattice type, default to fcc\n");
    part_cell = 4;
  }
  double cell_vol = part_cell / rho;
  double cell_size = pow(cell_vol, 1. / 3.);
  int min_n = nx;
  if (ny < min_n) min_n = ny;
  if (nz < min_n) min_n = nz;
  sim_box.cell_x = nx;
  sim_box.cell_y = ny;
  sim_box.cell_z = nz;
  sim_box.lx = nx * cell_size;
  sim_box.ly = ny * cell_size;
  sim_box.lz = nz * cell_size;
  sim_box.min_size = min_n * cell_size;
  sim_box.vol = nx * ny * nz * cell_vol;
  sim_box.cell_size = cell_size;
  sim_box.cell_type = cell_type;
}

/* part_alloc (sim_info.c:76-109) */
void part_alloc(void) {
  int part_cell;
  if (sim_box.cell_type == 1) part_cell = 1;
  else if (sim_box.cell_type == 2) part_cell = 4;
  else {
    printf("Unknown lattice type, default to fcc\n");
    part_cell = 4;
  }
  int part_tot = sim_box.cell_x * sim_box.cell_y * sim_box.cell_z * part_cell;
  free(part_conf);
  part_conf = host_alloc(part_tot * sizeof(*part_conf));
  if (part_conf == NULL) {
    printf("ERROR: Failed particle allocation\n");
    exit(EXIT_FAILURE);
  }
  part_inf.Ncell = part_cell;
  part_inf.NN = part_tot;
  dev_behind = 1;
}

void add_particle(int id, double xx, double yy, double zz) {
  part_conf[id][0] = id;
  part_conf[id][1] = xx;
  part_conf[id][2] = yy;
  part_conf[id][3] = zz;
}

void part_init_err(void) {
  printf("Overlap in the initial configuration. Possible solutions:\n");
  printf("-- If SC lattice was selected, try to change to FCC\n");
  printf("-- If FCC lattice was selected, the selected value of density is unphysical\n");
  exit(EXIT_FAILURE);
}

void part_init_sc(void) {   /* sim_info.c:125-138 */
  double aa = sim_box.cell_size;
  if (aa < 1) part_init_err();
  int pid = 0;
  for (int i = 0; i < sim_box.cell_x; i++)
    for (int j = 0; j < sim_box.cell_y; j++)
      for (int k = 0; k < sim_box.cell_z; k++) {
        add_particle(pid, i * aa, j * aa, k * aa);
        pid++;
      }
}

void part_init_fcc(void) {  /* sim_info.c:140-156 */
  double aa = sim_box.cell_size;
  if (aa / sqrt(2.0) < 1) part_init_err();
  int pid = 0;
  for (int i = 0; i < sim_box.cell_x; i++)
    for (int j = 0; j < sim_box.cell_y; j++)
      for (int k = 0; k < sim_box.cell_z; k++) {
        add_particle(pid, i * aa, j * aa, k * aa);
        add_particle(pid + 1, (i + 0.5) * aa, (j + 0.5) * aa, k * aa);
        add_particle(pid + 2, (i + 0.5) * aa, j * aa, (k + 0.5) * aa);
        add_particle(pid + 3, i * aa, (j + 0.5) * aa, (k + 0.5) * aa);
        pid += 4;
      }
}

void part_init(void) {
  if (sim_box.cell_type == 1) part_init_sc();
  else part_init_fcc();
  dev_behind = 1;
  host_behind = 0;
  cells_behind = 1;
}

box_info sim_box_info_get(void) { return sim_box; }
p_info part_info_get(void) { return part_inf; }

config part_config_get(void) {
  if (host_behind && gpu) {
    host_gpu_check(hsmc_gpu_download(gpu, 0, &part_conf[0][0]), "download");
    host_behind = 0;
  }
  return part_conf;
}

void part_config_changed(void) {
  dev_behind = 1;
  cells_behind = 1;
}

void host_sync_device(void) {
  if (dev_behind && gpu) {
    host_gpu_check(hsmc_gpu_upload(gpu, 0, &part_conf[0][0]), "upload");
    dev_behind = 0;
  }
}

void host_device_moved(void) {
  host_behind = 1;
  cells_behind = 1;
}

void part_free(void) {
  free(part_conf);
  part_conf = NULL;
}

void print_sim_info(void) {   /* sim_info.c:193-216 */
  printf("Simulation box size (x, y, z): %.5f %.5f %.5f\n", sim_box.lx, sim_box.ly, sim_box.lz);
  printf("Number of particles: %d\n", part_inf.NN);
  if (G_IN.press > 0) printf("Pressure: %.8f\n", G_IN.press);
}

void sim_box_info_write(FILE *fid) { fwrite(&sim_box, sizeof(box_info), 1, fid); }
void sim_box_info_read(FILE *fid) {
  if (fread(&sim_box, sizeof(box_info), 1, fid) != 1) {
    printf("ERROR: short restart file (box)\n");
    exit(EXIT_FAILURE);
  }
}
void part_info_write(FILE *fid) { fwrite(&part_inf, sizeof(p_info), 1, fid); }
void part_info_read(FILE *fid) {
  if (fread(&part_inf, sizeof(p_info), 1, fid) != 1) {
    printf("ERROR: short restart file (particles)\n");
    exit(EXIT_FAILURE);
  }
}
void part_conf_write(FILE *fid) {
  part_config_get();
  fwrite(part_conf, sizeof(double), (size_t)part_inf.NN * 4, fid);
}
void part_conf_read(FILE *fid) {
  if (fread(part_conf, sizeof(double), (size_t)part_inf.NN * 4, fid) != (size_t)part_inf.NN * 4) {
    printf("ERROR: short restart file (configuration)\n");
    exit(EXIT_FAILURE);
  }
  dev_behind = 1;
  host_behind = 0;
  cells_behind = 1;
}

/* ---- RNG: MT19937 with GSL's seeding and draws (rng.c) ---- */
static unsigned long mt[624];   /* the layout gsl_rng_fwrite writes */
static int mti = 624;

void rng_seed(unsigned long s) {
  if (s == 0) s = 4357;   /* GSL's default seed for mt19937 */
  mt[0] = s & 0xffffffffUL;
  for (int i = 1; i < 624; i++)
    mt[i] = (1812433253UL * (mt[i - 1] ^ (mt[i - 1] >> 30)) + (unsigned long)i) & 0xffffffffUL;
  mti = 624;
}

static unsigned long mt_next(void) {
  if (mti >= 624) {
    for (int i = 0; i < 624; i++) {
      unsigned long y = (mt[i] & 0x80000000UL) | (mt[(i + 1) % 624] & 0x7fffffffUL);
      unsigned long v = mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1UL) v ^= 0x9908b0dfUL;
      mt[i] = v;
    }
    mti = 0;
  }
  unsigned long y = mt[mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680UL;
  y ^= (y << 15) & 0xefc60000UL;
  y ^= (y >> 18);
  return y & 0xffffffffUL;
}

void rng_init(void) { rng_seed(G_IN.seed); }
double rng_get_double(void) { return (double)mt_next() / (double)4294967295.0; }
int rng_get_int(int MM) {
  unsigned long scale = 0xffffffffUL / (unsigned long)MM;
  unsigned long k;
  do { k = mt_next() / scale; } while (k >= (unsigned long)MM);
  return (int)k;
}
void rng_free(void) {}
void rng_write(FILE *fid) {
  /* mt19937 state as GSL lays it out: unsigned long mt[624]; int mti */
  struct { unsigned long m[624]; int i; } st;
  memcpy(st.m, mt, sizeof(mt));
  st.i = mti;
  fwrite(&st, sizeof(st), 1, fid);
}
void rng_read(FILE *fid) {
  struct { unsigned long m[624]; int i; } st;
  if (fread(&st, sizeof(st), 1, fid) != 1) {
    printf("ERROR: short restart file (random number generator)\n");
    exit(EXIT_FAILURE);
  }
  memcpy(mt, st.m, sizeof(mt));
  mti = st.i;
}

/* Philox key of the GPU streams: the input seed (0 selects 4357 as for MT) */
uint64_t host_philox_seed(void) { return G_IN.seed ? (uint64_t)G_IN.seed : 4357u; }

/* ---- cell list (cell_list.c) ---- */
static cl_info neigh_info;
static int cells_alloc_rows = 0;

void cell_list_alloc_arr(int ***arr, int rows, int cols) {
  *arr = (int **)host_alloc(rows * sizeof(int *));
  for (int i = 0; i < rows; i++) (*arr)[i] = (int *)host_alloc(cols * sizeof(int));
}

void cell_list_free_arr(int ***arr, int rows) {
  if (!*arr) return;
  for (int i = 0; i < rows; i++) free((*arr)[i]);
  free(*arr);
  *arr = NULL;
}

/* compute_cell_list_info (cell_list.c:92-131) */
void compute_cell_list_info(bool init) {
  box_info b = sim_box;
  if (init) {
    neigh_info.num_x = (int)floor(b.lx / G_IN.neigh_dr);
    neigh_info.num_y = (int)floor(b.ly / G_IN.neigh_dr);
    neigh_info.num_z = (int)floor(b.lz / G_IN.neigh_dr);
    neigh_info.num_tot = neigh_info.num_x * neigh_info.num_y * neigh_info.num_z;
    if (neigh_info.num_tot < 27) {
      printf("WARNING: The specified cell size has been reduced in order to accomodate 27 cells in the simulation box\n");
      neigh_info.num_x = 3;
      neigh_info.num_y = 3;
      neigh_info.num_z = 3;
      neigh_info.num_tot = 27;   /* kept consistent with the 3x3x3 grid */
    }
    neigh_info.neigh_num = 27;
  }
  neigh_info.size_x = b.lx / neigh_info.num_x;
  neigh_info.size_y = b.ly / neigh_info.num_y;
  neigh_info.size_z = b.lz / neigh_info.num_z;
  if (neigh_info.size_x < 1.0 || neigh_info.size_y < 1.0 || neigh_info.size_z < 1.0) {
    printf("ERROR: size of cells in cell list is smaller than the size of the particles\n");
    exit(EXIT_FAILURE);
  }
}

static void ensure_gpu(void) {
  if (gpu) return;
  hsmc_gpu_replica rep;
  memset(&rep, 0, sizeof(rep));
  rep.n_part = part_inf.NN;
  rep.lx = sim_box.lx;
  rep.ly = sim_box.ly;
  rep.lz = sim_box.lz;
  rep.dr_max = G_IN.dr_max;
  rep.seed = host_philox_seed();
  host_gpu_check(hsmc_gpu_create(&gpu, G_IN.gpu_device, 1, &rep), "hsmc_gpu_create");
  dev_behind = 1;
  if (pending_sweep_counter) {
    host_gpu_check(hsmc_gpu_set_sweep_counter(gpu, pending_sweep_counter), "set_sweep_counter");
    pending_sweep_counter = 0;
  }
}

/* Sweep-stream position for restart files: read from the context, or set
 * (now, or when the context is created). */
uint64_t host_sweep_counter(void) {
  if (!gpu) return pending_sweep_counter;
  uint64_t c[4];
  host_gpu_check(hsmc_gpu_counters(gpu, 0, c), "counters");
  return c[3];
}

void host_set_sweep_counter(uint64_t sweep) {
  if (gpu) host_gpu_check(hsmc_gpu_set_sweep_counter(gpu, sweep), "set_sweep_counter");
  else pending_sweep_counter = sweep;
}

/* Box of the GPU context follows sim_box (after restart or volume moves). */
void host_gpu_box_sync(void) {
  if (gpu) host_gpu_check(hsmc_gpu_set_box(gpu, 0, sim_box.lx, sim_box.ly, sim_box.lz), "set_box");
}

/* cell_list_new (cell_list.c:142-175), built on the GPU in the reference's
 * fresh-build layout (ascending ids per cell). */
void cell_list_new(void) {
  compute_cell_list_info(false);
  ensure_gpu();
  host_sync_device();
  const int C = neigh_info.num_tot, mp = G_IN.neigh_max_part;
  int *pc = (int *)host_alloc(sizeof(int) * (size_t)C * mp);
  int st = hsmc_gpu_cell_list(gpu, 0, neigh_info.num_x, neigh_info.num_y, neigh_info.num_z, mp, pc,
                              NULL);
  if (st == HSMC_ERR_CELL_OVERFLOW) {
    printf("ERROR: More than %d particles in one cell\n", G_IN.neigh_max_part);
    exit(EXIT_FAILURE);
  }
  host_gpu_check(st, "cell_list");
  for (int c = 0; c < C; c++) memcpy(neigh_info.part_cell[c], pc + (size_t)c * mp, sizeof(int) * mp);
  free(pc);
  cells_behind = 0;
}

/* neigh_id (cell_list.c:266-301), index ix*ny*nz + iy*nz + iz */
void neigh_id(int a, int b, int c) {
  const int nx = neigh_info.num_x, ny = neigh_info.num_y, nz = neigh_info.num_z;
  int ref = (a * ny + b) * nz + c, counter = 0;
  for (int i = -1; i < 2; i++)
    for (int j = -1; j < 2; j++)
      for (int k = -1; k < 2; k++) {
        int ix = a + i, iy = b + j, iz = c + k;
        if (ix > nx - 1) ix -= nx; else if (ix < 0) ix += nx;
        if (iy > ny - 1) iy -= ny; else if (iy < 0) iy += ny;
        if (iz > nz - 1) iz -= nz; else if (iz < 0) iz += nz;
        neigh_info.neigh_mat[ref][counter++] = (ix * ny + iy) * nz + iz;
      }
}

void cell_neigh_init(void) {
  for (int a = 0; a < neigh_info.num_x; a++)
    for (int b = 0; b < neigh_info.num_y; b++)
      for (int c = 0; c < neigh_info.num_z; c++) neigh_id(a, b, c);
}

/* cell_list_init (cell_list.c:33-58): also creates the GPU context and
 * sends the configuration. */
void cell_list_init(bool alloc) {
  compute_cell_list_info(true);
  if (alloc) {
    if (gpu) {
      hsmc_gpu_destroy(gpu);
      gpu = NULL;
    }
    cell_list_free_arr(&neigh_info.part_cell, cells_alloc_rows);
    cell_list_free_arr(&neigh_info.neigh_mat, cells_alloc_rows);
    cell_list_alloc_arr(&neigh_info.part_cell, neigh_info.num_tot, G_IN.neigh_max_part);
    cell_list_alloc_arr(&neigh_info.neigh_mat, neigh_info.num_tot, neigh_info.neigh_num);
    cells_alloc_rows = neigh_info.num_tot;
  }
  ensure_gpu();
  host_gpu_box_sync();
  dev_behind = 1;
  cell_list_new();
  cell_neigh_init();
}

void cell_list_free(void) {
  cell_list_free_arr(&neigh_info.part_cell, cells_alloc_rows);
  cell_list_free_arr(&neigh_info.neigh_mat, cells_alloc_rows);
  cells_alloc_rows = 0;
  if (gpu) {
    hsmc_gpu_destroy(gpu);
    gpu = NULL;
  }
}

cl_info get_cell_list_info(void) {
  if (cells_behind) cell_list_new();
  return neigh_info;
}

/* cell_list_check / add / del / update (cell_list.c:180-233): host
 * bookkeeping of the serial chain */
void cell_list_check(int c) {
  if (neigh_info.part_cell[c][0] < 0) {
    printf("ERROR: Trying to remove one particle from an empty cell\n");
    exit(EXIT_FAILURE);
  }
  if (neigh_info.part_cell[c][0] > G_IN.neigh_max_part - 1) {
    printf("ERROR: More than %d particles in one cell\n", G_IN.neigh_max_part);
    exit(EXIT_FAILURE);
  }
}

void cell_list_add(int c, int id) {
  int n = neigh_info.part_cell[c][0];
  neigh_info.part_cell[c][0] += 1;
  cell_list_check(c);
  neigh_info.part_cell[c][n + 1] = id;
}

void cell_list_del(int c, int id) {
  int n = neigh_info.part_cell[c][0];
  int idx_remove = n;
  bool shift = false;
  neigh_info.part_cell[c][0] -= 1;
  cell_list_check(c);
  for (int i = 1; i <= n; i++) {
    if (neigh_info.part_cell[c][i] == id) {
      idx_remove = i;
      neigh_info.part_cell[c][i] = -1;
      shift = true;
    }
    if (shift && i > idx_remove) {
      neigh_info.part_cell[c][i - 1] = neigh_info.part_cell[c][i];
      neigh_info.part_cell[c][i] = -1;
    }
  }
  if (!shift) {
    printf("ERROR: Attempt to delete particle %d from cell %d, but particle is not in cell\n", id, c);
    exit(EXIT_FAILURE);
  }
}

void cell_list_update(int cdel, int cadd, int id) {
  cell_list_del(cdel, id);
  cell_list_add(cadd, id);
}

static int cell_of(double x, double y, double z) {   /* cell_list.c:244-246 */
  int ix = (int)(x / neigh_info.size_x), iy = (int)(y / neigh_info.size_y),
      iz = (int)(z / neigh_info.size_z);
  if (ix >= neigh_info.num_x) ix = neigh_info.num_x - 1;   /* x == L */
  if (iy >= neigh_info.num_y) iy = neigh_info.num_y - 1;
  if (iz >= neigh_info.num_z) iz = neigh_info.num_z - 1;
  return (ix * neigh_info.num_y + iy) * neigh_info.num_z + iz;
}

int cell_part_idx(int id) {
  config c = part_config_get();
  return cell_of(c[id][1], c[id][2], c[id][3]);
}

/* ---- moves (moves.c) ---- */
static uint64_t part_moves = 0, acc_part_moves = 0, rej_part_moves = 0;
static uint64_t vol_moves = 0, acc_vol_moves = 0, rej_vol_moves = 0;

/* apply_pbc (moves.c:215-226) on the host copy */
void apply_pbc(int idx) {
  config c = part_config_get();
  if (c[idx][1] > sim_box.lx) c[idx][1] -= sim_box.lx;
  else if (c[idx][1] < 0.0) c[idx][1] += sim_box.lx;
  if (c[idx][2] > sim_box.ly) c[idx][2] -= sim_box.ly;
  else if (c[idx][2] < 0.0) c[idx][2] += sim_box.ly;
  if (c[idx][3] > sim_box.lz) c[idx][3] -= sim_box.lz;
  else if (c[idx][3] < 0.0) c[idx][3] += sim_box.lz;
  part_config_changed();
}

/* check_overlap (moves.c:157-212): overlap of particle idx at its current
 * coordinates with every other particle after scaling by sf. */
bool check_overlap(int idx_ref, double sf_x, double sf_y, double sf_z) {
  if (sf_x != sf_y || sf_y != sf_z) {
    printf("ERROR: check_overlap supports one scale factor for all axes\n");
    exit(EXIT_FAILURE);
  }
  ensure_gpu();
  host_sync_device();
  int32_t i = idx_ref;
  uint8_t out = 0;
  host_gpu_check(hsmc_gpu_overlap_scaled(gpu, 0, &i, &sf_x, 1, &out), "overlap_scaled");
  return out != 0;
}

/* compute_dist (moves.c:400-431) */
double compute_dist(int idx1, int idx2, double sf_x, double sf_y, double sf_z) {
  if (sf_x != sf_y || sf_y != sf_z) {
    printf("ERROR: compute_dist supports one scale factor for all axes\n");
    exit(EXIT_FAILURE);
  }
  ensure_gpu();
  host_sync_device();
  int32_t i = idx1, j = idx2;
  double d = 0;
  host_gpu_check(hsmc_gpu_pair_dist(gpu, 0, &i, &j, 1, sf_x, &d), "pair_dist");
  return d;
}

/* part_move (moves.c:27-80): one move of a random particle, decided by the
 * GPU overlap test; the host cell list is updated as the reference does. */
void part_move(void) {
  ensure_gpu();
  config c = part_config_get();
  host_sync_device();
  if (cells_behind) cell_list_new();
  int r_idx = rng_get_int(part_inf.NN);
  double r_x = rng_get_double(), r_y = rng_get_double(), r_z = rng_get_double();
  int cell_old = cell_of(c[r_idx][1], c[r_idx][2], c[r_idx][3]);
  double q[3] = {c[r_idx][1] + (r_x - 0.5) * G_IN.dr_max, c[r_idx][2] + (r_y - 0.5) * G_IN.dr_max,
                 c[r_idx][3] + (r_z - 0.5) * G_IN.dr_max};
  if (q[0] > sim_box.lx) q[0] -= sim_box.lx; else if (q[0] < 0.0) q[0] += sim_box.lx;
  if (q[1] > sim_box.ly) q[1] -= sim_box.ly; else if (q[1] < 0.0) q[1] += sim_box.ly;
  if (q[2] > sim_box.lz) q[2] -= sim_box.lz; else if (q[2] < 0.0) q[2] += sim_box.lz;
  int32_t idx = r_idx;
  uint8_t ov = 0;
  host_gpu_check(hsmc_gpu_overlap_batch(gpu, 0, &idx, q, 1, &ov), "overlap_batch");
  if (ov) {
    rej_part_moves += 1;
  } else {
    c[r_idx][1] = q[0];
    c[r_idx][2] = q[1];
    c[r_idx][3] = q[2];
    host_gpu_check(hsmc_gpu_set_particles(gpu, 0, &idx, q, 1), "set_particles");
    int cell_new = cell_of(q[0], q[1], q[2]);
    if (cell_new != cell_old) cell_list_update(cell_old, cell_new, r_idx);
    acc_part_moves += 1;
  }
  part_moves += 1;
}

/* vol_move (moves.c:83-153): proposal and acceptance on the host (glibc
 * log/exp/pow and the MT stream), overlap scan and rescale on the GPU. */
void vol_move(void) {
  ensure_gpu();
  host_sync_device();
  double r_dv = rng_get_double();
  double log_vol_new = log(sim_box.vol) + (r_dv - 0.5) * G_IN.dv_max;
  double vol_new = exp(log_vol_new);
  double vol_ratio = vol_new / sim_box.vol;
  double sf = pow(vol_ratio, 1. / 3.);
  uint8_t no_ov = 0;
  host_gpu_check(hsmc_gpu_any_overlap_scaled(gpu, 0, &sf, 1, &no_ov), "any_overlap_scaled");
  if (no_ov) {
    double boltz = exp(G_IN.press * (sim_box.vol - vol_new) + (part_inf.NN + 1) * log(vol_ratio));
    double r_acc = rng_get_double();
    if (r_acc >= boltz) {
      rej_vol_moves += 1;
    } else {
      acc_vol_moves += 1;
      G_IN.rho = part_inf.NN / vol_new;
      sim_box_init(G_IN.type, G_IN.nx, G_IN.ny, G_IN.nz, G_IN.rho);
      host_gpu_check(hsmc_gpu_scale(gpu, 0, sf, sim_box.lx, sim_box.ly, sim_box.lz), "scale");
      host_device_moved();
      cell_list_new();
    }
  } else {
    rej_vol_moves += 1;
  }
  vol_moves += 1;
}

void host_count_sweep_moves(uint64_t attempts, uint64_t accepted) {
  part_moves += attempts;
  acc_part_moves += accepted;
  rej_part_moves += attempts - accepted;
}

void get_moves_counters64(uint64_t out[6]) {
  out[0] = part_moves; out[1] = acc_part_moves; out[2] = rej_part_moves;
  out[3] = vol_moves; out[4] = acc_vol_moves; out[5] = rej_vol_moves;
}

/* get_moves_counters (moves.c:231-241): truncated to int like the reference */
void get_moves_counters(int *pm, int *apm, int *rpm, int *vm, int *avm, int *rvm) {
  if (pm) *pm = (int)part_moves;
  if (apm) *apm = (int)acc_part_moves;
  if (rpm) *rpm = (int)rej_part_moves;
  if (vm) *vm = (int)vol_moves;
  if (avm) *avm = (int)acc_vol_moves;
  if (rvm) *rvm = (int)rej_vol_moves;
}

void reset_moves_counters(void) {
  part_moves = acc_part_moves = rej_part_moves = 0;
  vol_moves = acc_vol_moves = rej_vol_moves = 0;
}
