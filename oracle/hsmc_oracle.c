/*
 * hsmc_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see hsmc_oracle.h).
 *
 * Clean-room restatement of fedluc/HSMC's hot path.  Compiled with
 * gcc -O2 -std=gnu99 -ffp-contract=off (the reference's flags, src/Makefile:3,
 * plus no FMA contraction so every FP64 decision has the reference's
 * rounding).  Citations are to the reference's src/ files.
 */
#define _GNU_SOURCE
#include "hsmc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ===================================================================== */
/* MT19937 (Matsumoto & Nishimura, 2002 seeding) with GSL's conventions.  */
/* ===================================================================== */

/* gsl_rng_set(rng_mt, G_IN.seed) (rng.c:24): GSL documents that a zero
 * seed selects its default seed 4357 for mt19937.  That rule is taken from
 * GSL's documentation; GSL is absent here (parity unpinned). */
void orc_mt_seed(orc_mt *m, unsigned long seed) {
  uint32_t s = (uint32_t)(seed & 0xffffffffUL);
  if (seed == 0) s = 4357u;
  m->mt[0] = s;
  for (int i = 1; i < 624; i++) {
    uint32_t p = m->mt[i - 1];
    m->mt[i] = (uint32_t)(1812433253u * (p ^ (p >> 30)) + (uint32_t)i);
  }
  m->mti = 624;
}

uint32_t orc_mt_next(orc_mt *m) {
  if (m->mti >= 624) {
    for (int i = 0; i < 624; i++) {
      uint32_t y = (m->mt[i] & 0x80000000u) | (m->mt[(i + 1) % 624] & 0x7fffffffu);
      uint32_t v = m->mt[(i + 397) % 624] ^ (y >> 1);
      if (y & 1u) v ^= 0x9908b0dfu;
      m->mt[i] = v;
    }
    m->mti = 0;
  }
  uint32_t y = m->mt[m->mti++];
  y ^= (y >> 11);
  y ^= (y << 7) & 0x9d2c5680u;
  y ^= (y << 15) & 0xefc60000u;
  y ^= (y >> 18);
  return y;
}

/* rng_get_double (rng.c:29-31): gsl_rng_get / gsl_rng_max, in [0,1]. */
double orc_mt_double(orc_mt *m) {
  return (double)orc_mt_next(m) / (double)4294967295.0;
}

/* rng_get_int (rng.c:34-36) = gsl_rng_uniform_int(r, n): GSL's published
 * scale-and-reject rule: scale = range / n, k = draw / scale, retry k >= n. */
int orc_mt_int(orc_mt *m, int n) {
  if (n <= 0) return 0;
  uint32_t scale = 0xffffffffu / (uint32_t)n;
  uint32_t k;
  do { k = orc_mt_next(m) / scale; } while (k >= (uint32_t)n);
  return (int)k;
}

orc_mt *orc_mt_new(unsigned long seed) {
  orc_mt *m = (orc_mt *)malloc(sizeof(orc_mt));
  if (m) orc_mt_seed(m, seed);
  return m;
}
void orc_mt_free(orc_mt *m) { free(m); }

/* ===================================================================== */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).                     */
/* ===================================================================== */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ===================================================================== */
/* Box and lattice (sim_info.c)                                           */
/* ===================================================================== */

/* sim_box_init (sim_info.c:32-71) */
int orc_box_init(orc_sys *s, int type, int nx, int ny, int nz, double rho) {
  int part_cell = (type == 1) ? 1 : 4;        /* :40-49 (unknown -> fcc) */
  double cell_vol = part_cell / rho;          /* :51 */
  double cell_size = pow(cell_vol, 1. / 3.);  /* :52 */
  int min_n = nx;
  if (ny < min_n) min_n = ny;
  if (nz < min_n) min_n = nz;
  s->cell_x = nx; s->cell_y = ny; s->cell_z = nz;
  s->lx = nx * cell_size;                     /* :63-65 */
  s->ly = ny * cell_size;
  s->lz = nz * cell_size;
  s->min_size = min_n * cell_size;
  s->vol = nx * ny * nz * cell_vol;           /* :67 */
  s->cell_size = cell_size;
  s->cell_type = type;
  return ORC_OK;
}

static void add_particle(orc_sys *s, int id, double x, double y, double z) {
  double *p = s->conf + 4 * (size_t)id;       /* sim_info.c:159-166 */
  p[0] = id; p[1] = x; p[2] = y; p[3] = z;
}

/* part_init / part_init_sc / part_init_fcc (sim_info.c:114-156) */
int orc_part_init(orc_sys *s) {
  double aa = s->cell_size;
  int pid = 0;
  if (s->cell_type == 1) {
    if (aa < 1) return ORC_INIT_OVERLAP;                 /* :128 */
    for (int i = 0; i < s->cell_x; i++)
      for (int j = 0; j < s->cell_y; j++)
        for (int k = 0; k < s->cell_z; k++) {
          add_particle(s, pid, i * aa, j * aa, k * aa);
          pid++;
        }
  } else {
    if (aa / sqrt(2.0) < 1) return ORC_INIT_OVERLAP;     /* :143 */
    for (int i = 0; i < s->cell_x; i++)
      for (int j = 0; j < s->cell_y; j++)
        for (int k = 0; k < s->cell_z; k++) {
          add_particle(s, pid, i * aa, j * aa, k * aa);                       /* :148 */
          add_particle(s, pid + 1, (i + 0.5) * aa, (j + 0.5) * aa, k * aa);   /* :149 */
          add_particle(s, pid + 2, (i + 0.5) * aa, j * aa, (k + 0.5) * aa);   /* :150 */
          add_particle(s, pid + 3, i * aa, (j + 0.5) * aa, (k + 0.5) * aa);   /* :151 */
          pid += 4;
        }
  }
  return ORC_OK;
}

/* ===================================================================== */
/* Cell list (cell_list.c)                                                */
/* ===================================================================== */

/* compute_cell_list_info (cell_list.c:92-131) */
static int cell_info(orc_sys *s, int init) {
  if (init) {
    s->num_x = (int)floor(s->lx / s->neigh_dr);     /* :100-102 */
    s->num_y = (int)floor(s->ly / s->neigh_dr);
    s->num_z = (int)floor(s->lz / s->neigh_dr);
    s->num_tot = s->num_x * s->num_y * s->num_z;
    /* :108-113 forces a 3x3x3 grid without updating num_tot, after which the
     * reference indexes past its allocation.  The oracle refuses that case. */
    if (s->num_tot < 27) return ORC_NOT_SUPPORTED;
    s->neigh_num = 27;                               /* :116 */
    /* the reference's index formula (:244-246) is only a bijection for cubic
     * grids; the oracle restates it and therefore supports only those. */
    if (s->num_x != s->num_y || s->num_y != s->num_z) return ORC_NOT_SUPPORTED;
  }
  s->size_x = s->lx / s->num_x;                      /* :121-123 */
  s->size_y = s->ly / s->num_y;
  s->size_z = s->lz / s->num_z;
  if (s->size_x < 1.0 || s->size_y < 1.0 || s->size_z < 1.0)  /* :126-129 */
    return ORC_CELL_TOO_SMALL;
  return ORC_OK;
}

/* cell_part_idx (cell_list.c:237-249); same formula for a free point
 * (widom_cell_part_idx, compute_widom_chem_pot.c:122-128). */
static inline int cell_idx_xyz(const orc_sys *s, double x, double y, double z) {
  return (int)(x / s->size_x) * s->num_x * s->num_x +
         (int)(y / s->size_y) * s->num_y +
         (int)(z / s->size_z);
}
int orc_cell_part_idx(const orc_sys *s, int id) {
  const double *p = s->conf + 4 * (size_t)id;
  return cell_idx_xyz(s, p[1], p[2], p[3]);
}

/* cell_list_check (cell_list.c:224-233) */
static int cell_check(const orc_sys *s, int c) {
  int n = s->part_cell[(size_t)c * s->max_part];
  if (n < 0) return ORC_CELL_DEL;
  if (n > s->max_part - 1) return ORC_CELL_OVERFLOW;
  return ORC_OK;
}

/* cell_list_new (cell_list.c:142-175) */
int orc_cell_list_new(orc_sys *s) {
  int st = cell_info(s, 0);
  if (st) return st;
  const int mp = s->max_part;
  for (int c = 0; c < s->num_tot; c++) {             /* :148-153 */
    s->part_cell[(size_t)c * mp] = 0;
    for (int j = 1; j < mp; j++) s->part_cell[(size_t)c * mp + j] = -1;
  }
  for (int i = 0; i < s->N; i++) {                   /* :162-171 */
    int c = orc_cell_part_idx(s, i);
    if (c < 0 || c >= s->num_tot) return ORC_BAD_ARG;  /* UB in the reference */
    int *row = s->part_cell + (size_t)c * mp;
    row[0] += 1;
    st = cell_check(s, c);
    if (st) return st;
    row[row[0]] = i;   /* idx_col[c] == count after the increment */
  }
  return ORC_OK;
}

/* cell_neigh_init / neigh_id (cell_list.c:253-301) */
static void cell_neigh_init(orc_sys *s) {
  const int nx = s->num_x, ny = s->num_y, nz = s->num_z;
  for (int a = 0; a < nx; a++)
    for (int b = 0; b < ny; b++)
      for (int c = 0; c < nz; c++) {
        int ref = a * nx * nx + b * ny + c;
        int counter = 0;
        for (int i = -1; i < 2; i++)
          for (int j = -1; j < 2; j++)
            for (int k = -1; k < 2; k++) {
              int ix = a + i, iy = b + j, iz = c + k;
              if (ix > nx - 1) ix -= nx; else if (ix < 0) ix += nx;
              if (iy > ny - 1) iy -= ny; else if (iy < 0) iy += ny;
              if (iz > nz - 1) iz -= nz; else if (iz < 0) iz += nz;
              s->neigh_mat[(size_t)ref * 27 + counter] = ix * nx * nx + iy * ny + iz;
              counter++;
            }
      }
}

/* cell_list_init(alloc=true) (cell_list.c:33-58) */
int orc_cell_list_init(orc_sys *s) {
  int st = cell_info(s, 1);
  if (st) return st;
  free(s->part_cell);
  free(s->neigh_mat);
  s->part_cell = (int *)malloc(sizeof(int) * (size_t)s->num_tot * s->max_part);
  s->neigh_mat = (int *)malloc(sizeof(int) * (size_t)s->num_tot * 27);
  if (!s->part_cell || !s->neigh_mat) return ORC_NO_MEMORY;
  st = orc_cell_list_new(s);
  if (st) return st;
  cell_neigh_init(s);
  return ORC_OK;
}

/* cell_list_add (cell_list.c:185-192) */
int orc_cell_list_add(orc_sys *s, int c, int id) {
  int *row = s->part_cell + (size_t)c * s->max_part;
  int n = row[0];
  row[0] += 1;
  int st = cell_check(s, c);
  if (st) return st;
  row[n + 1] = id;
  return ORC_OK;
}

/* cell_list_del (cell_list.c:194-220): remove and shift left */
int orc_cell_list_del(orc_sys *s, int c, int id) {
  int *row = s->part_cell + (size_t)c * s->max_part;
  int n = row[0];
  int idx_remove = n;
  int shift = 0;
  row[0] -= 1;
  int st = cell_check(s, c);
  if (st) return st;
  for (int i = 1; i <= n; i++) {
    if (row[i] == id) { idx_remove = i; row[i] = -1; shift = 1; }
    if (shift && i > idx_remove) { row[i - 1] = row[i]; row[i] = -1; }
  }
  return shift ? ORC_OK : ORC_CELL_DEL;
}

/* cell_list_update (cell_list.c:180-183) */
int orc_cell_list_update(orc_sys *s, int cdel, int cadd, int id) {
  int st = orc_cell_list_del(s, cdel, id);
  if (st) return st;
  return orc_cell_list_add(s, cadd, id);
}

/* ===================================================================== */
/* System life cycle                                                      */
/* ===================================================================== */

orc_sys *orc_create(int type, int nx, int ny, int nz, double rho,
                    double neigh_dr, int max_part, double dr_max,
                    unsigned long seed, int *status) {
  int st = ORC_OK;
  orc_sys *s = (orc_sys *)calloc(1, sizeof(orc_sys));
  if (!s) { if (status) *status = ORC_NO_MEMORY; return NULL; }
  s->rho = rho;
  s->neigh_dr = neigh_dr;
  s->max_part = max_part;
  s->dr_max = dr_max;
  s->dv_max = 0.001;                       /* read_input.c:100 default */
  orc_box_init(s, type, nx, ny, nz, rho);
  s->Ncell = (s->cell_type == 1) ? 1 : 4;  /* part_alloc, sim_info.c:76-109 */
  s->N = s->cell_x * s->cell_y * s->cell_z * s->Ncell;
  s->conf = (double *)malloc(sizeof(double) * 4 * (size_t)s->N);
  if (!s->conf) { st = ORC_NO_MEMORY; goto fail; }
  st = orc_part_init(s);
  if (st) goto fail;
  orc_mt_seed(&s->rng, seed);
  st = orc_cell_list_init(s);
  if (st) goto fail;
  if (status) *status = ORC_OK;
  return s;
fail:
  orc_destroy(s);
  if (status) *status = st;
  return NULL;
}

void orc_destroy(orc_sys *s) {
  if (!s) return;
  free(s->conf);
  free(s->part_cell);
  free(s->neigh_mat);
  free(s);
}

void orc_set_params(orc_sys *s, double dr_max, double dv_max, double press) {
  s->dr_max = dr_max; s->dv_max = dv_max; s->press = press;
}
void orc_set_seed(orc_sys *s, unsigned long seed) { orc_mt_seed(&s->rng, seed); }

int orc_set_config(orc_sys *s, const double *conf4) {
  memcpy(s->conf, conf4, sizeof(double) * 4 * (size_t)s->N);
  return orc_cell_list_new(s);
}
void orc_get_config(const orc_sys *s, double *conf4) {
  memcpy(conf4, s->conf, sizeof(double) * 4 * (size_t)s->N);
}
void orc_get_cell_list(const orc_sys *s, int *part_cell, int *neigh_mat) {
  if (part_cell) memcpy(part_cell, s->part_cell, sizeof(int) * (size_t)s->num_tot * s->max_part);
  if (neigh_mat) memcpy(neigh_mat, s->neigh_mat, sizeof(int) * (size_t)s->num_tot * 27);
}
void orc_get_box(const orc_sys *s, double out[8]) {
  out[0] = s->vol; out[1] = s->lx; out[2] = s->ly; out[3] = s->lz;
  out[4] = s->size_x; out[5] = s->size_y; out[6] = s->size_z; out[7] = s->rho;
}
void orc_get_counters(const orc_sys *s, uint64_t out[6]) {
  out[0] = s->part_moves; out[1] = s->acc_part_moves; out[2] = s->rej_part_moves;
  out[3] = s->vol_moves; out[4] = s->acc_vol_moves; out[5] = s->rej_vol_moves;
}
void orc_reset_counters(orc_sys *s) {   /* moves.c:243-251 */
  s->part_moves = s->acc_part_moves = s->rej_part_moves = 0;
  s->vol_moves = s->acc_vol_moves = s->rej_vol_moves = 0;
}

/* ===================================================================== */
/* Moves (moves.c)                                                        */
/* ===================================================================== */

/* compute_dist (moves.c:400-431) */
double orc_compute_dist(const orc_sys *s, int i, int j, double sfx, double sfy, double sfz) {
  double lx = s->lx * sfx, ly = s->ly * sfy, lz = s->lz * sfz;
  double lx_2 = lx / 2.0, ly_2 = ly / 2.0, lz_2 = lz / 2.0;
  const double *a = s->conf + 4 * (size_t)i, *b = s->conf + 4 * (size_t)j;
  double dx = (a[1] - b[1]) * sfx;
  double dy = (a[2] - b[2]) * sfy;
  double dz = (a[3] - b[3]) * sfz;
  if (dx > lx_2) dx -= lx; else if (dx < -lx_2) dx += lx;
  if (dy > ly_2) dy -= ly; else if (dy < -ly_2) dy += ly;
  if (dz > lz_2) dz -= lz; else if (dz < -lz_2) dz += lz;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

/* check_overlap (moves.c:157-212) */
int orc_check_overlap(const orc_sys *s, int idx, double sfx, double sfy, double sfz) {
  int c = orc_cell_part_idx(s, idx);
  if (c < 0 || c >= s->num_tot) return -1;
  const int *nb = s->neigh_mat + (size_t)c * 27;
  for (int ii = 0; ii < 27; ii++) {
    const int *row = s->part_cell + (size_t)nb[ii] * s->max_part;
    int n = row[0];
    for (int jj = 1; jj <= n; jj++) {
      int p = row[jj];
      double dr = orc_compute_dist(s, idx, p, sfx, sfy, sfz);
      if (dr < 1.0 && p != idx) return 1;
    }
  }
  return 0;
}

/* apply_pbc (moves.c:215-226) */
void orc_apply_pbc(orc_sys *s, int idx) {
  double *p = s->conf + 4 * (size_t)idx;
  if (p[1] > s->lx) p[1] -= s->lx; else if (p[1] < 0.0) p[1] += s->lx;
  if (p[2] > s->ly) p[2] -= s->ly; else if (p[2] < 0.0) p[2] += s->ly;
  if (p[3] > s->lz) p[3] -= s->lz; else if (p[3] < 0.0) p[3] += s->lz;
}

int orc_overlap_at(orc_sys *s, int idx, double x, double y, double z) {
  double *p = s->conf + 4 * (size_t)idx;
  double ox = p[1], oy = p[2], oz = p[3];
  p[1] = x; p[2] = y; p[3] = z;
  int r = orc_check_overlap(s, idx, 1.0, 1.0, 1.0);
  p[1] = ox; p[2] = oy; p[3] = oz;
  return r;
}

/* part_move body (moves.c:44-78) with the four random numbers given.
 * Returns 1 accepted, 0 rejected, <0 error. */
int orc_part_move_scripted(orc_sys *s, int idx, double rx, double ry, double rz) {
  double *p = s->conf + 4 * (size_t)idx;
  double xo = p[1], yo = p[2], zo = p[3];               /* :46-48 */
  int cold = orc_cell_part_idx(s, idx);                 /* :49 */
  p[1] += (rx - 0.5) * s->dr_max;                       /* :52-54 */
  p[2] += (ry - 0.5) * s->dr_max;
  p[3] += (rz - 0.5) * s->dr_max;
  orc_apply_pbc(s, idx);                                /* :57 */
  int ov = orc_check_overlap(s, idx, 1.0, 1.0, 1.0);    /* :60 */
  s->part_moves += 1;                                   /* :78 */
  if (ov) {
    p[1] = xo; p[2] = yo; p[3] = zo;                    /* :62-65 */
    s->rej_part_moves += 1;
    return ov < 0 ? ov : 0;
  }
  int cnew = orc_cell_part_idx(s, idx);                 /* :69-72 */
  if (cnew != cold) {
    int st = orc_cell_list_update(s, cold, cnew, idx);
    if (st) return st;
  }
  s->acc_part_moves += 1;                               /* :74 */
  return 1;
}

/* part_move (moves.c:27-80) */
int orc_part_move(orc_sys *s) {
  int idx = orc_mt_int(&s->rng, s->N);                  /* :38 */
  double rx = orc_mt_double(&s->rng);                   /* :41-43 */
  double ry = orc_mt_double(&s->rng);
  double rz = orc_mt_double(&s->rng);
  return orc_part_move_scripted(s, idx, rx, ry, rz);
}

/* vol_move (moves.c:83-153).  Returns 1 accepted, 0 rejected, <0 error. */
int orc_vol_move(orc_sys *s) {
  double r_dv = orc_mt_double(&s->rng);                             /* :91 */
  double log_vol_new = log(s->vol) + (r_dv - 0.5) * s->dv_max;      /* :95 */
  double vol_new = exp(log_vol_new);                                /* :96 */
  double vol_ratio = vol_new / s->vol;                              /* :99 */
  double sf = pow(vol_ratio, 1. / 3.);                              /* :102 */
  int overlap = 0;
  for (int ii = 0; ii < s->N; ii++) {                               /* :106-112 */
    overlap = orc_check_overlap(s, ii, sf, sf, sf);
    if (overlap) break;
  }
  int accepted = 0;
  if (!overlap) {
    double boltz = exp(s->press * (s->vol - vol_new) +
                       (s->N + 1) * log(vol_ratio));                /* :117-118 */
    double r_acc = orc_mt_double(&s->rng);                          /* :119 */
    if (r_acc >= boltz) {
      s->rej_vol_moves += 1;
    } else {
      s->acc_vol_moves += 1;
      s->rho = s->N / vol_new;                                      /* :129 */
      orc_box_init(s, s->cell_type, s->cell_x, s->cell_y, s->cell_z, s->rho); /* :131 */
      for (int ii = 0; ii < s->N; ii++) {                           /* :135-140 */
        double *p = s->conf + 4 * (size_t)ii;
        p[1] *= sf; p[2] *= sf; p[3] *= sf;
        orc_apply_pbc(s, ii);
      }
      int st = orc_cell_list_new(s);                                /* :142 */
      if (st) return st;
      accepted = 1;
    }
  } else {
    s->rej_vol_moves += 1;
  }
  s->vol_moves += 1;                                                /* :151 */
  return accepted;
}

/* sweep_nvt (nvt.c:201-209) */
int orc_sweep_nvt(orc_sys *s, int n_sweeps) {
  for (int k = 0; k < n_sweeps; k++)
    for (int ii = 0; ii < s->N; ii++) {
      int r = orc_part_move(s);
      if (r < 0) return r;
    }
  return ORC_OK;
}

/* sweep_npt (npt.c:177-196) */
int orc_sweep_npt(orc_sys *s, int n_sweeps) {
  for (int k = 0; k < n_sweeps; k++)
    for (int ii = 0; ii < s->N; ii++) {
      int id = orc_mt_int(&s->rng, s->N + 1);                     /* :185 */
      int r = (id < s->N) ? orc_part_move(s) : orc_vol_move(s);   /* :187-192 */
      if (r < 0) return r;
    }
  return ORC_OK;
}

int orc_time_sweeps(orc_sys *s, int n_sweeps, double *wall_s, double *cpu_s) {
  struct timespec t0, t1;
  clock_t c0 = clock();
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int st = orc_sweep_nvt(s, n_sweeps);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  clock_t c1 = clock();
  if (wall_s) *wall_s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  if (cpu_s) *cpu_s = (double)(c1 - c0) / CLOCKS_PER_SEC;
  return st;
}

/* ===================================================================== */
/* Observables                                                            */
/* ===================================================================== */

/* rdf_hist_compute (compute_rdf.c:110-128): every pair i<j, bin
 * (int)((dr-1)/dr_bin), weight 2 in the reference; here an integer count of
 * pairs (the reference's double histogram is exactly 2 * count).  Bins the
 * reference would index outside [0, nbins) are counted in *out_of_range. */
int orc_rdf_counts(const orc_sys *s, double dr_bin, int nbins, double rmax_eff,
                   int64_t *counts, int64_t *out_of_range) {
  memset(counts, 0, sizeof(int64_t) * (size_t)nbins);
  int64_t oor = 0;
  for (int i = 0; i < s->N; i++)
    for (int j = i + 1; j < s->N; j++) {
      double dr = orc_compute_dist(s, i, j, 1.0, 1.0, 1.0);
      if (dr < rmax_eff) {
        int bin = (int)((dr - 1.0) / dr_bin);
        if (bin >= 0 && bin < nbins) counts[bin]++; else oor++;
      }
    }
  if (out_of_range) *out_of_range = oor;
  return ORC_OK;
}

/* pressv_compute_hist (compute_press.c:123-165): pairs j>i found through
 * the cell list, dr < rmax_eff, bin (int)((dr-1)/dr_bin). */
int orc_contact_counts(const orc_sys *s, double dr_bin, int nbins, double rmax_eff,
                       int64_t *counts, int64_t *out_of_range) {
  memset(counts, 0, sizeof(int64_t) * (size_t)nbins);
  int64_t oor = 0;
  for (int i = 0; i < s->N; i++) {
    int c = orc_cell_part_idx(s, i);
    const int *nb = s->neigh_mat + (size_t)c * 27;
    for (int jj = 0; jj < 27; jj++) {
      const int *row = s->part_cell + (size_t)nb[jj] * s->max_part;
      int n = row[0];
      for (int kk = 1; kk <= n; kk++) {
        int p = row[kk];
        double dr = orc_compute_dist(s, i, p, 1.0, 1.0, 1.0);
        if (dr < rmax_eff && p > i) {
          int bin = (int)((dr - 1.0) / dr_bin);
          if (bin >= 0 && bin < nbins) counts[bin]++; else oor++;
        }
      }
    }
  }
  if (out_of_range) *out_of_range = oor;
  return ORC_OK;
}

/* presst_compute_hist (compute_press.c:243-273) for the given scale
 * factors: no_overlap[k] = 1 when no particle overlaps after scaling. */
int orc_presst_flags(const orc_sys *s, const double *sf, int n, uint8_t *no_overlap) {
  for (int k = 0; k < n; k++) {
    int overlap = 0;
    for (int j = 0; j < s->N; j++) {
      overlap = orc_check_overlap(s, j, sf[k], sf[k], sf[k]);
      if (overlap) break;
    }
    no_overlap[k] = overlap ? 0 : 1;
  }
  return ORC_OK;
}

/* widom_check_overlap + widom_compute_dist (compute_widom_chem_pot.c:82-160)
 * for one trial point.  The point's cell index is clamped into the grid
 * (the reference reads outside its tables for a point exactly at L). */
static int widom_overlap(const orc_sys *s, double rx, double ry, double rz) {
  int ix = (int)(rx / s->size_x), iy = (int)(ry / s->size_y), iz = (int)(rz / s->size_z);
  if (ix >= s->num_x) ix = s->num_x - 1;
  if (iy >= s->num_y) iy = s->num_y - 1;
  if (iz >= s->num_z) iz = s->num_z - 1;
  int c = ix * s->num_x * s->num_x + iy * s->num_y + iz;   /* :124-126 */
  double lx = s->lx, ly = s->ly, lz = s->lz;
  double lx_2 = lx / 2.0, ly_2 = ly / 2.0, lz_2 = lz / 2.0;
  const int *nb = s->neigh_mat + (size_t)c * 27;
  for (int ii = 0; ii < 27; ii++) {
    const int *row = s->part_cell + (size_t)nb[ii] * s->max_part;
    int n = row[0];
    for (int jj = 1; jj <= n; jj++) {
      const double *b = s->conf + 4 * (size_t)row[jj];
      double dx = rx - b[1], dy = ry - b[2], dz = rz - b[3];   /* :143-145 */
      if (dx > lx_2) dx -= lx; else if (dx < -lx_2) dx += lx;
      if (dy > ly_2) dy -= ly; else if (dy < -ly_2) dy += ly;
      if (dz > lz_2) dz -= lz; else if (dz < -lz_2) dz += lz;
      double dr = sqrt(dx * dx + dy * dy + dz * dz);
      if (dr < 1.0) return 1;                                  /* :109 */
    }
  }
  return 0;
}

int64_t orc_widom_points(const orc_sys *s, const double *xyz, int m, uint8_t *hit) {
  int64_t hits = 0;
  for (int i = 0; i < m; i++) {
    int h = !widom_overlap(s, xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    if (hit) hit[i] = (uint8_t)h;
    hits += h;
  }
  return hits;
}

/* widom_insertion (compute_widom_chem_pot.c:44-71): n trials on the MT stream */
int64_t orc_widom_insertion(orc_sys *s, int n) {
  int64_t hits = 0;
  for (int i = 0; i < n; i++) {
    double rx = orc_mt_double(&s->rng) * s->lx;   /* :76-78 */
    double ry = orc_mt_double(&s->rng) * s->ly;
    double rz = orc_mt_double(&s->rng) * s->lz;
    if (!widom_overlap(s, rx, ry, rz)) hits++;
  }
  return hits;
}

/* Trial point i of sample `sample` on the GPU's Philox stream:
 * ctr = {i_lo, i_hi, sample_lo, 0x57490000 | (sample_hi & 0xffff)},
 * key = seed; coordinates u/(2^32-1) * L as widom_rand_pos does. */
void orc_widom_philox_point(const orc_sys *s, uint64_t seed, uint64_t sample,
                            int64_t i, double out[3]) {
  uint32_t ctr[4] = {(uint32_t)(uint64_t)i, (uint32_t)((uint64_t)i >> 32),
                     (uint32_t)sample, 0x57490000u | (uint32_t)((sample >> 32) & 0xffffu)};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t u[4];
  orc_philox4x32_10(ctr, key, u);
  out[0] = ((double)u[0] / (double)4294967295.0) * s->lx;
  out[1] = ((double)u[1] / (double)4294967295.0) * s->ly;
  out[2] = ((double)u[2] / (double)4294967295.0) * s->lz;
}

int64_t orc_widom_philox(const orc_sys *s, uint64_t seed, uint64_t sample,
                         int64_t first, int64_t n) {
  int64_t hits = 0;
  double p[3];
  for (int64_t i = first; i < first + n; i++) {
    orc_widom_philox_point(s, seed, sample, i, p);
    if (!widom_overlap(s, p[0], p[1], p[2])) hits++;
  }
  return hits;
}

/* Cell-stratified uniform insertion (no reference counterpart; restates
 * hsmc_gpu_widom_cells): per_cell points uniform in each cell of an
 * nx*ny*nz grid of the box, point j of cell c = (ix*ny + iy)*nz + iz from
 * Philox counter {c, j, sample_lo, 0x5743xxxx}, coordinate
 * (i + (u + 0.5) 2^-32) * (L/n) per axis; each tested by the reference's
 * Widom overlap test (compute_widom_chem_pot.c:82-160). */
int64_t orc_widom_cells(const orc_sys *s, uint64_t seed, uint64_t sample, int nx, int ny,
                        int nz, int per_cell) {
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const double sx = s->lx / nx, sy = s->ly / ny, sz = s->lz / nz;
  int64_t hits = 0;
  for (int ix = 0; ix < nx; ix++)
    for (int iy = 0; iy < ny; iy++)
      for (int iz = 0; iz < nz; iz++) {
        const uint32_t c = (uint32_t)((ix * ny + iy) * nz + iz);
        for (int j = 0; j < per_cell; j++) {
          const uint32_t ctr[4] = {c, (uint32_t)j, (uint32_t)sample,
                                   0x57430000u | (uint32_t)((sample >> 32) & 0xFFFFu)};
          uint32_t u[4];
          orc_philox4x32_10(ctr, key, u);
          const double x = ((double)ix + ((double)u[0] + 0.5) * 0x1p-32) * sx;
          const double y = ((double)iy + ((double)u[1] + 0.5) * 0x1p-32) * sy;
          const double z = ((double)iz + ((double)u[2] + 0.5) * 0x1p-32) * sz;
          if (!widom_overlap(s, x, y, z)) hits++;
        }
      }
  return hits;
}

/* ===================================================================== */
/* Checkerboard sweep: CPU restatement of the HIP kernel's algorithm.     */
/* (No reference counterpart; the reference sweep is serial, nvt.c:201.)  */
/* ===================================================================== */

orc_cb *orc_cb_create(int N, const double *conf4, double lx, double ly, double lz,
                      int gx, int gy, int gz, double dr_max, uint64_t seed) {
  orc_cb *c = (orc_cb *)calloc(1, sizeof(orc_cb));
  if (!c) return NULL;
  c->N = N; c->lx = lx; c->ly = ly; c->lz = lz;
  c->gx = gx; c->gy = gy; c->gz = gz; c->dr_max = dr_max;
  c->key[0] = (uint32_t)seed; c->key[1] = (uint32_t)(seed >> 32);
  c->x = (double *)malloc(sizeof(double) * (size_t)N);
  c->y = (double *)malloc(sizeof(double) * (size_t)N);
  c->z = (double *)malloc(sizeof(double) * (size_t)N);
  if (!c->x || !c->y || !c->z) { orc_cb_destroy(c); return NULL; }
  for (int i = 0; i < N; i++) {
    c->x[i] = conf4[4 * i + 1]; c->y[i] = conf4[4 * i + 2]; c->z[i] = conf4[4 * i + 3];
  }
  return c;
}

void orc_cb_destroy(orc_cb *c) {
  if (!c) return;
  free(c->x); free(c->y); free(c->z); free(c);
}

void orc_cb_get_config(const orc_cb *c, double *conf4) {
  for (int i = 0; i < c->N; i++) {
    conf4[4 * i] = i; conf4[4 * i + 1] = c->x[i]; conf4[4 * i + 2] = c->y[i]; conf4[4 * i + 3] = c->z[i];
  }
}
void orc_cb_get_counters(const orc_cb *c, uint64_t out[4]) {
  out[0] = c->attempts; out[1] = c->accepted; out[2] = c->left_cell; out[3] = c->sweep;
}

/* Shift of the cell grid (in [0, side) per axis) and order of the 8
 * sublattices for one sweep. */
void orc_cb_sweep_randoms(const orc_cb *c, uint64_t sweep, double shift[3], int perm[8]) {
  uint32_t ctr[4] = {0xFFFFFFFFu, 0u, (uint32_t)sweep, (uint32_t)(sweep >> 32)};
  uint32_t u[4];
  orc_philox4x32_10(ctr, c->key, u);
  double sx = c->lx / c->gx, sy = c->ly / c->gy, sz = c->lz / c->gz;
  shift[0] = ((double)u[0] * 0x1p-32) * sx;
  shift[1] = ((double)u[1] * 0x1p-32) * sy;
  shift[2] = ((double)u[2] * 0x1p-32) * sz;
  ctr[0] = 0xFFFFFFFEu;
  orc_philox4x32_10(ctr, c->key, u);
  for (int i = 0; i < 8; i++) perm[i] = i;
  for (int i = 7, h = 0; i >= 1; i--, h++) {
    uint32_t half = (u[h >> 1] >> (16 * (h & 1))) & 0xFFFFu;
    int j = (int)((half * (uint32_t)(i + 1)) >> 16);
    int t = perm[i]; perm[i] = perm[j]; perm[j] = t;
  }
}

/* Occupant pick of attempt t of cell cid in sweep `sweep` among n occupants:
 * rng_get_int's rule (rng.c:34-36, GSL uniform_int: scale = (2^32-1)/n,
 * k = draw/scale, redraw while k >= n) on the Philox stream, the first draw
 * from counter (cid, t, sweep) and redraw r from (cid, t | r << 16, sweep). */
uint32_t orc_cb_pick(const uint32_t key[2], uint32_t cid, uint32_t t, uint64_t sweep, uint32_t n) {
  const uint32_t scale = 0xFFFFFFFFu / n;
  uint32_t ctr[4] = {cid, t, (uint32_t)sweep, (uint32_t)(sweep >> 32)};
  uint32_t w[4];
  orc_philox4x32_10(ctr, key, w);
  uint32_t k = w[0] / scale;
  for (uint32_t r = 1; k >= n; r++) {
    ctr[1] = t | (r << 16);
    orc_philox4x32_10(ctr, key, w);
    k = w[0] / scale;
  }
  return k;
}

/* cell index on the shifted grid: (int)((x - shift, wrapped) * (1/side)) */
static inline int cb_axis_cell(double x, double shift, double L, double inv_side, int g) {
  double t = x - shift;
  if (t < 0.0) t += L;
  int ci = (int)(t * inv_side);
  if (ci >= g) ci = g - 1;
  return ci;
}

static inline double cb_dist2(double ax, double ay, double az, double bx, double by, double bz,
                              double lx, double ly, double lz, double hx, double hy, double hz) {
  double dx = ax - bx, dy = ay - by, dz = az - bz;
  if (dx > hx) dx -= lx; else if (dx < -hx) dx += lx;
  if (dy > hy) dy -= ly; else if (dy < -hy) dy += ly;
  if (dz > hz) dz -= lz; else if (dz < -hz) dz += lz;
  return dx * dx + dy * dy + dz * dz;
}

int orc_cb_sweeps(orc_cb *c, int n_sweeps) {
  const int gx = c->gx, gy = c->gy, gz = c->gz, C = gx * gy * gz, N = c->N;
  const double lx = c->lx, ly = c->ly, lz = c->lz;
  const double hx = lx / 2.0, hy = ly / 2.0, hz = lz / 2.0;
  const double sx = lx / gx, sy = ly / gy, sz = lz / gz;
  const double ix = 1.0 / sx, iy = 1.0 / sy, iz = 1.0 / sz;
  int *cnt = (int *)malloc(sizeof(int) * (size_t)(C + 1));
  int *start = (int *)malloc(sizeof(int) * (size_t)(C + 1));
  int *order = (int *)malloc(sizeof(int) * (size_t)N);
  int *cell = (int *)malloc(sizeof(int) * (size_t)N);
  if (!cnt || !start || !order || !cell) { free(cnt); free(start); free(order); free(cell); return ORC_NO_MEMORY; }
  for (int sw = 0; sw < n_sweeps; sw++) {
    const uint64_t s = c->sweep;
    double shift[3]; int perm[8];
    orc_cb_sweep_randoms(c, s, shift, perm);
    /* bin: cells hold their occupants in ascending particle id */
    memset(cnt, 0, sizeof(int) * (size_t)(C + 1));
    for (int i = 0; i < N; i++) {
      int cx = cb_axis_cell(c->x[i], shift[0], lx, ix, gx);
      int cy = cb_axis_cell(c->y[i], shift[1], ly, iy, gy);
      int cz = cb_axis_cell(c->z[i], shift[2], lz, iz, gz);
      cell[i] = (cx * gy + cy) * gz + cz;
      cnt[cell[i]]++;
    }
    start[0] = 0;
    for (int k = 0; k < C; k++) start[k + 1] = start[k] + cnt[k];
    memset(cnt, 0, sizeof(int) * (size_t)(C + 1));
    for (int i = 0; i < N; i++) order[start[cell[i]] + cnt[cell[i]]++] = i;
    for (int ph = 0; ph < 8; ph++) {
      const int col = perm[ph];
      const int bx = col & 1, by = (col >> 1) & 1, bz = (col >> 2) & 1;
      for (int cx = bx; cx < gx; cx += 2)
        for (int cy = by; cy < gy; cy += 2)
          for (int cz = bz; cz < gz; cz += 2) {
            const int cid = (cx * gy + cy) * gz + cz;
            const int n = start[cid + 1] - start[cid];
            for (int t = 0; t < n; t++) {
              uint32_t ctr[4] = {(uint32_t)cid, (uint32_t)t, (uint32_t)s, (uint32_t)(s >> 32)};
              uint32_t w[4];
              orc_philox4x32_10(ctr, c->key, w);
              const uint32_t k = orc_cb_pick(c->key, (uint32_t)cid, (uint32_t)t, s, (uint32_t)n);
              const int pid = order[start[cid] + (int)k];
              double nx = c->x[pid] + ((double)w[1] / (double)4294967295.0 - 0.5) * c->dr_max;
              double ny = c->y[pid] + ((double)w[2] / (double)4294967295.0 - 0.5) * c->dr_max;
              double nz = c->z[pid] + ((double)w[3] / (double)4294967295.0 - 0.5) * c->dr_max;
              if (nx > lx) nx -= lx; else if (nx < 0.0) nx += lx;   /* apply_pbc */
              if (ny > ly) ny -= ly; else if (ny < 0.0) ny += ly;
              if (nz > lz) nz -= lz; else if (nz < 0.0) nz += lz;
              c->attempts++;
              if (cb_axis_cell(nx, shift[0], lx, ix, gx) != cx ||
                  cb_axis_cell(ny, shift[1], ly, iy, gy) != cy ||
                  cb_axis_cell(nz, shift[2], lz, iz, gz) != cz) {
                c->left_cell++;
                continue;
              }
              int overlap = 0;
              for (int d = 0; d < 27 && !overlap; d++) {
                const int ddx = d / 9 - 1, ddy = (d / 3) % 3 - 1, ddz = d % 3 - 1;
                if ((gx == 2 && ddx == -1) || (gy == 2 && ddy == -1) || (gz == 2 && ddz == -1)) continue;
                int ax = cx + ddx, ay = cy + ddy, az = cz + ddz;
                if (ax < 0) ax += gx; else if (ax >= gx) ax -= gx;
                if (ay < 0) ay += gy; else if (ay >= gy) ay -= gy;
                if (az < 0) az += gz; else if (az >= gz) az -= gz;
                const int nc = (ax * gy + ay) * gz + az;
                for (int q = start[nc]; q < start[nc + 1]; q++) {
                  const int o = order[q];
                  if (o == pid) continue;
                  if (cb_dist2(nx, ny, nz, c->x[o], c->y[o], c->z[o], lx, ly, lz, hx, hy, hz) < 1.0) {
                    overlap = 1; break;
                  }
                }
              }
              if (!overlap) {
                c->x[pid] = nx; c->y[pid] = ny; c->z[pid] = nz;
                c->accepted++;
              }
            }
          }
    }
    c->sweep++;
  }
  free(cnt); free(start); free(order); free(cell);
  return ORC_OK;
}
