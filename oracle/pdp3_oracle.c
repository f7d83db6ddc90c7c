/*
 * pdp3_oracle.c -- CPU oracle for the PDP3 r-z particle-in-cell step loop.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and CPU baseline for the HIP
 * product library, never part of it.  Clean-room FP64 restatement of the
 * reference (cosmonaut-ok/pdp3); every function cites the reference file:line
 * it follows.  Arithmetic is written so that each floating-point expression
 * is evaluated in the same order as the reference's (including the order its
 * unparenthesised macros expand to), and the file is compiled without FMA
 * contraction, so that results agree with the reference's IEEE build bit for
 * bit up to libm differences.
 *
 * Quirks of the reference are reproduced on purpose (non-normalised gather
 * weights, Vay scaling, the strict-radial precedence slip, non-depositing
 * equality cases).  They are part of the behaviour being checked.
 */
#include "pdp3_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---- constants: include/constant.h:10-27 ---- */
#define PI 3.1415926535897932
#define EPSILON0 8.85E-12
#define EL_MASS 9.1E-31
#define EL_CHARGE 1.6E-19
#define LIGHT_SPEED 3.0E8
#define MAGN_CONST 1.26E-6
#define MNZL 1e-15

#define MAX_SPECIES 64

typedef struct {
  int n;
  int is_bunch;
  double charge; /* per real particle, C   (src/particles.cpp:30) */
  double mass;   /* per real particle, kg  (src/particles.cpp:32) */
  double *r, *y, *z, *vr, *vphi, *vz, *q, *m;
  double *r_old, *z_old, *sin_t;
  unsigned char *alive;
  /* bunch parameters (include/bunch.h:22-28) */
  double duration, hole_duration, velocity;
  int bunch_number;
  int created;
  double *inject_u;
} species_t;

struct or_sim {
  int nr, nz;
  double r_size, z_size, dr, dz;
  double current_time, start_time, relaxation_time, end_time, dt;
  int pusher;
  int threads;
  double *a[OR_NCOMP];
  int nsp;
  species_t sp[MAX_SPECIES];
  int order[MAX_SPECIES]; /* species list order (ParticlesList::part_list) */
  int norder;
  int next_bunch; /* init_beam's current_bunch_number */
  int reverse_deposit; /* test hook: serial deposits in reverse particle order */
};

#define AT(s, arr, i, k) ((s)->a[arr][(size_t)(i) * (size_t)(s)->nz + (size_t)(k)])

/* ---- geometry macros: include/geometry.h:33-44 ---- */
int or_cell_number(double x, double dx) { return (int)ceil(x / dx) - 1; }
static double cell_volume(int i, double dr, double dz) {
  return PI * dz * dr * dr * 2.0 * i;
}
static double cyl_vol(double z, double r) { return PI * z * r * r / 4.; }

/* ---- lib::get_gamma / get_gamma_inv: src/lib.cpp:45-80 ---- */
static const double C2 = LIGHT_SPEED * LIGHT_SPEED; /* pow(3e8,2) exact */
static int get_gamma(double sq_velocity, double *gamma) {
  double beta = sq_velocity / C2;
  if (beta > 1) return OR_ERR_SUPERLUMINAL;
  double g = 1 / sqrt(1.0 - beta);
  if (isinf(g)) g = 1e100;
  *gamma = g;
  return OR_OK;
}
/* Test hook (off by default): or_set_gamma_inv_nearest(1) replaces glibc's
 * pow(y, -0.5) by the double nearest to 1/sqrt(y), picked among 1/sqrt(y)
 * and its two neighbours on the residual 1 - y c^2 evaluated with
 * error-free products.  glibc's pow returns that double in all but a few
 * hundredths of a percent of arguments; the switch measures how far the
 * loop moves when only those last-bit cases change. */
static int g_gamma_inv_nearest = 0;
void or_set_gamma_inv_nearest(int on) { g_gamma_inv_nearest = on != 0; }
static double rsqrt_residual(double y, double c) {
  double p = c * c, pe = fma(c, c, -p);
  double q = y * p, qe = fma(y, p, -q);
  return ((1.0 - q) - qe) - y * pe;
}
static double rsqrt_nearest(double y) {
  double c = 1.0 / sqrt(y), best = c, rb = fabs(rsqrt_residual(y, c));
  double lo = nextafter(c, 0.0), hi = nextafter(c, 1e308);
  double rl = fabs(rsqrt_residual(y, lo)), rh = fabs(rsqrt_residual(y, hi));
  if (rl < rb) {
    best = lo;
    rb = rl;
  }
  if (rh < rb) best = hi;
  return best;
}
static double get_gamma_inv(double sq_velocity) {
  double y = 1.0 + sq_velocity / C2;
  return g_gamma_inv_nearest ? rsqrt_nearest(y) : pow(y, -0.5);
}
/* tinyvec3d::tv_dot (lib/tinyvec3d/tinyvec3d.cpp:54-61): ((0+a0b0)+a1b1)+a2b2 */
static double dot3(const double *a, const double *b) {
  double res = 0;
  res += a[0] * b[0];
  res += a[1] * b[1];
  res += a[2] * b[2];
  return res;
}

/* ------------------------------------------------------------------------ */
/* simulation object                                                        */
/* ------------------------------------------------------------------------ */

or_sim *or_sim_new(int n_grid_r, int n_grid_z, double r_size, double z_size) {
  if (n_grid_r < 4 || n_grid_z < 4) return NULL;
  or_sim *s = (or_sim *)calloc(1, sizeof(or_sim));
  s->nr = n_grid_r;
  s->nz = n_grid_z;
  s->r_size = r_size;
  s->z_size = z_size;
  /* Geometry::get_dr/get_dz, src/geometry.cpp:45-53 */
  s->dr = r_size / (n_grid_r - 1);
  s->dz = z_size / (n_grid_z - 1);
  s->pusher = OR_PUSHER_BORIS;
  s->threads = 1;
  size_t n = (size_t)n_grid_r * (size_t)n_grid_z;
  for (int c = 0; c < OR_NCOMP; c++) s->a[c] = (double *)calloc(n, sizeof(double));
  /* Geometry::set_epsilon, src/geometry.cpp:33-39 */
  for (size_t j = 0; j < n; j++) s->a[OR_EPSILON][j] = 1;
  return s;
}

static void free_species(species_t *p) {
  free(p->r); free(p->y); free(p->z); free(p->vr); free(p->vphi); free(p->vz);
  free(p->q); free(p->m); free(p->r_old); free(p->z_old); free(p->sin_t);
  free(p->alive); free(p->inject_u);
}

void or_sim_free(or_sim *s) {
  if (!s) return;
  for (int c = 0; c < OR_NCOMP; c++) free(s->a[c]);
  for (int j = 0; j < s->nsp; j++) free_species(&s->sp[j]);
  free(s);
}

void or_sim_set_time(or_sim *s, double current_time, double start_time,
                     double relaxation_time, double end_time, double delta_t) {
  /* Time::Time, src/pdp3Time.cpp:11-18 */
  s->current_time = current_time;
  s->start_time = start_time;
  s->relaxation_time = relaxation_time;
  s->end_time = end_time;
  s->dt = delta_t;
}

void or_sim_set_pusher(or_sim *s, int pusher) { s->pusher = pusher; }
void or_sim_set_threads(or_sim *s, int threads) { s->threads = threads < 1 ? 1 : threads; }
void or_sim_set_reverse_deposit(or_sim *s, int reverse) { s->reverse_deposit = reverse != 0; }
double or_sim_time(const or_sim *s) { return s->current_time; }
double or_sim_dr(const or_sim *s) { return s->dr; }
double or_sim_dz(const or_sim *s) { return s->dz; }
double *or_sim_field(or_sim *s, int comp) {
  return (comp >= 0 && comp < OR_NCOMP) ? s->a[comp] : NULL;
}

/* Geometry::set_pml, src/geometry.cpp:56-107 */
void or_sim_set_pml(or_sim *s, double l1, double l2, double l3, double sigma1,
                    double sigma2) {
  const double dr = s->dr, dz = s->dz;
  const int nr = s->nr, nz = s->nz;
  double len_left = dz * (floor(nz * l1));
  double len_right = dz * (floor(nz * l2));
  double len_ext = dr * (floor(nr * l3));
  if ((l1 == 0) && (l2 == 0) && (l3 != 0)) {
    for (int i = 0; i < nr; i++)
      for (int k = 0; k < nz; k++)
        if (s->r_size - dr * i <= len_ext)
          AT(s, OR_SIGMA, i, k) = sigma1 + (sigma2 - sigma1) / pow(len_ext, 2) *
                                               pow((dr * (i + 1) - s->r_size + len_ext), 2);
  } else {
    for (int k = 0; k < nz; k++)
      for (int i = 0; i < nr; i++) {
        if (dz * k < len_left)
          AT(s, OR_SIGMA, i, k) = sigma1 + (sigma2 - sigma1) / pow(len_left, 2) *
                                               pow(len_left - dz * k, 2);
        if (s->z_size - dz * k < len_right)
          AT(s, OR_SIGMA, i, k) = sigma1 + (sigma2 - sigma1) / pow(len_right, 2) *
                                               pow(dz * (k + 1) - s->z_size + len_right, 2);
      }
  }
  if (l3 != 0) {
    for (int i = 0; i < nr; i++)
      for (int k = 0; k < nz; k++)
        if (s->r_size - dr * i <= len_ext)
          if ((s->r_size - dr * (i + 1) < len_ext * dz * k / len_left) &&
              (s->r_size - dr * i <= len_ext / len_right * (s->z_size - dz * k)))
            AT(s, OR_SIGMA, i, k) = sigma1 + (sigma2 - sigma1) / pow(len_ext, 2) *
                                                 pow(dr * (i + 1) - s->r_size + len_ext, 2);
  }
}

static double *dup_or_fill(const double *src, int n, double fill) {
  double *d = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; j++) d[j] = src ? src[j] : fill;
  return d;
}

static species_t *new_species(or_sim *s, int n, int charge_e, double mass_e) {
  if (s->nsp >= MAX_SPECIES) return NULL;
  species_t *p = &s->sp[s->nsp];
  memset(p, 0, sizeof(*p));
  p->n = n;
  /* Particles::Particles, src/particles.cpp:23-55 */
  p->charge = charge_e * EL_CHARGE;
  p->mass = mass_e * EL_MASS;
  p->r_old = dup_or_fill(NULL, n, 0);
  p->z_old = dup_or_fill(NULL, n, 0);
  p->sin_t = dup_or_fill(NULL, n, 0);
  p->y = dup_or_fill(NULL, n, 0);
  p->alive = (unsigned char *)calloc((size_t)(n > 0 ? n : 1), 1);
  return p;
}

int or_sim_add_plasma(or_sim *s, int charge_e, double mass_e, int n,
                      const double *r, const double *z, const double *vr,
                      const double *vphi, const double *vz, const double *w,
                      const unsigned char *alive) {
  species_t *p = new_species(s, n, charge_e, mass_e);
  if (!p) return OR_ERR_BAD_ARG;
  p->is_bunch = 0;
  p->r = dup_or_fill(r, n, 0);
  p->z = dup_or_fill(z, n, 0);
  p->vr = dup_or_fill(vr, n, 0);
  p->vphi = dup_or_fill(vphi, n, 0);
  p->vz = dup_or_fill(vz, n, 0);
  p->q = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  p->m = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; j++) {
    /* src/particles.cpp:328-329 */
    p->q[j] = p->charge * w[j];
    p->m[j] = p->mass * w[j];
    p->alive[j] = alive ? alive[j] : 1;
  }
  p->created = 1;
  s->order[s->norder++] = s->nsp;
  return s->nsp++;
}

int or_sim_add_bunch(or_sim *s, int charge_e, double mass_e, int n,
                     const double *r, const double *w, const double *inject_u,
                     double duration, double hole_duration, double velocity,
                     int bunch_number) {
  species_t *p = new_species(s, n, charge_e, mass_e);
  if (!p) return OR_ERR_BAD_ARG;
  p->is_bunch = 1;
  /* Bunch::Bunch, src/bunch.cpp:6-80: z=0, v=0, not alive */
  p->r = dup_or_fill(r, n, 0);
  p->z = dup_or_fill(NULL, n, 0);
  p->vr = dup_or_fill(NULL, n, 0);
  p->vphi = dup_or_fill(NULL, n, 0);
  p->vz = dup_or_fill(NULL, n, 0);
  p->q = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  p->m = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; j++) {
    p->q[j] = p->charge * w[j]; /* src/bunch.cpp:74-75 */
    p->m[j] = p->mass * w[j];
  }
  p->inject_u = dup_or_fill(inject_u, n, 0);
  p->duration = duration;
  p->hole_duration = hole_duration;
  p->velocity = velocity;
  p->bunch_number = bunch_number;
  p->created = 0;
  return s->nsp++;
}

/* a bunch already inside the grid: created and alive at once, nothing left
 * to inject (duration 0); absorbed at the walls and charge-weighted like any
 * bunch.  No reference counterpart (its bunches enter by injection only):
 * the adjoint checks differentiate a bunch given as a state. */
int or_sim_add_bunch_loaded(or_sim *s, int charge_e, double mass_e, int n,
                            const double *r, const double *z, const double *vr,
                            const double *vphi, const double *vz, const double *w) {
  species_t *p = new_species(s, n, charge_e, mass_e);
  if (!p) return OR_ERR_BAD_ARG;
  p->is_bunch = 1;
  p->r = dup_or_fill(r, n, 0);
  p->z = dup_or_fill(z, n, 0);
  p->vr = dup_or_fill(vr, n, 0);
  p->vphi = dup_or_fill(vphi, n, 0);
  p->vz = dup_or_fill(vz, n, 0);
  p->q = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  p->m = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; j++) {
    p->q[j] = p->charge * w[j];
    p->m[j] = p->mass * w[j];
    p->alive[j] = 1;
  }
  p->inject_u = dup_or_fill(NULL, n, 0);
  p->duration = 0;
  p->created = 1;
  s->order[s->norder++] = s->nsp;
  return s->nsp++;
}

int or_sim_num_species(const or_sim *s) { return s->nsp; }
int or_sim_species_count(const or_sim *s, int idx) {
  return (idx >= 0 && idx < s->nsp) ? s->sp[idx].n : -1;
}

void or_sim_get_species(const or_sim *s, int idx, double *r, double *z,
                        double *vr, double *vphi, double *vz, double *q,
                        double *m, unsigned char *alive) {
  const species_t *p = &s->sp[idx];
  size_t b = sizeof(double) * (size_t)p->n;
  if (r) memcpy(r, p->r, b);
  if (z) memcpy(z, p->z, b);
  if (vr) memcpy(vr, p->vr, b);
  if (vphi) memcpy(vphi, p->vphi, b);
  if (vz) memcpy(vz, p->vz, b);
  if (q) memcpy(q, p->q, b);
  if (m) memcpy(m, p->m, b);
  if (alive) memcpy(alive, p->alive, (size_t)p->n);
}

void or_sim_set_species(or_sim *s, int idx, const double *r, const double *z,
                        const double *vr, const double *vphi, const double *vz,
                        const unsigned char *alive) {
  species_t *p = &s->sp[idx];
  size_t b = sizeof(double) * (size_t)p->n;
  if (r) memcpy(p->r, r, b);
  if (z) memcpy(p->z, z, b);
  if (vr) memcpy(p->vr, vr, b);
  if (vphi) memcpy(p->vphi, vphi, b);
  if (vz) memcpy(p->vz, vz, b);
  if (alive) memcpy(p->alive, alive, (size_t)p->n);
}

/* ------------------------------------------------------------------------ */
/* fields                                                                   */
/* ------------------------------------------------------------------------ */

/* EField::boundary_conditions, src/eField.cpp:46-82 */
void or_sim_e_boundary_conditions(or_sim *s) {
  const int nr = s->nr, nz = s->nz;
  for (int i = 0; i < nr - 1; i++) {
    AT(s, OR_E_R, i, 0) = 0;
    AT(s, OR_E_R, i, nz - 1) = 0;
  }
  for (int k = 0; k < nz; k++) {
    AT(s, OR_E_PHI, 0, k) = 0;
    AT(s, OR_E_PHI, nr - 1, k) = 0;
  }
  for (int i = 0; i < nr; i++) {
    AT(s, OR_E_PHI, i, 0) = 0;
    AT(s, OR_E_PHI, i, nz - 1) = 0;
  }
  for (int k = 0; k < nz - 1; k++) {
    AT(s, OR_E_Z, 0, k) = 0;
    AT(s, OR_E_Z, nr - 1, k) = 0;
  }
}

/* HField::calc_field, src/hField.cpp:93-157 */
void or_sim_h_calc(or_sim *s) {
  const int nr = s->nr, nz = s->nz;
  const double dr = s->dr, dz = s->dz, dt = s->dt;
  /* H_r on the last radial row (:105-117) and on all other rows (:119-128) */
  for (int i = 0; i < nr; i++)
    for (int k = 0; k < nz - 1; k++) {
      double alpha_t = dt * (AT(s, OR_E_PHI, i, k + 1) - AT(s, OR_E_PHI, i, k)) /
                       (dz * MAGN_CONST);
      AT(s, OR_HH_R, i, k) = AT(s, OR_H_R, i, k) + alpha_t / 2;
      AT(s, OR_H_R, i, k) = AT(s, OR_H_R, i, k) + alpha_t;
    }
  /* H_phi (:130-141) */
  for (int i = 0; i < nr - 1; i++)
    for (int k = 0; k < nz - 1; k++) {
      double alpha_t = dt *
                       ((AT(s, OR_E_Z, i + 1, k) - AT(s, OR_E_Z, i, k)) / dr -
                        (AT(s, OR_E_R, i, k + 1) - AT(s, OR_E_R, i, k)) / dz) /
                       MAGN_CONST;
      AT(s, OR_HH_PHI, i, k) = AT(s, OR_H_PHI, i, k) + alpha_t / 2;
      AT(s, OR_H_PHI, i, k) = AT(s, OR_H_PHI, i, k) + alpha_t;
    }
  /* H_z (:143-155) */
  for (int i = 0; i < nr - 1; i++)
    for (int k = 0; k < nz - 1; k++) {
      double alpha_t = dt *
                       ((AT(s, OR_E_PHI, i + 1, k) + AT(s, OR_E_PHI, i, k)) /
                            (2.0 * dr * (i + 0.5)) +
                        (AT(s, OR_E_PHI, i + 1, k) - AT(s, OR_E_PHI, i, k)) / dr) /
                       MAGN_CONST;
      AT(s, OR_HH_Z, i, k) = AT(s, OR_H_Z, i, k) - alpha_t / 2;
      AT(s, OR_H_Z, i, k) = AT(s, OR_H_Z, i, k) - alpha_t;
    }
}

/* conductivity coefficients, src/eField.cpp:105-110 */
static void koef(const or_sim *s, int i, int k, double *ke, double *kh) {
  double epsilonx2 = 2 * s->a[OR_EPSILON][(size_t)i * s->nz + k] * EPSILON0;
  double sigma_t = s->a[OR_SIGMA][(size_t)i * s->nz + k] * s->dt;
  *ke = (epsilonx2 - sigma_t) / (epsilonx2 + sigma_t);
  *kh = 2 * s->dt / (epsilonx2 + sigma_t);
}

/* EField::calc_field, src/eField.cpp:85-170 */
void or_sim_e_calc(or_sim *s) {
  const int nr = s->nr, nz = s->nz;
  const double dr = s->dr, dz = s->dz;
  double ke, kh;
  /* E_r on the axis row (:99-112) */
  for (int k = 1; k < nz - 1; k++) {
    int i = 0;
    koef(s, i, k, &ke, &kh);
    AT(s, OR_E_R, i, k) = AT(s, OR_E_R, i, k) * ke -
                          (AT(s, OR_J_R, i, k) +
                           (AT(s, OR_H_PHI, i, k) - AT(s, OR_H_PHI, i, k - 1)) / dz) * kh;
  }
  /* E_z on the axis (:114-129) */
  for (int k = 0; k < nz - 1; k++) {
    int i = 0;
    koef(s, i, k, &ke, &kh);
    AT(s, OR_E_Z, i, k) = AT(s, OR_E_Z, i, k) * ke -
                          (AT(s, OR_J_Z, i, k) - 4.0 / dr * AT(s, OR_H_PHI, i, k)) * kh;
  }
  /* interior (:131-152) */
  for (int i = 1; i < nr - 1; i++)
    for (int k = 1; k < nz - 1; k++) {
      koef(s, i, k, &ke, &kh);
      AT(s, OR_E_R, i, k) = AT(s, OR_E_R, i, k) * ke -
                            (AT(s, OR_J_R, i, k) +
                             (AT(s, OR_H_PHI, i, k) - AT(s, OR_H_PHI, i, k - 1)) / dz) * kh;
      AT(s, OR_E_PHI, i, k) =
          AT(s, OR_E_PHI, i, k) * ke -
          (AT(s, OR_J_PHI, i, k) - (AT(s, OR_H_R, i, k) - AT(s, OR_H_R, i, k - 1)) / dz +
           (AT(s, OR_H_Z, i, k) - AT(s, OR_H_Z, i - 1, k)) / dr) * kh;
      AT(s, OR_E_Z, i, k) =
          AT(s, OR_E_Z, i, k) * ke -
          (AT(s, OR_J_Z, i, k) - (AT(s, OR_H_PHI, i, k) - AT(s, OR_H_PHI, i - 1, k)) / dr -
           (AT(s, OR_H_PHI, i, k) + AT(s, OR_H_PHI, i - 1, k)) / (2.0 * dr * i)) * kh;
    }
  /* E_z on the z=0 column (:154-168) */
  for (int i = 1; i < nr - 1; i++) {
    int k = 0;
    koef(s, i, k, &ke, &kh);
    AT(s, OR_E_Z, i, k) =
        AT(s, OR_E_Z, i, k) * ke -
        (AT(s, OR_J_Z, i, k) - (AT(s, OR_H_PHI, i, k) - AT(s, OR_H_PHI, i - 1, k)) / dr -
         (AT(s, OR_H_PHI, i, k) + AT(s, OR_H_PHI, i - 1, k)) / (2.0 * dr * i)) * kh;
  }
}

void or_sim_reset_j(or_sim *s) {
  /* Current::reset_j, src/current.cpp:125-135 */
  size_t n = (size_t)s->nr * (size_t)s->nz;
  memset(s->a[OR_J_R], 0, n * sizeof(double));
  memset(s->a[OR_J_PHI], 0, n * sizeof(double));
  memset(s->a[OR_J_Z], 0, n * sizeof(double));
}

/* one ring-volume weighted term as the reference's unparenthesised
 * `F * CYL_RNG_VOL(dzx, ra, rb) / vol` expands: (((F*PI)*dzx)*(rb^2-ra^2))/vol */
static double ring_term(double F, double dzx, double ra, double rb, double vol) {
  return F * PI * dzx * (rb * rb - ra * ra) / vol;
}

/* the four-node weighted sum shared by the gathers.  `lo` selects the
 * E_r / H_z / H_phi family (cell index from r - dr/2), otherwise the
 * E_z / E_phi / H_r family (cell index from r). `zhalf` selects the staggered
 * axial index (from z - dz/2). */
static double gather4(const or_sim *s, const double *F, double radius,
                      double longitude, int lo, int zhalf) {
  const double dr = s->dr, dz = s->dz;
  const int nz = s->nz;
  double r1 = radius - 0.5 * dr;
  double r3 = radius + 0.5 * dr;
  int i_r = lo ? or_cell_number(radius - 0.5 * dr, dr) : or_cell_number(radius, dr);
  int k_z = zhalf ? or_cell_number(longitude - 0.5 * dz, dz) : or_cell_number(longitude, dz);
  if (i_r < 0) i_r = 0;
  if (k_z < 0) k_z = 0;
  double vol_1, vol_2, r2, dz1, dz2;
  if (lo) {
    vol_1 = cell_volume(i_r + 1, dr, dz);
    vol_2 = cell_volume(i_r + 3, dr, dz);
    r2 = (i_r + 1) * dr;
  } else {
    if (radius > dr)
      vol_1 = cell_volume(i_r, dr, dz);
    else
      vol_1 = cyl_vol(dz, dr);
    r2 = (i_r + 0.5) * dr;
    vol_2 = cell_volume(i_r + 2, dr, dz);
  }
  if (zhalf) {
    dz1 = (k_z + 1.5) * dz - longitude;
    dz2 = longitude - (k_z + 0.5) * dz;
  } else {
    dz1 = (k_z + 1) * dz - longitude;
    dz2 = longitude - k_z * dz;
  }
  const double *row0 = F + (size_t)i_r * nz;
  const double *row1 = F + (size_t)(i_r + 1) * nz;
  double acc = 0;
  acc = acc + ring_term(row0[k_z], dz1, r1, r2, vol_1);
  acc = acc + ring_term(row1[k_z], dz1, r2, r3, vol_2);
  acc = acc + ring_term(row0[k_z + 1], dz2, r1, r2, vol_1);
  acc = acc + ring_term(row1[k_z + 1], dz2, r2, r3, vol_2);
  return acc;
}

/* EField::get_field, src/eField.cpp:204-316 */
void or_sim_get_e(const or_sim *s, double r, double z, double out[3]) {
  out[0] = gather4(s, s->a[OR_E_R], r, z, 1, 0);   /* :220-246 */
  out[2] = gather4(s, s->a[OR_E_Z], r, z, 0, 1);   /* :248-278 */
  out[1] = gather4(s, s->a[OR_E_PHI], r, z, 0, 0); /* :280-310 */
}

/* HField::get_field, src/hField.cpp:159-267 */
void or_sim_get_h(const or_sim *s, double r, double z, double out[3]) {
  out[2] = gather4(s, s->a[OR_HH_Z], r, z, 1, 0);   /* :176-200 */
  out[0] = gather4(s, s->a[OR_HH_R], r, z, 0, 1);   /* :202-233 */
  out[1] = gather4(s, s->a[OR_HH_PHI], r, z, 1, 1); /* :235-262 */
}

/* ------------------------------------------------------------------------ */
/* pushers                                                                  */
/* ------------------------------------------------------------------------ */

/* boris_pusher_single with PUSHER_BORIS_RELATIVISTIC,
 * src/particles.cpp:1025-1150 (the velocity part, after the gathers) */
int or_boris_velocity(double q, double m, double dt, const double e_in[3],
                      const double h_in[3], double vel[3]) {
  double e[3] = {e_in[0], e_in[1], e_in[2]};
  double b[3] = {h_in[0], h_in[1], h_in[2]};
  double velocity[3], vtmp[3];
  double const1 = q * dt / (2 * m);
  for (int j = 0; j < 3; j++) e[j] *= const1;
  double bscale = MAGN_CONST * const1;
  for (int j = 0; j < 3; j++) b[j] *= bscale;
  for (int j = 0; j < 3; j++) velocity[j] = (fabs(vel[j]) < MNZL) ? 0 : vel[j];
  double gamma;
  int st = get_gamma(dot3(velocity, velocity), &gamma);
  if (st) return st;
  for (int j = 0; j < 3; j++) velocity[j] *= gamma;
  for (int j = 0; j < 3; j++) velocity[j] += e[j];
  gamma = get_gamma_inv(dot3(velocity, velocity));
  for (int j = 0; j < 3; j++) b[j] *= gamma;
  double const2 = 2. / (1. + dot3(b, b));
  for (int j = 0; j < 3; j++) vtmp[j] = velocity[j];
  velocity[0] = vtmp[0] + const2 * ((vtmp[1] - vtmp[0] * b[2] + vtmp[2] * b[0]) * b[2] -
                                    (vtmp[2] + vtmp[0] * b[1] - vtmp[1] * b[0]) * b[1]);
  velocity[1] = vtmp[1] + const2 * (-(vtmp[0] + vtmp[1] * b[2] - vtmp[2] * b[1]) * b[2] +
                                    (vtmp[2] + vtmp[0] * b[1] - vtmp[1] * b[0]) * b[0]);
  velocity[2] = vtmp[2] + const2 * ((vtmp[0] + vtmp[1] * b[2] - vtmp[2] * b[1]) * b[1] -
                                    (vtmp[1] - vtmp[0] * b[2] + vtmp[2] * b[0]) * b[0]);
  for (int j = 0; j < 3; j++) velocity[j] += e[j];
  gamma = get_gamma_inv(dot3(velocity, velocity));
  for (int j = 0; j < 3; j++) vel[j] = velocity[j] * gamma;
  return OR_OK;
}

/* vay_pusher_single, src/particles.cpp:1152-1254, including its scaling of
 * b by mu0*c1, then 2*c1, then c1 and the missing 1/c^2 in us2 */
int or_vay_velocity(double q, double m, double dt, const double e_in[3],
                    const double h_in[3], double vel[3]) {
  double e[3] = {e_in[0], e_in[1], e_in[2]};
  double b[3] = {h_in[0], h_in[1], h_in[2]};
  double velocity[3], vtmp[3];
  double const1 = q * dt / (2 * m);
  double s2 = 2 * const1;
  for (int j = 0; j < 3; j++) e[j] *= s2;
  double bscale = MAGN_CONST * const1;
  for (int j = 0; j < 3; j++) b[j] *= bscale;
  for (int j = 0; j < 3; j++) velocity[j] = (fabs(vel[j]) < MNZL) ? 0 : vel[j];
  double gamma;
  int st = get_gamma(dot3(velocity, velocity), &gamma);
  if (st) return st;
  for (int j = 0; j < 3; j++) velocity[j] *= gamma;
  for (int j = 0; j < 3; j++) vtmp[j] = velocity[j];
  for (int j = 0; j < 3; j++) b[j] *= s2;
  for (int j = 0; j < 3; j++) velocity[j] += e[j];
  for (int j = 0; j < 3; j++) b[j] *= const1;
  gamma = get_gamma_inv(dot3(vtmp, vtmp));
  velocity[0] += gamma * (vtmp[1] * b[2] - vtmp[2] * b[1]);
  velocity[1] += gamma * (vtmp[2] * b[0] - vtmp[0] * b[2]);
  velocity[2] += gamma * (vtmp[0] * b[1] - vtmp[1] * b[0]);
  gamma = get_gamma_inv(dot3(velocity, velocity));
  double alpha = 1. / gamma / gamma;
  double B2 = dot3(b, b);
  double sg = alpha - B2;
  double us2 = pow(dot3(velocity, b), 2);
  alpha = 1. / sqrt(0.5 * (sg + sqrt(sg * sg + 4. * (B2 + us2))));
  for (int j = 0; j < 3; j++) b[j] *= alpha;
  sg = 1. / (1. + dot3(b, b));
  alpha = dot3(velocity, b);
  vtmp[0] = sg * (velocity[0] + alpha * b[0] + b[2] * velocity[1] - b[1] * velocity[2]);
  vtmp[1] = sg * (velocity[1] + alpha * b[1] + b[0] * velocity[2] - b[2] * velocity[0]);
  vtmp[2] = sg * (velocity[2] + alpha * b[2] + b[1] * velocity[0] - b[0] * velocity[1]);
  gamma = get_gamma_inv(dot3(vtmp, vtmp));
  for (int j = 0; j < 3; j++) vel[j] = vtmp[j] * gamma;
  return OR_OK;
}

static int finite_pos(double r, double z) { return isfinite(r) && isfinite(z); }

/* Particles::step_v, src/particles.cpp:1296-1316 (push, dump old position,
 * half step, back to rz) */
static int step_v_one(or_sim *s, species_t *p, int i) {
  if (!finite_pos(p->r[i], p->z[i])) return OR_ERR_NONFINITE;
  double e[3], h[3], v[3] = {p->vr[i], p->vphi[i], p->vz[i]};
  or_sim_get_e(s, p->r[i], p->z[i], e);
  or_sim_get_h(s, p->r[i], p->z[i], h);
  int st = (s->pusher == OR_PUSHER_VAY) ? or_vay_velocity(p->q[i], p->m[i], s->dt, e, h, v)
                                        : or_boris_velocity(p->q[i], p->m[i], s->dt, e, h, v);
  if (st) return st;
  p->vr[i] = v[0];
  p->vphi[i] = v[1];
  p->vz[i] = v[2];
  /* dump_position_to_old_single :1256-1260 */
  p->r_old[i] = p->r[i];
  p->z_old[i] = p->z[i];
  /* half_step_pos_single :1270-1284 */
  double half_dt = s->dt / 2.;
  p->r[i] = p->r[i] + p->vr[i] * half_dt;
  p->y[i] = p->y[i] + p->vphi[i] * half_dt;
  p->z[i] = p->z[i] + p->vz[i] * half_dt;
  /* back_position_to_rz_single :1286-1294 */
  double rr = sqrt(p->r[i] * p->r[i] + p->y[i] * p->y[i]);
  p->sin_t[i] = p->y[i] / rr;
  p->r[i] = rr;
  p->y[i] = 0;
  return OR_OK;
}

int or_sim_step_v(or_sim *s, int idx) {
  species_t *p = &s->sp[idx];
  int status = OR_OK;
#ifdef _OPENMP
#pragma omp parallel for num_threads(s->threads) schedule(static) if (s->threads > 1)
#endif
  for (int i = 0; i < p->n; i++) {
    if (!p->alive[i]) continue;
    int st = step_v_one(s, p, i);
    if (st) {
#ifdef _OPENMP
#pragma omp critical
#endif
      status = st;
    }
  }
  return status;
}

/* ------------------------------------------------------------------------ */
/* deposits                                                                 */
/* ------------------------------------------------------------------------ */

/* azimuthal_current_distribution, one particle, src/particles.cpp:688-784 */
void or_azimuthal_deposit(const or_sim *s, double q, double r, double z,
                          double vphi, double *jphi) {
  const double dr = s->dr, dz = s->dz;
  const size_t nz = (size_t)s->nz;
  int r_i = or_cell_number(r, dr);
  int z_k = or_cell_number(z, dz);
  if (r_i < 0) r_i = 0;
  if (z_k < 0) z_k = 0;
  double r1 = r - 0.5 * dr;
  double r2 = (r_i + 0.5) * dr;
  double r3 = r + 0.5 * dr;
  double ro_v = q / (2. * PI * dz * dr * r);
  double v_1 = (r > dr) ? cell_volume(r_i, dr, dz) : cyl_vol(dz, dr);
  double v_2 = cell_volume(r_i + 1, dr, dz);
  double dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
  double dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
  double rho;
  rho = ring_term(ro_v, dz1, r1, r2, v_1);
  jphi[r_i * nz + z_k] += rho * vphi;
  rho = ring_term(ro_v, dz1, r2, r3, v_2);
  jphi[(r_i + 1) * nz + z_k] += rho * vphi;
  rho = ring_term(ro_v, dz2, r1, r2, v_1);
  jphi[r_i * nz + z_k + 1] += rho * vphi;
  rho = ring_term(ro_v, dz2, r2, r3, v_2);
  jphi[(r_i + 1) * nz + z_k + 1] += rho * vphi;
}

typedef struct {
  double dr, dz, dt;
  size_t nz;
  double *jr, *jz;
} dep_ctx;

#define JR(d, i, k) (d)->jr[(size_t)(i) * (d)->nz + (size_t)(k)]
#define JZ(d, i, k) (d)->jz[(size_t)(i) * (d)->nz + (size_t)(k)]

/* density prefactors, include/particles.h:169-178 */
static double density_strict(double q, double r, double dr, double dz, double dt) {
  return q / (PI * 4. * r * dz * dz * dr * dt);
}
static double density_r(double q, double r, double dr, double dz, double dt) {
  return q / (2 * PI * r * dr * dz * dt * 2 * dr);
}
static double density_z(double q, double r, double dr, double dz, double dt) {
  return q / (2 * PI * r * dr * dz * dt * dz);
}

/* simple_current_distribution, src/particles.cpp:333-474: one straight
 * segment inside cell (i_n, k_n) */
static void simple_segment(const dep_ctx *d, double q, double radius_new,
                           double longitude_new, double radius_old,
                           double longitude_old, int i_n, int k_n) {
  const double dr = d->dr, dz = d->dz, delta_t = d->dt;
  double delta_r = radius_new - radius_old;
  double delta_z = longitude_new - longitude_old;
  if ((fabs(delta_r) < MNZL) || (fabs(delta_z) < MNZL)) return;
  double wj, ssd;
  if (i_n >= 1) {
    double k = delta_r / delta_z;
    double b = radius_old;
    ssd = density_r(q, i_n * dr, dr, dz, delta_t);
    wj = ssd * (dr * delta_z - k * delta_z * delta_z / 2. - delta_z * b +
                dr * dr / k * ((i_n + 0.5) * (i_n + 0.5) - 0.25) *
                    log((k * delta_z + b) / b));
    JZ(d, i_n, k_n) += wj;
    ssd = density_r(q, (i_n + 1) * dr, dr, dz, delta_t);
    wj = ssd * (k * delta_z * delta_z / 2. + delta_z * b + delta_z * dr +
                dr * dr / k * (0.25 - (i_n + 0.5) * (i_n + 0.5)) *
                    log((k * delta_z + b) / b));
    JZ(d, i_n + 1, k_n) += wj;
    k = -delta_z / delta_r;
    double r0 = (i_n + 0.5) * dr;
    double r1 = radius_old;
    b = (k_n + 1.) * dz - longitude_old;
    ssd = density_z(q, r0, dr, dz, delta_t);
    wj = ssd * (r0 * k * delta_r + k / 2. * delta_r * (radius_old + delta_r / 2.) +
                0.5 * delta_r * (b - k * (2 * r0 + r1)) +
                delta_r * (b - k * r1) * (4 * r0 * r0 - dr * dr) /
                    (8 * radius_old * (radius_old + delta_r)) +
                (k * (r0 * r0 / 2. - dr * dr / 8.)) * log((radius_old + delta_r) / radius_old));
    JR(d, i_n, k_n) += wj;
    b = longitude_old - k_n * dz;
    wj = ssd * (-r0 * k * delta_r - k / 2. * delta_r * (radius_old + delta_r / 2.) +
                0.5 * delta_r * (b + k * (2 * r0 + r1)) +
                delta_r * (b + k * r1) * (4 * r0 * r0 - dr * dr) /
                    (8 * radius_old * (radius_old + delta_r)) -
                (k * (r0 * r0 / 2. - dr * dr / 8.)) * log((radius_old + delta_r) / radius_old));
    JR(d, i_n, k_n + 1) += wj;
  } else {
    double k = delta_r / delta_z;
    double b = radius_old;
    ssd = q / (2. * PI * dr / 4. * dr * dz * delta_t * dr);
    wj = ssd * (dr * delta_z - k * delta_z * delta_z / 2. - delta_z * b);
    JZ(d, i_n, k_n) += wj;
    ssd = density_r(q, dr, dr, dz, delta_t);
    wj = ssd * (k * delta_z * delta_z / 2. + delta_z * dr + delta_z * b);
    JZ(d, i_n + 1, k_n) += wj;
    k = -delta_z / delta_r;
    double r0 = (i_n + 0.5) * dr;
    double r1 = radius_old;
    b = (k_n + 1.) * dz - longitude_old;
    ssd = density_z(q, r0, dr, dz, delta_t);
    wj = ssd * (r0 * k * delta_r + k / 2. * delta_r * (radius_old + delta_r / 2.) +
                0.5 * delta_r * (b - k * (2 * r0 + r1)) +
                delta_r * (b - k * r1) * (4 * r0 * r0 - dr * dr) /
                    (8 * radius_old * (radius_old + delta_r)) +
                (k * (r0 * r0 / 2. - dr * dr / 8.)) * log((radius_old + delta_r) / radius_old));
    JR(d, i_n, k_n) += wj;
    b = longitude_old - k_n * dz;
    wj = ssd * (-r0 * k * delta_r - k / 2. * delta_r * (radius_old + delta_r / 2.) +
                0.5 * delta_r * (b + k * (2 * r0 + r1)) +
                delta_r * (b + k * r1) * (4 * r0 * r0 - dr * dr) /
                    (8 * radius_old * (radius_old + delta_r)) -
                (k * (r0 * r0 / 2. - dr * dr / 8.)) * log((radius_old + delta_r) / radius_old));
    JR(d, i_n, k_n + 1) += wj;
  }
}

/* strict_motion_weighting, src/particles.cpp:786-988 */
static void strict_motion(const dep_ctx *d, double q, double radius_new,
                          double longitude_new, double radius_old,
                          double longitude_old) {
  const double dr = d->dr, dz = d->dz, delta_t = d->dt;
  int i_n = or_cell_number(radius_new, dr);
  int k_n = or_cell_number(longitude_new, dz);
  int i_o = or_cell_number(radius_old, dr);
  int k_o = or_cell_number(longitude_old, dz);
  if (i_n < 0) i_n = 0;
  if (k_n < 0) k_n = 0;
  if (i_o < 0) i_o = 0;
  if (k_o < 0) k_o = 0;
  if ((fabs(radius_new - radius_old) < MNZL) && (fabs(longitude_new - longitude_old) < MNZL))
    return;
  if (fabs(radius_new - radius_old) < MNZL) {
    /* strictly axial motion (:810-887) */
    double value_part = 2. * PI * radius_new * dr * dz;
    double r1 = radius_new - 0.5 * dr;
    double r2 = (i_n + 0.5) * dr;
    double r3 = radius_new + 0.5 * dr;
    double wj_lower, wj_upper, wj, delta_z;
    if (i_n == 0)
      wj_lower = q / (delta_t * PI * dr * dr / 4.) * PI * (r2 * r2 - r1 * r1) / value_part;
    else
      wj_lower = q / (delta_t * 2. * PI * i_n * dr * dr) * PI * (r2 * r2 - r1 * r1) / value_part;
    wj_upper = q / (delta_t * 2 * PI * (i_n + 1) * dr * dr) * PI * (r3 * r3 - r2 * r2) / value_part;
    /* the reference adds 0 to J_r[i_n][k_n], J_r[i_n][k_n+1] here (:831-832) */
    JR(d, i_n, k_n) += 0.;
    JR(d, i_n, k_n + 1) += 0.;
    int res_k = k_n - k_o;
    switch (res_k) {
      case 0:
        delta_z = longitude_new - longitude_old;
        wj = wj_lower * delta_z;
        JZ(d, i_n, k_n) += wj;
        wj = wj_upper * delta_z;
        JZ(d, i_n + 1, k_n) += wj;
        break;
      case 1:
        delta_z = k_n * dz - longitude_old;
        wj = wj_lower * delta_z;
        JZ(d, i_n, k_n - 1) += wj;
        wj = wj_upper * delta_z;
        JZ(d, i_n + 1, k_n - 1) += wj;
        delta_z = longitude_new - k_n * dz;
        wj = wj_lower * delta_z;
        JZ(d, i_n, k_n) += wj;
        wj = wj_upper * delta_z;
        JZ(d, i_n + 1, k_n) += wj;
        break;
      case -1:
        delta_z = (k_n + 1) * dz - longitude_old;
        wj = wj_lower * delta_z;
        JZ(d, i_n, k_n + 1) += wj;
        wj = wj_upper * delta_z;
        JZ(d, i_n + 1, k_n + 1) += wj;
        delta_z = longitude_new - (k_n + 1) * dz;
        wj = wj_lower * delta_z;
        JZ(d, i_n, k_n) += wj;
        wj = wj_upper * delta_z;
        JZ(d, i_n + 1, k_n) += wj;
        break;
      default:
        break;
    }
  } else if (fabs(longitude_new - longitude_old) < MNZL) {
    /* strictly radial motion (:890-986) */
    double wj, delta_r, left_delta_z, right_delta_z, res_j, r0, ssd;
    int res_i = i_n - i_o;
    switch (res_i) {
      case 0:
        delta_r = radius_new - radius_old;
        left_delta_z = (k_n + 1) * dz - longitude_new;
        right_delta_z = longitude_new - k_n * dz;
        r0 = (i_n + 0.5) * dr;
        ssd = density_strict(q, r0, dr, dz, delta_t);
        wj = ssd * (delta_r - r0 * r0 / (radius_old + delta_r) + r0 * r0 / radius_old +
                    dr * dr / (4. * (radius_old + delta_r)) - dr * dr / (4. * radius_old));
        res_j = wj * left_delta_z;
        JR(d, i_n, k_n) += res_j;
        res_j = wj * right_delta_z;
        JR(d, i_n, k_n + 1) += res_j;
        break;
      case 1:
        delta_r = (i_n)*dr - radius_old;
        left_delta_z = (k_n + 1) * dz - longitude_new;
        right_delta_z = longitude_new - k_n * dz;
        r0 = (i_n - 0.5) * dr;
        ssd = density_strict(q, r0, dr, dz, delta_t);
        wj = ssd * (delta_r - r0 * r0 / (radius_old + delta_r) + r0 * r0 / radius_old +
                    dr * dr / (4. * (radius_old + delta_r)) - dr * dr / (4. * radius_old));
        res_j = wj * left_delta_z;
        JR(d, i_n - 1, k_n) += res_j;
        res_j = wj * right_delta_z;
        JR(d, i_n - 1, k_n + 1) += res_j;
        delta_r = radius_new - i_n * dr;
        r0 = (i_n + 0.5) * dr;
        ssd = density_strict(q, r0, dr, dz, delta_t);
        /* `r0 * r0 / i_n * dr` is ((r0*r0)/i_n)*dr in the reference (:941) */
        wj = ssd * (delta_r - r0 * r0 / (i_n * dr + delta_r) + r0 * r0 / i_n * dr +
                    dr * dr / (4. * (i_n * dr + delta_r)) - dr * dr / (4. * i_n * dr));
        res_j = wj * left_delta_z;
        JR(d, i_n, k_n) += res_j;
        res_j = wj * right_delta_z;
        JR(d, i_n, k_n + 1) += res_j;
        break;
      case -1:
        delta_r = (i_n + 1) * dr - radius_old;
        left_delta_z = (k_n + 1) * dz - longitude_new;
        right_delta_z = longitude_new - k_n * dz;
        r0 = (i_n + 1.5) * dr;
        ssd = density_strict(q, r0, dr, dz, delta_t);
        wj = ssd * (delta_r - r0 * r0 / (radius_old + delta_r) + r0 * r0 / radius_old +
                    dr * dr / (4. * (radius_old + delta_r)) - dr * dr / (4. * radius_old));
        res_j = wj * left_delta_z;
        JR(d, i_n + 1, k_n) += res_j;
        res_j = wj * right_delta_z;
        JR(d, i_n + 1, k_n + 1) += res_j;
        delta_r = radius_new - (i_n + 1) * dr;
        r0 = (i_n + 0.5) * dr;
        ssd = density_strict(q, r0, dr, dz, delta_t);
        /* same precedence as above (:976) */
        wj = ssd * (delta_r - r0 * r0 / ((i_n + 1) * dr + delta_r) + r0 * r0 / (i_n + 1) * dr +
                    dr * dr / (4. * ((i_n + 1) * dr + delta_r)) - dr * dr / (4. * (i_n + 1) * dr));
        res_j = wj * left_delta_z;
        JR(d, i_n, k_n) += res_j;
        res_j = wj * right_delta_z;
        JR(d, i_n, k_n + 1) += res_j;
        break;
      default:
        break;
    }
  }
}

/* current_distribution body for one particle, src/particles.cpp:480-685 */
static void deposit_one(const dep_ctx *d, double q, double pr, double pz,
                        double por, double poz) {
  const double dr = d->dr, dz = d->dz;
  int i_n = or_cell_number(pr, dr);
  int k_n = or_cell_number(pz, dz);
  int i_o = or_cell_number(por, dr);
  int k_o = or_cell_number(poz, dz);
  if (i_n < 0) i_n = 0;
  if (k_n < 0) k_n = 0;
  if (i_o < 0) i_o = 0;
  if (k_o < 0) k_o = 0;
  /* boundary-equality remaps (:496-499) */
  if (por == (i_o + 1) * dr) i_o = i_n;
  if (poz == (k_o + 1) * dz) k_o = k_n;
  if (pr == (i_n + 1) * dr) i_n = i_o;
  if (pz == (k_n + 1) * dz) k_n = k_o;
  int res_cell = abs(i_n - i_o) + abs(k_n - k_o);
  if ((fabs(pr - por) < MNZL) || (fabs(pz - poz) < MNZL)) {
    strict_motion(d, q, pr, pz, por, poz);
    return;
  }
  switch (res_cell) {
    case 0:
      simple_segment(d, q, pr, pz, por, poz, i_n, k_n);
      break;
    case 1:
      if ((i_n != i_o) && (k_n == k_o)) {
        if (por > (i_n + 1) * dr) {
          /* inward across r = (i_n+1) dr (:517-528) */
          double a = (por - pr) / (poz - pz);
          double r_boundary = (i_n + 1) * dr;
          double delta_r = r_boundary - pr;
          double z_boundary = pz + delta_r / a;
          simple_segment(d, q, r_boundary, z_boundary, por, poz, i_n + 1, k_n);
          simple_segment(d, q, pr, pz, r_boundary, z_boundary, i_n, k_n);
        } else {
          /* outward across r = i_n dr (:530-539) */
          double a = (pr - por) / (pz - poz);
          double r_boundary = (i_n)*dr;
          double delta_r = r_boundary - por;
          double z_boundary = poz + delta_r / a;
          simple_segment(d, q, r_boundary, z_boundary, por, poz, i_n - 1, k_n);
          simple_segment(d, q, pr, pz, r_boundary, z_boundary, i_n, k_n);
        }
      } else if ((i_n == i_o) && (k_n != k_o)) {
        if (poz < k_n * dz) {
          /* forward across z = k_n dz (:545-553) */
          double z_boundary = k_n * dz;
          double delta_z = z_boundary - poz;
          double a = (pr - por) / (pz - poz);
          double r_boundary = por + a * delta_z;
          simple_segment(d, q, r_boundary, z_boundary, por, poz, i_n, k_n - 1);
          simple_segment(d, q, pr, pz, r_boundary, z_boundary, i_n, k_n);
        } else {
          /* backward across z = (k_n+1) dz (:555-563) */
          double z_boundary = (k_n + 1) * dz;
          double delta_z = z_boundary - pz;
          double a = (por - pr) / (poz - pz);
          double r_boundary = pr + a * delta_z;
          simple_segment(d, q, r_boundary, z_boundary, por, poz, i_n, k_n + 1);
          simple_segment(d, q, pr, pz, r_boundary, z_boundary, i_n, k_n);
        }
      }
      break;
    case 2:
      if (i_o < i_n) {
        if (k_o < k_n) {
          /* [i-1][k-1] -> [i][k] (:575-594) */
          double a = (pr - por) / (pz - poz);
          double r1 = i_n * dr;
          double delta_z1 = (r1 - por) / a;
          double z1 = poz + delta_z1;
          double z2 = k_n * dz;
          double delta_r2 = (z2 - poz) * a;
          double r2 = por + delta_r2;
          if (z1 < k_n * dz) {
            simple_segment(d, q, r1, z1, por, poz, i_n - 1, k_n - 1);
            simple_segment(d, q, r2, z2, r1, z1, i_n, k_n - 1);
            simple_segment(d, q, pr, pz, r2, z2, i_n, k_n);
          } else if (z1 > k_n * dz) {
            simple_segment(d, q, r2, z2, por, poz, i_n - 1, k_n - 1);
            simple_segment(d, q, r1, z1, r2, z2, i_n - 1, k_n);
            simple_segment(d, q, pr, pz, r1, z1, i_n, k_n);
          }
        } else {
          /* [i-1][k+1] -> [i][k] (:597-620) */
          double a = (pr - por) / (pz - poz);
          double r1 = i_n * dr;
          double delta_z1 = (r1 - por) / a;
          double z1 = poz + delta_z1;
          double z2 = (k_n + 1) * dz;
          double delta_r2 = -(poz - z2) * a;
          double r2 = por + delta_r2;
          if (z1 > (k_n + 1) * dz) {
            simple_segment(d, q, r1, z1, por, poz, i_n - 1, k_n + 1);
            simple_segment(d, q, r2, z2, r1, z1, i_n, k_n + 1);
            simple_segment(d, q, pr, pz, r2, z2, i_n, k_n);
          } else if (z1 < (k_n + 1) * dz) {
            simple_segment(d, q, r2, z2, por, poz, i_n - 1, k_n + 1);
            simple_segment(d, q, r1, z1, r2, z2, i_n - 1, k_n);
            simple_segment(d, q, pr, pz, r1, z1, i_n, k_n);
          }
        }
      } else if (i_o > i_n) {
        if (k_o < k_n) {
          /* [i+1][k-1] -> [i][k] (:627-649) */
          double a = (pr - por) / (pz - poz);
          double r1 = (i_n + 1) * dr;
          double delta_z1 = -(por - r1) / a;
          double z1 = poz + delta_z1;
          double z2 = k_n * dz;
          double delta_r2 = -(z2 - poz) * a;
          double r2 = por - delta_r2;
          if (z1 < (k_n)*dz) {
            simple_segment(d, q, r1, z1, por, poz, i_n + 1, k_n - 1);
            simple_segment(d, q, r2, z2, r1, z1, i_n, k_n - 1);
            simple_segment(d, q, pr, pz, r2, z2, i_n, k_n);
          } else if (z1 > (k_n)*dz) {
            simple_segment(d, q, r2, z2, por, poz, i_n + 1, k_n - 1);
            simple_segment(d, q, r1, z1, r2, z2, i_n + 1, k_n);
            simple_segment(d, q, pr, pz, r1, z1, i_n, k_n);
          }
        } else if (k_o > k_n) {
          /* [i+1][k+1] -> [i][k] (:652-675) */
          double a = (por - pr) / (poz - pz);
          double r1 = (i_n + 1) * dr;
          double delta_z1 = (r1 - pr) / a;
          double z1 = pz + delta_z1;
          double z2 = (k_n + 1) * dz;
          double delta_r2 = (z2 - pz) * a;
          double r2 = pr + delta_r2;
          if (z1 > (k_n + 1) * dz) {
            simple_segment(d, q, r1, z1, por, poz, i_n + 1, k_n + 1);
            simple_segment(d, q, r2, z2, r1, z1, i_n, k_n + 1);
            simple_segment(d, q, pr, pz, r2, z2, i_n, k_n);
          } else if (z1 < (k_n + 1) * dz) {
            simple_segment(d, q, r2, z2, por, poz, i_n + 1, k_n + 1);
            simple_segment(d, q, r1, z1, r2, z2, i_n + 1, k_n);
            simple_segment(d, q, pr, pz, r1, z1, i_n, k_n);
          }
        }
      }
      break;
    default:
      break;
  }
}

void or_current_deposit(const or_sim *s, double q, double r_new, double z_new,
                        double r_old, double z_old, double *jr, double *jz) {
  dep_ctx d = {s->dr, s->dz, s->dt, (size_t)s->nz, jr, jz};
  deposit_one(&d, q, r_new, z_new, r_old, z_old);
}

/* Particles::charge_weighting, one particle, src/particles.cpp:101-226 */
void or_charge_deposit(const or_sim *s, double q, double r, double z, double *rho) {
  const double dr = s->dr, dz = s->dz;
  const size_t nz = (size_t)s->nz;
  int r_i = or_cell_number(r, dr);
  int z_k = or_cell_number(z, dz);
  if (r_i < 0) r_i = 0;
  if (z_k < 0) z_k = 0;
  double r1, r2, r3, dz1, dz2, ro_v, v_1, v_2, value;
  if (r > dr) {
    r1 = r - 0.5 * dr;
    r2 = (r_i + 0.5) * dr;
    r3 = r + 0.5 * dr;
    ro_v = q / (2. * PI * dz * dr * r);
    v_1 = cell_volume(r_i, dr, dz);
    v_2 = cell_volume(r_i + 1, dr, dz);
    dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
    dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
    value = ring_term(ro_v, dz1, r1, r2, v_1);
    rho[r_i * nz + z_k] += value;
    value = ring_term(ro_v, dz1, r2, r3, v_2);
    rho[(r_i + 1) * nz + z_k] += value;
    value = ring_term(ro_v, dz2, r1, r2, v_1);
    rho[r_i * nz + z_k + 1] += value;
    value = ring_term(ro_v, dz2, r2, r3, v_2);
    rho[(r_i + 1) * nz + z_k + 1] += value;
  } else if (r <= dr / 2.) {
    r_i = 0;
    r2 = (r_i + 0.5) * dr;
    r3 = r + 0.5 * dr;
    dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
    dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
    ro_v = q / (PI * dz * (2. * r * r + dr * dr / 2.));
    v_1 = cyl_vol(dz, dr);
    v_2 = cell_volume(r_i + 1, dr, dz);
    value = ro_v * PI * dz1 * (dr * dr / 2. - r * dr + r * r) / v_1;
    rho[r_i * nz + z_k] += value;
    value = ring_term(ro_v, dz1, r2, r3, v_2);
    rho[(r_i + 1) * nz + z_k] += value;
    value = ro_v * PI * dz2 * (dr * dr / 2. - r * dr + r * r) / v_1;
    rho[r_i * nz + z_k + 1] += value;
    value = ring_term(ro_v, dz2, r2, r3, v_2);
    rho[(r_i + 1) * nz + z_k + 1] += value;
  } else {
    r1 = r - 0.5 * dr;
    r2 = (r_i + 0.5) * dr;
    r3 = r + 0.5 * dr;
    dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
    dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
    ro_v = q / (2. * PI * dz * dr * r);
    v_1 = cyl_vol(dz, dr);
    v_2 = cell_volume(r_i + 1, dr, dz);
    value = ring_term(ro_v, dz1, r1, r2, v_1);
    rho[r_i * nz + z_k] += value;
    value = ring_term(ro_v, dz1, r2, r3, v_2);
    rho[(r_i + 1) * nz + z_k] += value;
    value = ring_term(ro_v, dz2, r1, r2, v_1);
    rho[r_i * nz + z_k + 1] += value;
    value = ring_term(ro_v, dz2, r2, r3, v_2);
    rho[(r_i + 1) * nz + z_k + 1] += value;
  }
}

int or_sim_charge_weighting(or_sim *s, int idx, double *rho) {
  species_t *p = &s->sp[idx];
  if (!rho) rho = s->a[OR_RHO_BUNCH];
  for (int i = 0; i < p->n; i++) {
    if (!p->alive[i]) continue;
    if (!finite_pos(p->r[i], p->z[i])) return OR_ERR_NONFINITE;
    or_charge_deposit(s, p->q[i], p->r[i], p->z[i], rho);
  }
  return OR_OK;
}

/* ---- probes ---- */

/* TemperatureClassic::calc_t + normalize, src/temperatureClassic.cpp:57-158.
 * Per cell (CELL_NUMBER of r and z, :89-90) the sums of |v| (lib::sq_rt, the
 * plain sqrt of the non-AVX build, src/lib.cpp:25-32), v_r, v_phi, v_z and
 * the count over the living particles in index order; then per cell
 * v^2 = (S^2 - |V|^2) / N^2 and T = m v^2 / 2 below REL_LIMIT^2 (5e7 m/s,
 * :4-6), else m c^2 (gamma - 1) with lib::get_gamma (src/lib.cpp:45-70), in
 * eV (divided by |charge|, :147).  m and charge are the species' per real
 * particle.  out: nr x nz, row nr-1 and column nz-1 stay 0 (reset() clears
 * them and normalize never writes them); an empty cell gives NaN (0/0). */
int or_temperature(const or_sim *s, int idx, double *out) {
  const species_t *p = &s->sp[idx];
  const size_t cells = (size_t)s->nr * (size_t)s->nz;
  double *sum = calloc(4 * cells, sizeof(double));
  double *cnt = calloc(cells, sizeof(double));
  if (!sum || !cnt) {
    free(sum);
    free(cnt);
    return OR_ERR_OOM;
  }
  double *vr = sum + cells, *vphi = sum + 2 * cells, *vz = sum + 3 * cells;
  for (int j = 0; j < p->n; j++) {
    if (!p->alive[j]) continue;
    if (!finite_pos(p->r[j], p->z[j])) {
      free(sum);
      free(cnt);
      return OR_ERR_NONFINITE;
    }
    int i = or_cell_number(p->r[j], s->dr), k = or_cell_number(p->z[j], s->dz);
    if (i < 0 || k < 0 || i >= s->nr || k >= s->nz) continue;
    size_t c = (size_t)i * (size_t)s->nz + (size_t)k;
    double a = p->vr[j], b = p->vphi[j], e = p->vz[j];
    sum[c] += sqrt(a * a + b * b + e * e);
    vr[c] += a;
    vphi[c] += b;
    vz[c] += e;
    cnt[c] += 1;
  }
  const double rel_limit = 5e7;
  for (size_t c = 0; c < cells; c++) out[c] = 0.0;
  for (int i = 0; i < s->nr - 1; i++)
    for (int k = 0; k < s->nz - 1; k++) {
      size_t c = (size_t)i * (size_t)s->nz + (size_t)k;
      double vec2 = vr[c] * vr[c] + vphi[c] * vphi[c] + vz[c] * vz[c];
      double v2 = (sum[c] * sum[c] - vec2) / (cnt[c] * cnt[c]);
      double t;
      if (v2 < rel_limit * rel_limit) {
        t = p->mass * v2 / 2;
      } else {
        double g;
        if (get_gamma(v2, &g) != OR_OK) {
          free(sum);
          free(cnt);
          return OR_ERR_SUPERLUMINAL;
        }
        t = p->mass * C2 * (g - 1);
      }
      out[c] = t / fabs(p->charge);
    }
  free(sum);
  free(cnt);
  return OR_OK;
}

/* ParticlesDensity::calc_density, src/particlesDensity.cpp:40-225: each living
 * particle (index order) spreads ro_p = mass_array / EL_MASS / v_0 over the
 * four nodes (i..i+1, k..k+1) with ring-volume weights; three radial branches
 * (r > dr, r <= dr/2, else), each with its own v_0 and axis-cell volume, in
 * the reference's expression order (CYL_RNG_VOL, CELL_VOLUME and CYL_VOL of
 * include/geometry.h:33-39).  out: nr x nz, accumulated (the caller clears
 * it, as ParticlesDensity::reset). */
int or_density(const or_sim *s, int idx, double *out) {
  const species_t *p = &s->sp[idx];
  const double dr = s->dr, dz = s->dz;
  const size_t nz = (size_t)s->nz;
  for (int j = 0; j < p->n; j++) {
    if (!p->alive[j]) continue;
    double r = p->r[j], z = p->z[j];
    if (!finite_pos(r, z)) return OR_ERR_NONFINITE;
    int r_i = or_cell_number(r, dr), z_k = or_cell_number(z, dz);
    if (r_i < 0 || z_k < 0 || r_i > s->nr - 2 || z_k > s->nz - 2) continue;
    double r1, r2, r3, dz1, dz2, v_0, v_1, v_2, ro_p, w00, w10, w01, w11;
    if (r > dr) {
      r1 = r - 0.5 * dr;
      r2 = (r_i + 0.5) * dr;
      r3 = r + 0.5 * dr;
      v_0 = 2. * PI * dz * dr * r;
      ro_p = p->m[j] / EL_MASS / v_0;
      v_1 = cell_volume(r_i, dr, dz);
      v_2 = cell_volume(r_i + 1, dr, dz);
      dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
      dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
      w00 = PI * (dz1) * ((r2) * (r2) - (r1) * (r1)) / v_1;
      w10 = PI * (dz1) * ((r3) * (r3) - (r2) * (r2)) / v_2;
      w01 = PI * (dz2) * ((r2) * (r2) - (r1) * (r1)) / v_1;
      w11 = PI * (dz2) * ((r3) * (r3) - (r2) * (r2)) / v_2;
    } else if (r <= dr / 2.) {
      r_i = 0;
      r2 = (r_i + 0.5) * dr;
      r3 = r + 0.5 * dr;
      dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
      dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
      v_0 = PI * dz * (2. * r * r + dr * dr / 2.);
      ro_p = p->m[j] / EL_MASS / v_0;
      v_1 = cyl_vol(dz, dr);
      v_2 = cell_volume(r_i + 1, dr, dz);
      w00 = PI * dz1 * (dr * dr / 2. - r * dr + r * r) / v_1;
      w10 = PI * (dz1) * ((r3) * (r3) - (r2) * (r2)) / v_2;
      w01 = PI * dz2 * (dr * dr / 2. - r * dr + r * r) / v_1;
      w11 = PI * (dz2) * ((r3) * (r3) - (r2) * (r2)) / v_2;
    } else {
      r1 = r - 0.5 * dr;
      r2 = (r_i + 0.5) * dr;
      r3 = r + 0.5 * dr;
      dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
      dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
      v_0 = 2. * PI * dz * dr * r;
      ro_p = p->m[j] / EL_MASS / v_0;
      v_1 = cyl_vol(dz, dr);
      v_2 = cell_volume(r_i + 1, dr, dz);
      w00 = PI * (dz1) * ((r2) * (r2) - (r1) * (r1)) / v_1;
      w10 = PI * (dz1) * ((r3) * (r3) - (r2) * (r2)) / v_2;
      w01 = PI * (dz2) * ((r2) * (r2) - (r1) * (r1)) / v_1;
      w11 = PI * (dz2) * ((r3) * (r3) - (r2) * (r2)) / v_2;
    }
    size_t c = (size_t)r_i * nz + (size_t)z_k;
    out[c] += ro_p * w00;
    out[c + nz] += ro_p * w10;
    out[c + 1] += ro_p * w01;
    out[c + nz + 1] += ro_p * w11;
  }
  return OR_OK;
}

/* reflection_single (plasma, src/particles.cpp:1318-1353) and
 * Bunch::reflection_single (src/bunch.cpp:126-153), applied to the Cartesian
 * half-step position before it is rotated back to r-z */
static void reflect_one(const or_sim *s, species_t *p, int i) {
  double dr = s->dr, dz = s->dz;
  double radius_wall = s->r_size - dr / 2.;
  double longitude_wall = s->z_size - dz / 2.;
  double half_dr = dr / 2.;
  double half_dz = dz / 2.;
  if (p->is_bunch) {
    if (p->r[i] > radius_wall) p->alive[i] = 0;
    if (p->z[i] > longitude_wall) p->alive[i] = 0;
  } else {
    double radius_wallX2 = radius_wall * 2.;
    double longitude_wallX2 = longitude_wall * 2.;
    if (p->r[i] > radius_wall) {
      p->r[i] = radius_wallX2 - p->r[i];
      p->vr[i] = -p->vr[i];
    }
    if (p->z[i] > longitude_wall) {
      p->z[i] = longitude_wallX2 - p->z[i];
      p->vz[i] = -p->vz[i];
    }
  }
  if (p->r[i] < half_dr) {
    p->r[i] = dr - p->r[i];
    p->vr[i] = -p->vr[i];
  }
  if (p->z[i] < half_dz) {
    p->z[i] = dz - p->z[i];
    p->vz[i] = -p->vz[i];
  }
}

/* Particles::full_current_distribution, src/particles.cpp:1367-1373:
 * azimuthal J, second half step with walls, J_r/J_z, velocity rotation */
int or_sim_full_current_distribution(or_sim *s, int idx) {
  species_t *p = &s->sp[idx];
  const double half_dt = s->dt / 2.;
  int status = OR_OK;
  /* azimuthal_current_distribution (serial, :688-784) */
  for (int j = 0; j < p->n; j++) {
    int i = s->reverse_deposit ? p->n - 1 - j : j;
    if (p->alive[i])
      or_azimuthal_deposit(s, p->q[i], p->r[i], p->z[i], p->vphi[i], s->a[OR_J_PHI]);
  }
  /* move_half_reflect (:1355-1365) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(s->threads) schedule(static) if (s->threads > 1)
#endif
  for (int i = 0; i < p->n; i++) {
    if (!p->alive[i]) continue;
    if (!finite_pos(p->r[i], p->z[i])) {
#ifdef _OPENMP
#pragma omp critical
#endif
      status = OR_ERR_NONFINITE;
      continue;
    }
    p->r[i] = p->r[i] + p->vr[i] * half_dt;
    p->y[i] = p->y[i] + p->vphi[i] * half_dt;
    p->z[i] = p->z[i] + p->vz[i] * half_dt;
    reflect_one(s, p, i);
    double rr = sqrt(p->r[i] * p->r[i] + p->y[i] * p->y[i]);
    p->sin_t[i] = p->y[i] / rr;
    p->r[i] = rr;
    p->y[i] = 0;
  }
  if (status) return status;
  /* current_distribution (serial, :476-686) */
  dep_ctx d = {s->dr, s->dz, s->dt, (size_t)s->nz, s->a[OR_J_R], s->a[OR_J_Z]};
  for (int j = 0; j < p->n; j++) {
    int i = s->reverse_deposit ? p->n - 1 - j : j;
    if (p->alive[i]) deposit_one(&d, p->q[i], p->r[i], p->z[i], p->r_old[i], p->z_old[i]);
  }
  /* back_velocity_to_rz (:1005-1023) */
#ifdef _OPENMP
#pragma omp parallel for num_threads(s->threads) schedule(static) if (s->threads > 1)
#endif
  for (int i = 0; i < p->n; i++) {
    if (!p->alive[i]) continue;
    double st = p->sin_t[i];
    double ct = sqrt(1 - st * st);
    double u_2 = ct * p->vr[i] - st * p->vphi[i];
    double v_2 = st * p->vr[i] + ct * p->vphi[i];
    p->vr[i] = u_2;
    p->vphi[i] = v_2;
  }
  return OR_OK;
}

/* Bunch::bunch_inject, src/bunch.cpp:87-124 */
void or_sim_bunch_inject(or_sim *s, int idx) {
  species_t *p = &s->sp[idx];
  double bunch_time_begin = p->duration * p->bunch_number + p->hole_duration * p->bunch_number;
  double bunch_time_end = bunch_time_begin + p->duration;
  double dl = p->velocity * s->dt;
  int steps_amount = (int)ceil(p->duration / s->dt);
  if (steps_amount <= 0) return; /* a loaded bunch: nothing to inject */
  unsigned int particles_in_step = (unsigned int)p->n / (unsigned int)steps_amount;
  int start_number = (int)((s->current_time - bunch_time_begin) / s->dt * particles_in_step);
  double half_z_cell_size = s->dz / 2.0;
  if (s->current_time >= bunch_time_begin && s->current_time < bunch_time_end) {
    for (unsigned int i = 0; i < particles_in_step; i++) {
      unsigned int j = i + (unsigned int)start_number;
      if (j < (unsigned int)p->n) {
        p->z[j] = dl * p->inject_u[j] + half_z_cell_size;
        p->vz[j] = p->velocity;
        p->alive[j] = 1;
      }
    }
  }
}

/* LoadInitParam::init_beam, src/loadInitParam.cpp:168-248 (bunch creation in
 * order of bunch number, then injection of every created bunch) */
static void init_beam(or_sim *s) {
  int cand = -1;
  for (int j = 0; j < s->nsp; j++)
    if (s->sp[j].is_bunch && !s->sp[j].created && s->sp[j].bunch_number == s->next_bunch)
      cand = j;
  if (cand >= 0) {
    species_t *p = &s->sp[cand];
    double bunch_inject_time = s->next_bunch * (p->duration + p->hole_duration);
    if (s->current_time >= bunch_inject_time) {
      p->created = 1;
      s->order[s->norder++] = cand;
      s->next_bunch++;
    }
  }
  for (int o = 0; o < s->norder; o++) {
    int j = s->order[o];
    if (s->sp[j].is_bunch) or_sim_bunch_inject(s, j);
  }
}

/* LoadInitParam::run loop body, src/loadInitParam.cpp:471-587 */
int or_sim_step(or_sim *s, int nsteps) {
  for (int n = 0; n < nsteps; n++) {
    init_beam(s);
    or_sim_h_calc(s);
    or_sim_reset_j(s);
    size_t npts = (size_t)s->nr * (size_t)s->nz;
    memset(s->a[OR_RHO_BUNCH], 0, npts * sizeof(double));
    for (int o = 0; o < s->norder; o++) {
      int st = or_sim_step_v(s, s->order[o]);
      if (st) return st;
    }
    for (int o = 0; o < s->norder; o++) {
      int st = or_sim_full_current_distribution(s, s->order[o]);
      if (st) return st;
    }
    or_sim_e_calc(s);
    for (int o = 0; o < s->norder; o++) {
      int j = s->order[o];
      if (s->sp[j].is_bunch) {
        int st = or_sim_charge_weighting(s, j, s->a[OR_RHO_BUNCH]);
        if (st) return st;
      }
    }
    s->current_time = s->current_time + s->dt;
  }
  return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* initial loaders (only to replay the reference's smoke goldens)           */
/* ------------------------------------------------------------------------ */

/* math::random::random_reverse, src/math/rand.cpp:37-53: radical inverse */
double or_random_reverse(double vel, int power) {
  int int_vel = (int)floor(vel);
  double ost = 0;
  double r = 0;
  int order = 1;
  while (int_vel >= 1) {
    ost = int_vel % power;
    r = r + ost * pow((double)power, (-order));
    int_vel = (int)((int_vel - ost) / power);
    order = order + 1;
  }
  return r;
}

/* Particles::load_cylindrical_spatial_distribution, src/particles.cpp:278-331 */
void or_load_cylindrical(int n, double r_size, double z_size, double dr,
                         double dz, double n1, double n2,
                         double left_plasma_boundary, double *r, double *z,
                         double *w) {
  double dn = n2 - n1;
  double v_sum = 0;
  for (int j = 0; j < n; j++) {
    double rand_r = or_random_reverse(j, 13);
    double rand_z = or_random_reverse(n - 1 - j, 11);
    r[j] = (r_size - dr) * rand_r + dr / 2;
    if (dn == 0)
      z[j] = rand_z * (z_size - dz) + left_plasma_boundary + dz / 2;
    else
      z[j] = (z_size - left_plasma_boundary - dz) / dn *
                 (sqrt(pow(n1, 2) + rand_z * (2 * n1 * dn + pow(dn, 2))) - n1) +
             left_plasma_boundary + dz / 2;
    v_sum += 2 * PI * r[j] * dr * dz;
  }
  double v_avg = v_sum / n;
  double N_total = (n1 + n2) / 2 * PI * r_size * r_size * z_size;
  double n_per_macro_avg = N_total / n;
  for (int j = 0; j < n; j++) {
    double norm;
    if (r[j] > dr / 2)
      norm = 2 * PI * r[j] * dr * dz / v_avg;
    else
      norm = PI * (r[j] * r[j] + r[j] * dr + dr * dr / 4) * dz / v_avg;
    w[j] = n_per_macro_avg * norm;
  }
}

/* log of the inverse of H(x) = log Q(3, x) = -x + log(1 + x + x^2/2) at
 * y_i = -30 + 42 i / 999 (the tabulation used by the Maxwell-Juttner sampler
 * of src/math/maxwellJuttner.cpp:167, after Smilei, Derouillat et al. 2018).
 * Computed here from the closed form: solve F(x) = x - log(1+x+x^2/2) = e^y
 * by bisection on ln x, with F evaluated as -log1p(-P(3,x)) for x < 1 to
 * avoid cancellation. */
static double mj_F(double x) {
  if (x < 1.0) {
    double s = 0.0, t = x * x * x / 6.0;
    int n = 3;
    for (;;) {
      s += t;
      n++;
      t *= x / n;
      if (t < 1e-20 * s) break;
    }
    double P = exp(-x) * s;
    return -log1p(-P);
  }
  return x - log(1.0 + x + 0.5 * x * x);
}

static double mj_tab[1000];
static int mj_tab_ready = 0;
const double *or_mj_table(void) {
  if (!mj_tab_ready) {
    for (int i = 0; i < 1000; i++) {
      double y = -30.0 + i * 42.0 / 999.0;
      double target = exp(y);
      double lo = -40.0, hi = 20.0;
      for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        if (mj_F(exp(mid)) < target) lo = mid; else hi = mid;
      }
      mj_tab[i] = 0.5 * (lo + hi);
    }
    mj_tab_ready = 1;
  }
  return mj_tab;
}

/* maxwellJuttner TESTMODE draws, relativistic branch only
 * (src/math/maxwellJuttner.cpp:358-396); the classical branch (T < 0.1) is
 * not needed by the smoke configuration and reports OR_ERR_BAD_ARG. */
static int mj_energies(int npoints, double temperature, double *energies) {
  if (temperature < 0.1) return OR_ERR_BAD_ARG;
  const double *lnInvH = or_mj_table();
  const double invdU_H = 999. / (12. + 30.);
  double invT = 1. / temperature;
  double H0 = -invT + log(1. + invT + 0.5 * invT * invT);
  for (unsigned int i = 0; i < (unsigned int)npoints; i++) {
    int k = 0;
    double U, lnU, invH, I, remainder, gamma;
    unsigned int index;
    do {
      U = or_random_reverse((double)((unsigned int)npoints + i - (unsigned int)k), 13);
      lnU = log(-log(1. - U) - H0);
      if (lnU < -26.)
        invH = pow(-6. * U, 1. / 3.);
      else if (lnU > 12.)
        invH = -U + 11.35 * pow(-U, 0.06);
      else {
        I = (lnU + 30.) * invdU_H;
        index = (unsigned int)I;
        remainder = I - (double)index;
        invH = exp(lnInvH[index] + remainder * (lnInvH[index + 1] - lnInvH[index]));
      }
      gamma = temperature * invH;
      U = or_random_reverse((double)((unsigned int)npoints - i + (unsigned int)k), 15);
      k++;
    } while (U >= sqrt(1. - 1. / (gamma * gamma)));
    energies[i] = gamma - 1.;
  }
  return OR_OK;
}

/* Particles::velocity_distribution with TESTMODE, src/particles.cpp:228-276 */
int or_velocity_distribution_testmode(int n, double mass_kg, double temperature,
                                      double *vr, double *vphi, double *vz) {
  double *energies = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  int st = mj_energies(n, temperature, energies);
  if (st) {
    free(energies);
    return st;
  }
  for (int p = 0; p < n; p++) {
    double therm_vel_el = sqrt(2 * EL_CHARGE * energies[p] / mass_kg);
    double norm = 0.74;
    therm_vel_el *= norm;
    double rnd_0 = or_random_reverse(p, 3);
    double rnd_1 = or_random_reverse(p, 5);
    double rnd_2 = or_random_reverse(p, 7);
    double k0 = or_random_reverse(p, 4);
    double k1 = or_random_reverse(p, 6);
    double k2 = or_random_reverse(p, 8);
    if (k0 > 0.5) rnd_0 = -rnd_0;
    if (k1 > 0.5) rnd_1 = -rnd_1;
    if (k2 > 0.5) rnd_2 = -rnd_2;
    double v0 = rnd_0 * therm_vel_el;
    double v1 = rnd_1 * therm_vel_el;
    double v2 = rnd_2 * therm_vel_el;
    vr[p] = v0 * get_gamma_inv(v0 * v0);
    vphi[p] = v1 * get_gamma_inv(v1 * v1);
    vz[p] = v2 * get_gamma_inv(v2 * v2);
  }
  free(energies);
  return OR_OK;
}

/* Bunch::Bunch with TESTMODE, src/bunch.cpp:6-80 */
void or_bunch_ctor_testmode(int n, double radius, double dr, double dz,
                            double duration, double velocity, double density,
                            double *r, double *w) {
  double b_lenght = duration * velocity;
  double v_total = 0.;
  double N_total = PI * pow(radius, 2) * b_lenght * density;
  for (int i = 0; i < n; i++) {
    double rand_r = or_random_reverse(i, 9);
    r[i] = (radius - dr) * rand_r + dr / 2;
    if (r[i] > dr)
      v_total += 2 * PI * r[i] * dr * dz;
    else
      v_total += PI * r[i] * r[i] * dz;
  }
  double v_avg = v_total / n;
  double n_per_macro_avg = N_total / n;
  for (int i = 0; i < n; i++) {
    double norm;
    if (r[i] > dr / 2)
      norm = 2 * PI * r[i] * dr * dz / v_avg;
    else
      norm = PI * (r[i] * r[i] + r[i] * dr + dr * dr / 4) * dz / v_avg;
    w[i] = n_per_macro_avg * norm;
  }
}
