/*
 * pdp3_oracle.h -- CPU oracle for the PDP3 r-z particle-in-cell step loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a clean-room FP64 restatement of the
 * reference algorithm (cosmonaut-ok/pdp3), written in plain C so that it can be
 * compiled with gcc and driven from Python through ctypes.  It is the parity
 * checker for the HIP product library and the CPU baseline timed by bench.py.
 * Nothing in the product library links, loads or calls it.
 *
 * Layout: every grid array is row-major FP64, element (i, k) at i * n_z + k,
 * i = radial node (0..n_grid_r-1), k = axial node (0..n_grid_z-1), matching
 * the reference's `double **a; a[i][k]` (include/field.h:8-10).
 */
#ifndef PDP3_ORACLE_H
#define PDP3_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference exit()s instead: src/particles.cpp:1046-1053,
 * src/lib.cpp:54-59) */
#define OR_OK 0
#define OR_ERR_NONFINITE (-4)
#define OR_ERR_SUPERLUMINAL (-5)
#define OR_ERR_BAD_ARG (-1)
#define OR_ERR_OOM (-3)

/* pushers (configure.ac:142-154; "boris" = PUSHER_BORIS_RELATIVISTIC, which is
 * also what the default adaptive build compiles to) */
#define OR_PUSHER_BORIS 0
#define OR_PUSHER_VAY 1

/* grid components for or_sim_field() */
enum {
  OR_E_R = 0, OR_E_PHI, OR_E_Z,
  OR_H_R, OR_H_PHI, OR_H_Z,
  OR_HH_R, OR_HH_PHI, OR_HH_Z,      /* H "half time" arrays */
  OR_J_R, OR_J_PHI, OR_J_Z,
  OR_RHO_BUNCH,
  OR_SIGMA, OR_EPSILON,
  OR_NCOMP
};

typedef struct or_sim or_sim;

/* ---- simulation object (geometry + fields + species list + time) ---- */
or_sim *or_sim_new(int n_grid_r, int n_grid_z, double r_size, double z_size);
void or_sim_free(or_sim *s);
void or_sim_set_time(or_sim *s, double current_time, double start_time,
                     double relaxation_time, double end_time, double delta_t);
void or_sim_set_pml(or_sim *s, double l1, double l2, double l3,
                    double sigma1, double sigma2);
void or_sim_set_pusher(or_sim *s, int pusher);
double or_sim_time(const or_sim *s);
double or_sim_dr(const or_sim *s);
double or_sim_dz(const or_sim *s);
double *or_sim_field(or_sim *s, int comp);
/* reference-faithful OpenMP split: push / move loops parallel, deposits serial.
 * threads <= 1 keeps everything serial. */
void or_sim_set_threads(or_sim *s, int threads);
/* test hook: run the serial deposit loops in reverse particle order, to
 * measure how much the loop's result depends on the order of the grid sums */
void or_sim_set_reverse_deposit(or_sim *s, int reverse);
/* test hook, process-wide: 1 takes lib::get_gamma_inv as the double nearest
 * to 1/sqrt(y) instead of glibc's pow(y, -0.5) (they differ in the last bit
 * for a few hundredths of a percent of arguments) */
void or_set_gamma_inv_nearest(int on);

/* plasma species (specular walls).  charge_e / mass_e in units of the
 * electron charge / mass as in parameters.xml (src/particles.cpp:30-33).
 * w = real particles per macro-particle; charge_array = charge * w. */
int or_sim_add_plasma(or_sim *s, int charge_e, double mass_e, int n,
                      const double *r, const double *z, const double *vr,
                      const double *vphi, const double *vz, const double *w,
                      const unsigned char *alive);
/* bunch (absorbing at r_max / z_max, injected over time, src/bunch.cpp).
 * r, w from the bunch constructor; inject_u[i] in [0,1) is the uniform draw
 * used for particle i's injection z (src/bunch.cpp:106-117).  The bunch is
 * created in the species list at the first step with
 * t >= bunch_number * (duration + hole_duration) (src/loadInitParam.cpp:219-243). */
int or_sim_add_bunch_loaded(or_sim *s, int charge_e, double mass_e, int n,
                            const double *r, const double *z, const double *vr,
                            const double *vphi, const double *vz, const double *w);
int or_sim_add_bunch(or_sim *s, int charge_e, double mass_e, int n,
                     const double *r, const double *w, const double *inject_u,
                     double duration, double hole_duration, double velocity,
                     int bunch_number);
int or_sim_num_species(const or_sim *s);
int or_sim_species_count(const or_sim *s, int idx);
/* copies the species state out (any pointer may be NULL) */
void or_sim_get_species(const or_sim *s, int idx, double *r, double *z,
                        double *vr, double *vphi, double *vz, double *q,
                        double *m, unsigned char *alive);
void or_sim_set_species(or_sim *s, int idx, const double *r, const double *z,
                        const double *vr, const double *vphi, const double *vz,
                        const unsigned char *alive);

/* one full loop iteration each (src/loadInitParam.cpp:471-587): inject beam,
 * H, reset J/rho, push all species, deposit all species, E, bunch rho, t+=dt */
int or_sim_step(or_sim *s, int nsteps);

/* the split entry points the reference exposes per class */
void or_sim_h_calc(or_sim *s);                       /* HField::calc_field */
void or_sim_e_calc(or_sim *s);                       /* EField::calc_field */
void or_sim_e_boundary_conditions(or_sim *s);        /* EField::boundary_conditions */
void or_sim_reset_j(or_sim *s);                      /* Current::reset_j */
int or_sim_step_v(or_sim *s, int idx);               /* Particles::step_v */
int or_sim_full_current_distribution(or_sim *s, int idx);
int or_sim_charge_weighting(or_sim *s, int idx, double *rho);
/* TemperatureClassic::calc_t + normalize (src/temperatureClassic.cpp:57-158)
 * of species idx: nr x nz temperatures in eV (row nr-1, column nz-1: 0) */
int or_temperature(const or_sim *s, int idx, double *out);
/* ParticlesDensity::calc_density (src/particlesDensity.cpp:40-225) of
 * species idx, added into out (nr x nz) */
int or_density(const or_sim *s, int idx, double *out);
void or_sim_bunch_inject(or_sim *s, int idx);

/* point evaluations */
void or_sim_get_e(const or_sim *s, double r, double z, double out[3]);
void or_sim_get_h(const or_sim *s, double r, double z, double out[3]);
/* single-particle velocity update with given (already gathered) E, H */
int or_boris_velocity(double q, double m, double dt, const double e[3],
                      const double h[3], double v[3]);
int or_vay_velocity(double q, double m, double dt, const double e[3],
                    const double h[3], double v[3]);
/* one particle's charge-conserving J_r/J_z deposit from old to new position
 * (src/particles.cpp:476-686) into arrays jr, jz of the sim's shape */
void or_current_deposit(const or_sim *s, double q, double r_new, double z_new,
                        double r_old, double z_old, double *jr, double *jz);
void or_azimuthal_deposit(const or_sim *s, double q, double r, double z,
                          double vphi, double *jphi);
void or_charge_deposit(const or_sim *s, double q, double r, double z,
                       double *rho);
int or_cell_number(double x, double dx);

/* ---- initial loaders used only to replay the reference's smoke goldens ---- */
double or_random_reverse(double vel, int power);
/* src/particles.cpp:278-331 */
void or_load_cylindrical(int n, double r_size, double z_size, double dr,
                         double dz, double n1, double n2,
                         double left_plasma_boundary, double *r, double *z,
                         double *w);
/* src/particles.cpp:228-276 with the TESTMODE draws and the relativistic
 * Maxwell-Juttner branch (src/math/maxwellJuttner.cpp:314-399) */
int or_velocity_distribution_testmode(int n, double mass_kg,
                                      double temperature, double *vr,
                                      double *vphi, double *vz);
/* src/bunch.cpp:6-80 with TESTMODE radial draws */
void or_bunch_ctor_testmode(int n, double radius, double dr, double dz,
                            double duration, double velocity, double density,
                            double *r, double *w);
/* 1000-entry log-inverse table of H(x) = log Q(3,x) on [-30,12] */
const double *or_mj_table(void);

#ifdef __cplusplus
}
#endif
#endif
