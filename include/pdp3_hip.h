/*
 * pdp3_hip.h -- C ABI of the MI355X (gfx950) implementation of PDP3's r-z
 * particle-in-cell step loop.
 *
 * One opaque context owns the geometry, the Yee grid (E, H, H at half time,
 * J, rho of the bunches, PML conductivity), the species list and the model
 * time, all resident in HBM, plus one HIP stream on which every call is
 * enqueued.  Host arrays passed in are copied (borrowed for the call only);
 * nothing but the handle crosses the boundary.  Every function returns 0 on
 * success or a negative PDP3_ERR_* code; pdp3_last_error() gives the text.
 * The library never exits the process (the reference exit()s or abort()s).
 *
 * Each entry point names the reference interface it replaces
 * (cosmonaut-ok/pdp3, path:line).  Grid arrays are row-major FP64 with element
 * (i, k) at i * n_grid_z + k, i radial, k axial, as the reference's
 * `double **a; a[i][k]` (include/field.h:8-10).
 *
 * Not re-entrant per context: one host thread drives one context.
 */
#ifndef PDP3_HIP_H
#define PDP3_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define PDP3_OK 0
#define PDP3_ERR_BAD_ARG (-1)
#define PDP3_ERR_HIP (-2)
#define PDP3_ERR_OOM (-3)
#define PDP3_ERR_NONFINITE (-4)    /* NaN/inf particle position (src/particles.cpp:1046-1053) */
#define PDP3_ERR_SUPERLUMINAL (-5) /* |v| > c in the pusher (src/lib.cpp:54-59) */
#define PDP3_ERR_STATE (-6)

/* ---- precision forms ---- */
#define PDP3_PREC_FP64 0 /* plain form: FP64 throughout, reference arithmetic order */
#define PDP3_PREC_FP32 1 /* FP32 push/deposit on int cell + FP32 offset, FP64 grid */

/* ---- pushers (compile-time in the reference, configure.ac:142-154) ---- */
#define PDP3_PUSHER_BORIS 0 /* PUSHER_BORIS_RELATIVISTIC (the default adaptive build compiles to it) */
#define PDP3_PUSHER_VAY 1   /* PUSHER_VAY */

/* ---- species kinds ---- */
#define PDP3_SPECIES_PLASMA 0 /* Particles: specular walls (src/particles.cpp:1318-1353) */
#define PDP3_SPECIES_BUNCH 1  /* Bunch: absorbed at r_max / z_max (src/bunch.cpp:126-153) */

/* ---- grid components ---- */
#define PDP3_E_R 0
#define PDP3_E_PHI 1
#define PDP3_E_Z 2
#define PDP3_H_R 3
#define PDP3_H_PHI 4
#define PDP3_H_Z 5
#define PDP3_HH_R 6 /* HField::field_r_half_time (include/hField.h:15-17) */
#define PDP3_HH_PHI 7
#define PDP3_HH_Z 8
#define PDP3_J_R 9
#define PDP3_J_PHI 10
#define PDP3_J_Z 11
#define PDP3_RHO_BUNCH 12 /* LoadInitParam::c_rho_bunch */
#define PDP3_SIGMA 13     /* Geometry::sigma */
#define PDP3_EPSILON 14   /* Geometry::epsilon */
#define PDP3_NCOMP 15

/* ---- kernels with live timing (pdp3_profile_get) ---- */
#define PDP3_KERNEL_PARTICLES 0 /* fused push + deposit, plasma species */
#define PDP3_KERNEL_H 1
#define PDP3_KERNEL_E 2
#define PDP3_KERNEL_SORT 3
#define PDP3_KERNEL_BUNCH 4 /* fused push + deposit + charge weighting, bunches */
#define PDP3_NKERNELS 5

typedef struct pdp3_ctx pdp3_ctx;

/* Geometry(double rs, double zs, int ngr, int ngz) + Time(ct, st, rt, et, dt)
 * (include/geometry.h:20-24, src/geometry.cpp:8-31, src/pdp3Time.cpp:11-18).
 * epsilon is set to 1 (Geometry::set_epsilon, src/geometry.cpp:33-39), all
 * fields and sigma to 0. */
int pdp3_ctx_create(int n_grid_r, int n_grid_z, double r_size, double z_size,
                    double current_time, double start_time, double relaxation_time,
                    double end_time, double delta_t, int pusher, int precision,
                    pdp3_ctx **out);
int pdp3_ctx_destroy(pdp3_ctx *ctx);
const char *pdp3_last_error(const pdp3_ctx *ctx);
/* geometry and time read-back: out[0..5] = dr, dz, r_size, z_size, n_grid_r, n_grid_z */
int pdp3_geometry_get(const pdp3_ctx *ctx, double out[6]);
/* Time members (include/pdp3Time.h:6-11) */
int pdp3_time_get(const pdp3_ctx *ctx, double *current_time, double *delta_t);
int pdp3_time_set(pdp3_ctx *ctx, double current_time);

/* Geometry::set_pml(l1, l2, l3, sigma1, sigma2), src/geometry.cpp:56-107 */
int pdp3_geometry_set_pml(pdp3_ctx *ctx, double l1, double l2, double l3,
                          double sigma1, double sigma2);

/* EField::boundary_conditions, src/eField.cpp:46-82 */
int pdp3_efield_boundary_conditions(pdp3_ctx *ctx);
/* EField::set_homogeneous_efield, src/eField.cpp:33-44 */
int pdp3_efield_set_homogeneous(pdp3_ctx *ctx, double e_r, double e_phi, double e_z);
/* HField::set_homogeneous_h, src/hField.cpp:66-91 */
int pdp3_hfield_set_homogeneous(pdp3_ctx *ctx, double h_r, double h_phi, double h_z);
/* BoundaryMaxwellConditions::specify_initial_field, src/boundaryMaxwellConditions.cpp:17-38 */
int pdp3_boundary_specify_initial_field(pdp3_ctx *ctx, double e_phi_upper,
                                        double e_phi_left, double e_phi_right);

/* HField::calc_field(EField*, Time*), src/hField.cpp:93-157 */
int pdp3_hfield_calc(pdp3_ctx *ctx);
/* EField::calc_field(HField*, Time*, Current*), src/eField.cpp:85-170 */
int pdp3_efield_calc(pdp3_ctx *ctx);
/* Current::reset_j, src/current.cpp:125-135 */
int pdp3_current_reset(pdp3_ctx *ctx);
/* ChargeDensity::reset_rho on the bunch density, src/chargeDensity.cpp:43-49 */
int pdp3_rho_bunch_reset(pdp3_ctx *ctx);

/* EField::get_field / HField::get_field at n points (src/eField.cpp:204-316,
 * src/hField.cpp:159-267): out is n x 3 (r, phi, z components) */
int pdp3_efield_get_field(pdp3_ctx *ctx, const double *r, const double *z, int n, double *out);
int pdp3_hfield_get_field(pdp3_ctx *ctx, const double *r, const double *z, int n, double *out);

/* grid array transfer (the reference's public double** members,
 * include/field.h:8-10, include/current.h:23-25, include/chargeDensity.h:17,
 * include/geometry.h:14-15).  n_grid_r * n_grid_z doubles. */
int pdp3_grid_download(pdp3_ctx *ctx, int component, double *out);
int pdp3_grid_upload(pdp3_ctx *ctx, int component, const double *in);

/* Particles(name, charge, mass, number, geom), src/particles.cpp:23-55:
 * charge_e / mass_e are in electron charges / masses (parameters.xml).  The
 * state starts dead; upload it with pdp3_species_upload.  Species are kept in
 * ParticlesList order (include/particlesList.h:14): plasma species in the
 * order they are added. */
int pdp3_species_add(pdp3_ctx *ctx, int charge_e, double mass_e, int number, int *out_index);
/* Bunch(...) ctor + LoadInitParam::init_beam bookkeeping, src/bunch.cpp:6-80,
 * src/loadInitParam.cpp:168-248.  r and w as the constructor leaves them
 * (z = v = 0, dead); inject_u[i] in [0,1) is the uniform draw that places
 * particle i at injection (src/bunch.cpp:106-117).  The bunch joins the
 * species list at the first step with t >= bunch_number*(duration+hole). */
int pdp3_bunch_add(pdp3_ctx *ctx, int charge_e, double mass_e, int number,
                   const double *r, const double *w, const double *inject_u,
                   double duration, double hole_duration, double velocity,
                   int bunch_number, int *out_index);
/* a bunch that is already inside the grid (no reference counterpart: the
 * reference's bunches enter by injection only): alive particles at the given
 * positions and velocities, absorbed at r_max / z_max like Bunch
 * (src/bunch.cpp:126-153), charge-weighted into the bunch density every
 * step; it joins the species list at once. */
int pdp3_bunch_add_loaded(pdp3_ctx *ctx, int charge_e, double mass_e, int number, const double *r,
                          const double *z, const double *vr, const double *vphi, const double *vz,
                          const double *w, int *out_index);
/* upload / download of a species' state.  w = real particles per
 * macro-particle (charge_array = charge * w, src/particles.cpp:328-329);
 * alive as Particles::is_alive.  Any pointer may be NULL (kept / skipped). */
int pdp3_species_upload(pdp3_ctx *ctx, int index, const double *r, const double *z,
                        const double *vr, const double *vphi, const double *vz,
                        const double *w, const uint8_t *alive);
int pdp3_species_download(pdp3_ctx *ctx, int index, double *r, double *z, double *vr,
                          double *vphi, double *vz, double *w, uint8_t *alive);
int pdp3_species_count(const pdp3_ctx *ctx, int index, int *number);
/* synthetic uniform-in-cell plasma with an isotropic Maxwellian (counter-
 * based Philox-4x32-10 draws keyed on (seed, species, cell, slot), so any
 * device regenerates the same state).  Density is n0 for z-cells >= ramp_cells
 * and ramps linearly from 0 at z = 0 before that (ramp_cells = 0: uniform).
 * The species must have number == ppc * (n_grid_r-1) * (n_grid_z-1). */
int pdp3_species_init_uniform(pdp3_ctx *ctx, int index, uint64_t seed, int ppc,
                              double density, int ramp_cells, double temperature_ev);

/* ParticlesList::step_v -> Particles::step_v for one species,
 * src/particlesList.cpp:20-24, src/particles.cpp:1296-1316 */
int pdp3_particles_step_v(pdp3_ctx *ctx, int index);
/* Particles::full_current_distribution, src/particles.cpp:1367-1373 */
int pdp3_particles_full_current_distribution(pdp3_ctx *ctx, int index);
/* Particles::charge_weighting(ChargeDensity*) into the bunch density,
 * src/particles.cpp:101-226 */
int pdp3_particles_charge_weighting(pdp3_ctx *ctx, int index);
/* Bunch::bunch_inject(Time*), src/bunch.cpp:87-124 */
int pdp3_bunch_inject(pdp3_ctx *ctx, int index);

/* nsteps iterations of LoadInitParam::run's loop body,
 * src/loadInitParam.cpp:471-587: init_beam, H, reset J and rho, push and
 * deposit of every species (fused, one pass over each species' particles),
 * E, charge weighting of the bunches, t += dt.  Asynchronous except for the
 * error check at the end. */
int pdp3_step(pdp3_ctx *ctx, int nsteps);
/* waits for the context's stream and reports device-side errors */
int pdp3_sync(pdp3_ctx *ctx);

/* stable sort of every plasma species by cell (no reference counterpart:
 * keeps gathers and deposits local; bunches keep their index order, which
 * their injection relies on).  Particle indices change: download after a sort
 * returns the sorted order.  sort_interval > 0 makes pdp3_step sort every
 * sort_interval steps; 0 disables.  The periodic sorts do not wait for the
 * device: their record space follows the previous sort's count of runs of
 * equal cells, and a sort that finds more runs than that leaves the order as
 * it is (the next one is sized for it).  Their buffers (a second copy of the
 * species' particle arrays and a workspace) are allocated by the first step
 * taken with sort_interval > 0; later steps allocate only when the record
 * space has to grow.
 * pdp3_sort_particles always sorts. */
int pdp3_sort_particles(pdp3_ctx *ctx);
int pdp3_set_sort_interval(pdp3_ctx *ctx, int sort_interval);

/* No counterpart in the reference (an analysis option of the FP64 plain
 * form): on != 0 evaluates the second piece of a strictly radial move that
 * crosses a radial face (strict_motion_weighting, src/particles.cpp:937-945
 * and :972-980) with r0*r0/(i_n*dr), the expression of its other pieces,
 * where the reference's `r0 * r0 / i_n * dr` parses as ((r0*r0)/i_n)*dr and
 * leaves a spurious J_r of order q r0 / (dt dz^2) at such an event.  Default
 * 0 (the reference's arithmetic).  The FP32 form's closed forms have no such
 * term either way. */
int pdp3_set_strict_radial_fix(pdp3_ctx *ctx, int on);

/* Gauss-law residual R = div E - rho/eps0 on the grid nodes (no reference
 * counterpart: the reference allocates c_rho_new for it but never fills it,
 * src/loadInitParam.cpp:147,534).  rho is the charge density of every living
 * particle of every species in the loop, weighted as
 * Particles::charge_weighting (src/particles.cpp:101-226); div E is the
 * discrete divergence that EField::calc_field (src/eField.cpp:85-170) leaves
 * invariant, so R changes only by round-off when the deposit conserves charge
 * (outside the PML conductivity layer and the bunch injection nodes).
 * residual_out and rho_out (either may be NULL) receive n_grid_r x n_grid_z
 * doubles; R is 0 at k = 0, k = n_grid_z-1 and i = n_grid_r-1. */
int pdp3_gauss_residual(pdp3_ctx *ctx, double *residual_out, double *rho_out);

/* pdp3_gauss_residual without allocating device memory.  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  rho and
 * the residual are computed in a fixed device array of the library
 * (2 x PDP3_NOALLOC_MAX_NODES doubles per GPU, held until the process ends)
 * by the same kernels; the outputs are as pdp3_gauss_residual's (host
 * arrays, either may be NULL).  A grid of more than PDP3_NOALLOC_MAX_NODES
 * nodes (n_grid_r * n_grid_z) is rejected with PDP3_ERR_BAD_ARG.  Every call
 * on one GPU uses that GPU's one array, which keeps the last call's values:
 * calls that overlap on one GPU (contexts driven from several host threads)
 * race. */
#define PDP3_NOALLOC_MAX_NODES 4718592 /* 4.5 Mi nodes, e.g. 513 x 8193 */
int pdp3_gauss_residual_noalloc(pdp3_ctx *ctx, double *residual_out, double *rho_out);

/* TemperatureClassic::calc_t + normalize of one species
 * (src/temperatureClassic.cpp:57-158, the "T" probe of LoadInitParam::dump_data,
 * src/loadInitParam.cpp:369-379), computed on the device from the resident
 * particle state: per cell (CELL_NUMBER of r and z) the sums of |v|, v_r,
 * v_phi, v_z and the count, then T = m v^2/2 / |charge| in eV with
 * v^2 = (S^2 - |V|^2) / N^2 (m c^2 (gamma-1) / |charge| above 5e7 m/s).
 * out: n_grid_r x n_grid_z; row n_grid_r-1 and column n_grid_z-1 are 0, an
 * empty cell is NaN (0/0), as in the reference. */
int pdp3_temperature(pdp3_ctx *ctx, int index, double *out);
/* ParticlesDensity::calc_density of one species (src/particlesDensity.cpp:40-225,
 * the "density" probe, src/loadInitParam.cpp:381-390) on the device: every
 * particle's mass_array / EL_MASS spread over its four nodes with the
 * reference's ring-volume weights.  out: n_grid_r x n_grid_z. */
int pdp3_density(pdp3_ctx *ctx, int index, double *out);

/* tile shape of the FP32 particle kernel (no reference counterpart):
 * particles walked per lane (ppt, 4 / 8 / 16) and wavefronts per block
 * (waves); supported pairs: (16,4) (16,2) (16,1) (8,4) (8,2) (4,4).  The
 * FP64 form has one shape.  pdp3_autotune runs steps_per_candidate real steps
 * of the simulation with each shape, times the plasma kernel and keeps the
 * fastest (the steps advance the simulation; results differ between shapes
 * only by the order of FP64 sums). */
int pdp3_set_tile(pdp3_ctx *ctx, int ppt, int waves);
int pdp3_get_tile(const pdp3_ctx *ctx, int *ppt, int *waves);
int pdp3_autotune(pdp3_ctx *ctx, int steps_per_candidate, int *ppt, int *waves);

/* ---- adjoint (no reference counterpart) ----
 * Runs nsteps loop iterations from the current state and returns the
 * objective
 *   phi = dt dz sum over the last window_steps iterations of
 *         sum over axis nodes z_lo <= (k + 1/2) dz <= z_hi of E_z(0, k)^2
 * (E after each iteration) and its gradient with respect to the weight of
 * every particle (real particles per macro-particle; charge_array and
 * mass_array of src/particles.cpp:328-329 scale with it): dphi_dw[j] receives
 * species j's n values in its logical particle order (a NULL entry skips the
 * species).  Hand-written adjoint kernels of the FDTD / PML update and of the
 * push + deposit (the FP64 plain form's arithmetic, linearised around the
 * context's own forward trajectory: FP32 or FP64 form), with checkpoints of
 * the full state every checkpoint_every steps and recomputation between
 * them.  The context ends in the forward end state.  Requires
 * sort_interval 0 (a fixed particle order). */
int pdp3_adjoint_run(pdp3_ctx *ctx, int nsteps, int checkpoint_every, int window_steps, double z_lo,
                     double z_hi, double *phi_out, double *const *dphi_dw);

/* Tangent-linear model of the same loop and objective (no reference
 * counterpart): runs nsteps iterations from the current state, like
 * pdp3_objective_run, and returns phi and its directional derivatives along
 * PDP3_TANGENT_DIRS directions of the particle weights.  dw[j] holds species
 * j's PDP3_TANGENT_DIRS x n weight tangents, direction-major, in its logical
 * particle order (a NULL entry: zero).  The tangents are carried forward
 * through the FP64 plain form's particle step in dual numbers and through the
 * field updates themselves, i.e. the linearisation that pdp3_adjoint_run
 * transposes (around the context's own trajectory, FP32 or FP64 form): for
 * every direction d, sum over particles of dphi_dw * dw[d] equals dphi_out[d]
 * to round-off (the adjoint's dot-product test).  Requires sort_interval 0. */
#define PDP3_TANGENT_DIRS 4
int pdp3_tangent_run(pdp3_ctx *ctx, int nsteps, int window_steps, double z_lo, double z_hi,
                     const double *const *dw, double *phi_out, double *dphi_out);
/* pdp3_adjoint_run with pdp3_tangent_run's tangents carried along the
 * adjoint's own forward sweep (no reference counterpart): the tangents follow
 * exactly the trajectory the checkpoints record, which a separate run cannot
 * (the grid sums use FP64 atomics, so two runs differ by round-off that the
 * loop amplifies).  With checkpoint_every 1 the adjoint linearises around the
 * same states, and the dot-product identity holds to the round-off of the
 * sums; with longer segments the recomputed steps between checkpoints
 * deviate by amplified round-off. */
int pdp3_adjoint_tangent_run(pdp3_ctx *ctx, int nsteps, int checkpoint_every, int window_steps, double z_lo,
                             double z_hi, const double *const *dw, double *phi_out, double *const *dphi_dw,
                             double *dphi_tangent);

/* pdp3_adjoint_run over a batch of independent members (no reference
 * counterpart; e.g. the same problem at n parameter points): each member's
 * context is driven by its own host thread on its own stream, so the
 * members' forward sweeps, recomputations and adjoint kernels interleave on
 * the device.  Member m writes phi_out[m] and dphi_dw[m][j] (species j, as
 * pdp3_adjoint_run; dphi_dw or dphi_dw[m] may be NULL); status_out[m]
 * (may be NULL) receives its status, the return value is the first nonzero
 * one.  Every member holds its own checkpoints at once; a context may appear
 * only once. */
int pdp3_adjoint_run_batch(pdp3_ctx *const *ctxs, int n, int nsteps, int checkpoint_every, int window_steps,
                           double z_lo, double z_hi, double *phi_out, double *const *const *dphi_dw,
                           int *status_out);

/* the same objective without the gradient (forward only) */
int pdp3_objective_run(pdp3_ctx *ctx, int nsteps, int window_steps, double z_lo, double z_hi, double *phi_out);

/* ---- ensembles (no reference counterpart) ----
 * Independent instances (contexts set up with the calls above: any grid,
 * PML, plasma, loaded bunches; FP32 form, one common delta_t) advanced
 * together: per step one graph replay with an H launch, a J / rho reset, ONE
 * particle launch over the plasma tiles and bunch blocks of every instance,
 * and an E launch.  The contexts stay owned by the caller and must not be
 * stepped on their own while the ensemble lives; their clocks advance with
 * the ensemble.  Bunches may be loaded (pdp3_bunch_add_loaded) or injected
 * (pdp3_bunch_add): the injections run inside the graph from a device step
 * counter. */
typedef struct pdp3_ensemble pdp3_ensemble;
int pdp3_ensemble_create(pdp3_ctx *const *ctxs, int n, pdp3_ensemble **out);
int pdp3_ensemble_destroy(pdp3_ensemble *e);
const char *pdp3_ensemble_last_error(const pdp3_ensemble *e);
int pdp3_ensemble_step(pdp3_ensemble *e, int nsteps);
/* cell sort of every instance every sort_interval steps (outside the graph;
 * the periodic sorts of pdp3_set_sort_interval, enqueued for all instances
 * before any is waited for; their buffers are allocated by the first
 * pdp3_ensemble_step call after it is set) */
int pdp3_ensemble_set_sort_interval(pdp3_ensemble *e, int sort_interval);
/* device time of the stepped graphs since the last call (HIP events), and the step count */
int pdp3_ensemble_timing(pdp3_ensemble *e, double *total_ms, long long *steps);
/* tile shape of the ensemble's particle kernel, as pdp3_set_tile; the
 * autotuner times steps_per_candidate real steps per shape and keeps the fastest */
int pdp3_ensemble_set_tile(pdp3_ensemble *e, int ppt, int waves);
int pdp3_ensemble_get_tile(const pdp3_ensemble *e, int *ppt, int *waves);
int pdp3_ensemble_autotune(pdp3_ensemble *e, int steps_per_candidate, int *ppt, int *waves);

/* live timing with HIP events on the context's stream (no reference
 * counterpart): enable, then read the accumulated device time in ms and the
 * launch count of one kernel class */
int pdp3_profile_enable(pdp3_ctx *ctx, int enable);
int pdp3_profile_get(pdp3_ctx *ctx, int kernel, double *total_ms, long long *launches);
int pdp3_profile_reset(pdp3_ctx *ctx);

/* library version string */
const char *pdp3_version(void);

#ifdef __cplusplus
}
#endif
#endif
