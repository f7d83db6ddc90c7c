// Initial particle loaders of the reference's TESTMODE build (reproducible
// low-discrepancy draws instead of a random generator), restated for the
// drop-in driver.  All arithmetic is FP64 in the reference's order
// (compiled without FMA contraction), so a given configuration reproduces
// the reference's initial state.
#pragma once

#include <vector>

namespace pdp3 {
namespace loaders {

// math::random::random_reverse, src/math/rand.cpp:37-53 (radical inverse of
// floor(vel) in base `power`)
double random_reverse(double vel, int power);

// Particles::load_cylindrical_spatial_distribution, src/particles.cpp:278-331:
// r, z and w (real particles per macro-particle) of n particles between the
// densities n1 (z = 0) and n2 (z = z_size)
void load_cylindrical(int n, double r_size, double z_size, double dr, double dz, double n1, double n2,
                      double left_plasma_boundary, double *r, double *z, double *w);

// math::maxwell_juttner::maxwellJuttner with TESTMODE draws,
// src/math/maxwellJuttner.cpp:314-418: kinetic energies (in units of the
// temperature scale used by the reference) of n particles; Maxwell-Boltzmann
// below 0.1 eV, Maxwell-Juttner above
void maxwell_juttner(int n, double temperature_ev, double *energies);

// Particles::velocity_distribution with TESTMODE, src/particles.cpp:228-276
void velocity_distribution(int n, double mass_kg, double temperature_ev, double *vr, double *vphi, double *vz);

// Bunch::Bunch radial positions and weights with TESTMODE, src/bunch.cpp:6-80
void bunch_positions(int n, double radius, double dr, double dz, double duration, double velocity, double density,
                     double *r, double *w);

// the z draws of Bunch::bunch_inject with TESTMODE, src/bunch.cpp:106:
// particle j gets z = velocity * dt * u[j] + dz / 2 when injected
void bunch_inject_draws(int n, double *u);

// the two inverse-CDF tables of the sampler (1000 entries each), computed from
// their closed forms: log x with -log Q(3/2, x) = exp(-19 + 21 i / 999)
// (Maxwell-Boltzmann) and log x with x - log(1 + x + x^2/2) = exp(-30 + 42 i / 999)
// (Maxwell-Juttner), Q the regularised upper incomplete gamma function
const double *mb_table();
const double *mj_table();

}  // namespace loaders
}  // namespace pdp3
