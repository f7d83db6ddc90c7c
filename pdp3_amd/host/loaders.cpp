// Compiled with -ffp-contract=off: every product and sum rounds separately,
// as in the reference's IEEE build.
#include "loaders.hpp"

#include <cmath>

namespace pdp3 {
namespace loaders {

namespace {
// include/constant.h:10-18
constexpr double kPi = 3.1415926535897932;
constexpr double kElCharge = 1.6e-19;
constexpr double kLightSpeed2 = 3.0e8 * 3.0e8;

// lib::get_gamma_inv, src/lib.cpp:72-80
double gamma_inv(double sq_velocity) { return std::pow(1.0 + sq_velocity / kLightSpeed2, -0.5); }

// -log Q(3, x) = x - log(1 + x + x^2/2); for x < 1 through the lower
// series P(3, x) = e^-x sum_{n>=3} x^n / n! to keep digits
double mj_neg_log_q(double x) {
  if (x < 1.0) {
    double term = x * x * x / 6.0, sum = 0.0;
    for (int n = 4;; n++) {
      sum += term;
      term *= x / n;
      if (term < 1e-20 * sum) break;
    }
    return -std::log1p(-std::exp(-x) * sum);
  }
  return x - std::log(1.0 + x + 0.5 * x * x);
}

// -log Q(3/2, x); lower series P(3/2, x) = x^{3/2} e^-x / Gamma(5/2) *
// sum_n x^n / ((5/2)(7/2)...(3/2+n)) below x = 3, the closed form
// Q(3/2, x) = erfc(sqrt x) + 2 sqrt(x/pi) e^-x above
double mb_neg_log_q(double x) {
  if (x < 3.0) {
    double term = 1.0, sum = 0.0;
    for (int n = 1;; n++) {
      sum += term;
      term *= x / (1.5 + n);
      if (term < 1e-20 * sum) break;
    }
    const double gamma52 = 0.75 * std::sqrt(kPi);
    double p = std::pow(x, 1.5) * std::exp(-x) / gamma52 * sum;
    return -std::log1p(-p);
  }
  return -std::log(std::erfc(std::sqrt(x)) + 2.0 * std::sqrt(x / kPi) * std::exp(-x));
}

// log x solving f(x) = exp(y) for an increasing f, by bisection on log x
template <class F>
double solve_log(F f, double y) {
  const double target = std::exp(y);
  double lo = -40.0, hi = 20.0;
  for (int it = 0; it < 200; it++) {
    double mid = 0.5 * (lo + hi);
    if (f(std::exp(mid)) < target)
      lo = mid;
    else
      hi = mid;
  }
  return 0.5 * (lo + hi);
}

struct Tables {
  double mb[1000], mj[1000];
  Tables() {
    for (int i = 0; i < 1000; i++) {
      mb[i] = solve_log(mb_neg_log_q, -19.0 + i * 21.0 / 999.0);
      mj[i] = solve_log(mj_neg_log_q, -30.0 + i * 42.0 / 999.0);
    }
  }
};
const Tables &tables() {
  static const Tables t;
  return t;
}
}  // namespace

const double *mb_table() { return tables().mb; }
const double *mj_table() { return tables().mj; }

double random_reverse(double vel, int power) {
  int digits = (int)std::floor(vel);
  double result = 0.0;
  int order = 1;
  while (digits >= 1) {
    double digit = digits % power;
    result = result + digit * std::pow((double)power, -order);
    digits = (int)((digits - digit) / power);
    order++;
  }
  return result;
}

void load_cylindrical(int n, double r_size, double z_size, double dr, double dz, double n1, double n2,
                      double left_plasma_boundary, double *r, double *z, double *w) {
  const double dn = n2 - n1;
  double volume_sum = 0.0;
  for (int j = 0; j < n; j++) {
    double ur = random_reverse(j, 13);
    double uz = random_reverse(n - 1 - j, 11);
    r[j] = (r_size - dr) * ur + dr / 2;
    if (dn == 0) {
      z[j] = uz * (z_size - dz) + left_plasma_boundary + dz / 2;
    } else {
      // inverse CDF of the linear density profile
      z[j] = (z_size - left_plasma_boundary - dz) / dn *
                 (std::sqrt(std::pow(n1, 2) + uz * (2 * n1 * dn + std::pow(dn, 2))) - n1) +
             left_plasma_boundary + dz / 2;
    }
    volume_sum += 2 * kPi * r[j] * dr * dz;
  }
  const double volume_avg = volume_sum / n;
  const double total_particles = (n1 + n2) / 2 * kPi * r_size * r_size * z_size;
  const double per_macro_avg = total_particles / n;
  for (int j = 0; j < n; j++) {
    double norm;
    if (r[j] > dr / 2)
      norm = 2 * kPi * r[j] * dr * dz / volume_avg;
    else
      norm = kPi * (r[j] * r[j] + r[j] * dr + dr * dr / 4) * dz / volume_avg;
    w[j] = per_macro_avg * norm;
  }
}

void maxwell_juttner(int n, double temperature, double *energies) {
  const unsigned int npoints = (unsigned int)n;
  if (temperature < 0.1) {
    // Maxwell-Boltzmann: invert Q(3/2, x) = U
    const double *ln_inv_f = mb_table();
    const double inv_step = 999. / (2. + 19.);
    for (unsigned int i = 0; i < npoints; i++) {
      double u = random_reverse(npoints - i, 9 + i);
      double lnln_u = std::log(-std::log(u));
      double inv_f;
      if (lnln_u > 2.) {
        inv_f = 3. * std::sqrt(M_PI) / 4. * std::pow(u, 2. / 3.);
      } else if (lnln_u < -19.) {
        inv_f = 1.;
      } else {
        double pos = (lnln_u + 19.) * inv_step;
        unsigned int idx = (unsigned int)pos;
        double frac = pos - (double)idx;
        inv_f = std::exp(ln_inv_f[idx] + frac * (ln_inv_f[idx + 1] - ln_inv_f[idx]));
      }
      energies[i] = temperature * inv_f;
    }
    return;
  }
  // Maxwell-Juttner: invert H, then accept with probability beta
  const double *ln_inv_h = mj_table();
  const double inv_step = 999. / (12. + 30.);
  const double inv_t = 1. / temperature;
  const double h0 = -inv_t + std::log(1. + inv_t + 0.5 * inv_t * inv_t);
  for (unsigned int i = 0; i < npoints; i++) {
    int k = 0;
    double u, gamma;
    do {
      u = random_reverse(npoints + i - (unsigned int)k, 13);
      double ln_u = std::log(-std::log(1. - u) - h0);
      double inv_h;
      if (ln_u < -26.) {
        inv_h = std::pow(-6. * u, 1. / 3.);
      } else if (ln_u > 12.) {
        inv_h = -u + 11.35 * std::pow(-u, 0.06);
      } else {
        double pos = (ln_u + 30.) * inv_step;
        unsigned int idx = (unsigned int)pos;
        double frac = pos - (double)idx;
        inv_h = std::exp(ln_inv_h[idx] + frac * (ln_inv_h[idx + 1] - ln_inv_h[idx]));
      }
      gamma = temperature * inv_h;
      u = random_reverse(npoints - i + (unsigned int)k, 15);
      k++;
    } while (u >= std::sqrt(1. - 1. / (gamma * gamma)));
    energies[i] = gamma - 1.;
  }
}

void velocity_distribution(int n, double mass_kg, double temperature_ev, double *vr, double *vphi, double *vz) {
  std::vector<double> energies((size_t)(n > 0 ? n : 1));
  maxwell_juttner(n, temperature_ev, energies.data());
  for (int p = 0; p < n; p++) {
    double speed = std::sqrt(2 * kElCharge * energies[p] / mass_kg);
    speed *= 0.74;
    double a0 = random_reverse(p, 3), a1 = random_reverse(p, 5), a2 = random_reverse(p, 7);
    // signs from three more radical inverses
    if (random_reverse(p, 4) > 0.5) a0 = -a0;
    if (random_reverse(p, 6) > 0.5) a1 = -a1;
    if (random_reverse(p, 8) > 0.5) a2 = -a2;
    double u0 = a0 * speed, u1 = a1 * speed, u2 = a2 * speed;
    vr[p] = u0 * gamma_inv(u0 * u0);
    vphi[p] = u1 * gamma_inv(u1 * u1);
    vz[p] = u2 * gamma_inv(u2 * u2);
  }
}

void bunch_positions(int n, double radius, double dr, double dz, double duration, double velocity, double density,
                     double *r, double *w) {
  const double length = duration * velocity;
  const double total_particles = kPi * std::pow(radius, 2) * length * density;
  double volume_total = 0.;
  for (int i = 0; i < n; i++) {
    r[i] = (radius - dr) * random_reverse(i, 9) + dr / 2;
    if (r[i] > dr)
      volume_total += 2 * kPi * r[i] * dr * dz;
    else
      volume_total += kPi * r[i] * r[i] * dz;
  }
  const double volume_avg = volume_total / n;
  const double per_macro_avg = total_particles / n;
  for (int i = 0; i < n; i++) {
    double norm;
    if (r[i] > dr / 2)
      norm = 2 * kPi * r[i] * dr * dz / volume_avg;
    else
      norm = kPi * (r[i] * r[i] + r[i] * dr + dr * dr / 4) * dz / volume_avg;
    w[i] = per_macro_avg * norm;
  }
}

void bunch_inject_draws(int n, double *u) {
  for (int j = 0; j < n; j++) u[j] = random_reverse(j, 11);
}

}  // namespace loaders
}  // namespace pdp3
