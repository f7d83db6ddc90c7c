// FP64 plain form of the per-particle work of one step: gather, Boris/Vay
// push, two half steps with the x-y -> r-z rotation, azimuthal and
// charge-conserving current deposit, wall handling, velocity rotation and the
// bunch charge weighting.  Every expression keeps the reference's evaluation
// order (and this file is built with -ffp-contract=off) so that per-particle
// results match the reference's IEEE build; only the order of the grid sums
// differs (hardware FP64 atomics instead of the reference's serial loop).
#include "context.hpp"
#include "crmath.hpp"
#include "sink.hpp"

namespace pdp3 {

namespace {

struct Fields {
  const double *er, *ephi, *ez, *hhr, *hhphi, *hhz;
};

struct SpeciesArgs {
  double *r, *z, *vr, *vphi, *vz;
  const double *w;
  uint8_t *alive;
  double *r_old, *z_old, *sin_t;
  double charge, mass;
  int n;
  int is_bunch;
};

// CELL_NUMBER, include/geometry.h:44: x in (i dx, (i+1) dx] -> i; x / dx
// correctly rounded from the context's 1/dr, 1/dz (div_cr)
__device__ __forceinline__ int cell_r(const GridDesc &g, double x) { return (int)ceil(div_cr(x, g.dr, g.inv_dr)) - 1; }
__device__ __forceinline__ int cell_z(const GridDesc &g, double x) { return (int)ceil(div_cr(x, g.dz, g.inv_dz)) - 1; }
// CELL_VOLUME(i, dr, dz) and CYL_VOL(dz, dr) (include/geometry.h:33-41) with
// their reciprocals, from the context's table (computed there in the
// reference's expression order)
struct Vol {
  double v, inv;
};
__device__ __forceinline__ Vol vol_cell(const GridDesc &g, int i) {
  const double2 t = g.vol[i];
  return Vol{t.x, t.y};
}
__device__ __forceinline__ Vol vol_cyl(const GridDesc &g) { return Vol{g.cyl, g.inv_cyl}; }
// `F * CYL_RNG_VOL(dzx, ra, rb) / vol` as the reference's macro expands
__device__ __forceinline__ double ring_term(double F, double dzx, double ra, double rb, Vol vol) {
  return div_cr(F * kPi * dzx * (rb * rb - ra * ra), vol.v, vol.inv);
}
// sq / c^2 correctly rounded (the divisor of lib::get_gamma / get_gamma_inv)
constexpr double kC2 = kLightSpeed * kLightSpeed;
constexpr double kInvC2 = 1.0 / kC2;

// four-node ring-volume gather shared by EField::get_field
// (src/eField.cpp:204-316) and HField::get_field (src/hField.cpp:159-267).
// lo: radial index from r - dr/2 (E_r, H_z, H_phi); zhalf: axial index from
// z - dz/2 (E_z, H_r, H_phi).  Indices are clamped at 0 as the reference does
// and, beyond the reference, kept inside the array (flagged as a domain error).
__device__ double gather4(const GridDesc &g, const double *__restrict__ F, double radius, double longitude,
                          bool lo, bool zhalf, int *err) {
  const double dr = g.dr, dz = g.dz;
  double r1 = radius - 0.5 * dr;
  double r3 = radius + 0.5 * dr;
  int i_r = lo ? cell_r(g, radius - 0.5 * dr) : cell_r(g, radius);
  int k_z = zhalf ? cell_z(g, longitude - 0.5 * dz) : cell_z(g, longitude);
  if (i_r < 0) i_r = 0;
  if (k_z < 0) k_z = 0;
  if (i_r > g.nr - 2 || k_z > g.nz - 2) {
    atomicOr(err, kErrDomain);
    i_r = min(i_r, g.nr - 2);
    k_z = min(k_z, g.nz - 2);
  }
  Vol vol_1, vol_2;
  double r2, dz1, dz2;
  if (lo) {
    vol_1 = vol_cell(g, i_r + 1);
    vol_2 = vol_cell(g, i_r + 3);
    r2 = (i_r + 1) * dr;
  } else {
    vol_1 = (radius > dr) ? vol_cell(g, i_r) : vol_cyl(g);
    r2 = (i_r + 0.5) * dr;
    vol_2 = vol_cell(g, i_r + 2);
  }
  if (zhalf) {
    dz1 = (k_z + 1.5) * dz - longitude;
    dz2 = longitude - (k_z + 0.5) * dz;
  } else {
    dz1 = (k_z + 1) * dz - longitude;
    dz2 = longitude - k_z * dz;
  }
  const double *row0 = F + (size_t)i_r * (size_t)g.nz;
  const double *row1 = row0 + g.nz;
  double acc = 0;
  acc = acc + ring_term(row0[k_z], dz1, r1, r2, vol_1);
  acc = acc + ring_term(row1[k_z], dz1, r2, r3, vol_2);
  acc = acc + ring_term(row0[k_z + 1], dz2, r1, r2, vol_1);
  acc = acc + ring_term(row1[k_z + 1], dz2, r2, r3, vol_2);
  return acc;
}

__device__ __forceinline__ void gather_e(const GridDesc &g, const Fields &f, double r, double z, double e[3],
                                         int *err) {
  e[0] = gather4(g, f.er, r, z, true, false, err);
  e[1] = gather4(g, f.ephi, r, z, false, false, err);
  e[2] = gather4(g, f.ez, r, z, false, true, err);
}
__device__ __forceinline__ void gather_h(const GridDesc &g, const Fields &f, double r, double z, double h[3],
                                         int *err) {
  h[0] = gather4(g, f.hhr, r, z, false, true, err);
  h[1] = gather4(g, f.hhphi, r, z, true, true, err);
  h[2] = gather4(g, f.hhz, r, z, true, false, err);
}

__device__ __forceinline__ double dot3(const double *a, const double *b) {
  double res = 0;
  res += a[0] * b[0];
  res += a[1] * b[1];
  res += a[2] * b[2];
  return res;
}

// lib::get_gamma / get_gamma_inv, src/lib.cpp:45-80
__device__ __forceinline__ bool gamma_of(double sq_velocity, double &gamma) {
  double beta = div_cr(sq_velocity, kC2, kInvC2);
  if (beta > 1) return false;
  double gm = 1 / sqrt(1.0 - beta);
  gamma = isinf(gm) ? 1e100 : gm;
  return true;
}
// pow(y, -0.5) as glibc returns it: the correctly rounded 1/sqrt(y)
__device__ __forceinline__ double gamma_inv(double sq_velocity) {
  return rsqrt_cr(1.0 + div_cr(sq_velocity, kC2, kInvC2));
}

// relativistic Boris, src/particles.cpp:1025-1150 (velocity part)
__device__ bool boris(double q, double m, double dt, double e[3], double b[3], double v[3]) {
  double const1 = q * dt / (2 * m);
  for (int j = 0; j < 3; j++) e[j] *= const1;
  double bs = kMagnConst * const1;
  for (int j = 0; j < 3; j++) b[j] *= bs;
  double u[3], t[3];
  for (int j = 0; j < 3; j++) u[j] = (fabs(v[j]) < kMnzl) ? 0 : v[j];
  double gamma;
  if (!gamma_of(dot3(u, u), gamma)) return false;
  for (int j = 0; j < 3; j++) u[j] *= gamma;
  for (int j = 0; j < 3; j++) u[j] += e[j];
  gamma = gamma_inv(dot3(u, u));
  for (int j = 0; j < 3; j++) b[j] *= gamma;
  double const2 = 2. / (1. + dot3(b, b));
  for (int j = 0; j < 3; j++) t[j] = u[j];
  u[0] = t[0] + const2 * ((t[1] - t[0] * b[2] + t[2] * b[0]) * b[2] - (t[2] + t[0] * b[1] - t[1] * b[0]) * b[1]);
  u[1] = t[1] + const2 * (-(t[0] + t[1] * b[2] - t[2] * b[1]) * b[2] + (t[2] + t[0] * b[1] - t[1] * b[0]) * b[0]);
  u[2] = t[2] + const2 * ((t[0] + t[1] * b[2] - t[2] * b[1]) * b[1] - (t[1] - t[0] * b[2] + t[2] * b[0]) * b[0]);
  for (int j = 0; j < 3; j++) u[j] += e[j];
  gamma = gamma_inv(dot3(u, u));
  for (int j = 0; j < 3; j++) v[j] = u[j] * gamma;
  return true;
}

// Vay, src/particles.cpp:1152-1254, with the reference's b scaling
// (mu0 c1, then 2 c1, then c1) and its us2 without 1/c^2
__device__ bool vay(double q, double m, double dt, double e[3], double b[3], double v[3]) {
  double const1 = q * dt / (2 * m);
  double s2 = 2 * const1;
  for (int j = 0; j < 3; j++) e[j] *= s2;
  double bs = kMagnConst * const1;
  for (int j = 0; j < 3; j++) b[j] *= bs;
  double u[3], t[3];
  for (int j = 0; j < 3; j++) u[j] = (fabs(v[j]) < kMnzl) ? 0 : v[j];
  double gamma;
  if (!gamma_of(dot3(u, u), gamma)) return false;
  for (int j = 0; j < 3; j++) u[j] *= gamma;
  for (int j = 0; j < 3; j++) t[j] = u[j];
  for (int j = 0; j < 3; j++) b[j] *= s2;
  for (int j = 0; j < 3; j++) u[j] += e[j];
  for (int j = 0; j < 3; j++) b[j] *= const1;
  gamma = gamma_inv(dot3(t, t));
  u[0] += gamma * (t[1] * b[2] - t[2] * b[1]);
  u[1] += gamma * (t[2] * b[0] - t[0] * b[2]);
  u[2] += gamma * (t[0] * b[1] - t[1] * b[0]);
  gamma = gamma_inv(dot3(u, u));
  double alpha = 1. / gamma / gamma;
  double B2 = dot3(b, b);
  double sg = alpha - B2;
  double ub = dot3(u, b);
  double us2 = ub * ub;  // pow(x, 2) is exact as x*x
  alpha = 1. / sqrt(0.5 * (sg + sqrt(sg * sg + 4. * (B2 + us2))));
  for (int j = 0; j < 3; j++) b[j] *= alpha;
  sg = 1. / (1. + dot3(b, b));
  alpha = dot3(u, b);
  t[0] = sg * (u[0] + alpha * b[0] + b[2] * u[1] - b[1] * u[2]);
  t[1] = sg * (u[1] + alpha * b[1] + b[0] * u[2] - b[2] * u[0]);
  t[2] = sg * (u[2] + alpha * b[2] + b[1] * u[0] - b[0] * u[1]);
  gamma = gamma_inv(dot3(t, t));
  for (int j = 0; j < 3; j++) v[j] = t[j] * gamma;
  return true;
}

// azimuthal_current_distribution for one particle, src/particles.cpp:688-784
template <class S>
__device__ void deposit_jphi(const GridDesc &g, S &sk, double q, double r, double z, double vphi) {
  const double dr = g.dr, dz = g.dz;
  int r_i = cell_r(g, r);
  int z_k = cell_z(g, z);
  if (r_i < 0) r_i = 0;
  if (z_k < 0) z_k = 0;
  double r1 = r - 0.5 * dr;
  double r2 = (r_i + 0.5) * dr;
  double r3 = r + 0.5 * dr;
  double ro_v = q / (2. * kPi * dz * dr * r);
  Vol v_1 = (r > dr) ? vol_cell(g, r_i) : vol_cyl(g);
  Vol v_2 = vol_cell(g, r_i + 1);
  double dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
  double dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
  sk.template add<kCPhi>(r_i, z_k, ring_term(ro_v, dz1, r1, r2, v_1) * vphi);
  sk.template add<kCPhi>(r_i + 1, z_k, ring_term(ro_v, dz1, r2, r3, v_2) * vphi);
  sk.template add<kCPhi>(r_i, z_k + 1, ring_term(ro_v, dz2, r1, r2, v_1) * vphi);
  sk.template add<kCPhi>(r_i + 1, z_k + 1, ring_term(ro_v, dz2, r2, r3, v_2) * vphi);
}

// Particles::charge_weighting for one particle, src/particles.cpp:101-226
template <class S>
__device__ void deposit_rho(const GridDesc &g, S &sk, double q, double r, double z) {
  const double dr = g.dr, dz = g.dz;
  int r_i = cell_r(g, r);
  int z_k = cell_z(g, z);
  if (r_i < 0) r_i = 0;
  if (z_k < 0) z_k = 0;
  double dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
  double dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
  if (r > dr) {
    double r1 = r - 0.5 * dr, r2 = (r_i + 0.5) * dr, r3 = r + 0.5 * dr;
    double ro_v = q / (2. * kPi * dz * dr * r);
    Vol v_1 = vol_cell(g, r_i), v_2 = vol_cell(g, r_i + 1);
    sk.template add<kCRho>(r_i, z_k, ring_term(ro_v, dz1, r1, r2, v_1));
    sk.template add<kCRho>(r_i + 1, z_k, ring_term(ro_v, dz1, r2, r3, v_2));
    sk.template add<kCRho>(r_i, z_k + 1, ring_term(ro_v, dz2, r1, r2, v_1));
    sk.template add<kCRho>(r_i + 1, z_k + 1, ring_term(ro_v, dz2, r2, r3, v_2));
  } else if (r <= dr / 2.) {
    r_i = 0;
    double r2 = (r_i + 0.5) * dr, r3 = r + 0.5 * dr;
    double ro_v = q / (kPi * dz * (2. * r * r + dr * dr / 2.));
    Vol v_1 = vol_cyl(g), v_2 = vol_cell(g, r_i + 1);
    sk.template add<kCRho>(r_i, z_k, div_cr(ro_v * kPi * dz1 * (dr * dr / 2. - r * dr + r * r), v_1.v, v_1.inv));
    sk.template add<kCRho>(r_i + 1, z_k, ring_term(ro_v, dz1, r2, r3, v_2));
    sk.template add<kCRho>(r_i, z_k + 1, div_cr(ro_v * kPi * dz2 * (dr * dr / 2. - r * dr + r * r), v_1.v, v_1.inv));
    sk.template add<kCRho>(r_i + 1, z_k + 1, ring_term(ro_v, dz2, r2, r3, v_2));
  } else {
    double r1 = r - 0.5 * dr, r2 = (r_i + 0.5) * dr, r3 = r + 0.5 * dr;
    double ro_v = q / (2. * kPi * dz * dr * r);
    Vol v_1 = vol_cyl(g), v_2 = vol_cell(g, r_i + 1);
    sk.template add<kCRho>(r_i, z_k, ring_term(ro_v, dz1, r1, r2, v_1));
    sk.template add<kCRho>(r_i + 1, z_k, ring_term(ro_v, dz1, r2, r3, v_2));
    sk.template add<kCRho>(r_i, z_k + 1, ring_term(ro_v, dz2, r1, r2, v_1));
    sk.template add<kCRho>(r_i + 1, z_k + 1, ring_term(ro_v, dz2, r2, r3, v_2));
  }
}

// density prefactors, include/particles.h:169-178
__device__ __forceinline__ double density_strict(double q, double r, double dr, double dz, double dt) {
  return q / (kPi * 4. * r * dz * dz * dr * dt);
}
__device__ __forceinline__ double density_r(double q, double r, double dr, double dz, double dt) {
  return q / (2 * kPi * r * dr * dz * dt * 2 * dr);
}
__device__ __forceinline__ double density_z(double q, double r, double dr, double dz, double dt) {
  return q / (2 * kPi * r * dr * dz * dt * dz);
}

// simple_current_distribution: one straight segment inside cell (i_n, k_n),
// src/particles.cpp:333-474
template <class S>
__device__ void segment(const GridDesc &g, S &acc, double q, double radius_new, double longitude_new,
                        double radius_old, double longitude_old, int i_n, int k_n) {
  const double dr = g.dr, dz = g.dz, delta_t = g.dt;
  double delta_r = radius_new - radius_old;
  double delta_z = longitude_new - longitude_old;
  if ((fabs(delta_r) < kMnzl) || (fabs(delta_z) < kMnzl)) return;
  double k = delta_r / delta_z;
  double b = radius_old;
  double lg = log((k * delta_z + b) / b);
  double wj0, wj1;
  if (i_n >= 1) {
    double ssd = density_r(q, i_n * dr, dr, dz, delta_t);
    wj0 = ssd * (dr * delta_z - k * delta_z * delta_z / 2. - delta_z * b +
                 dr * dr / k * ((i_n + 0.5) * (i_n + 0.5) - 0.25) * lg);
    ssd = density_r(q, (i_n + 1) * dr, dr, dz, delta_t);
    wj1 = ssd * (k * delta_z * delta_z / 2. + delta_z * b + delta_z * dr +
                 dr * dr / k * (0.25 - (i_n + 0.5) * (i_n + 0.5)) * lg);
  } else {
    double ssd = q / (2. * kPi * dr / 4. * dr * dz * delta_t * dr);
    wj0 = ssd * (dr * delta_z - k * delta_z * delta_z / 2. - delta_z * b);
    ssd = density_r(q, dr, dr, dz, delta_t);
    wj1 = ssd * (k * delta_z * delta_z / 2. + delta_z * dr + delta_z * b);
  }
  acc.template add<kCZ>(i_n, k_n, wj0);
  acc.template add<kCZ>(i_n + 1, k_n, wj1);
  // J_r (identical in both radial branches)
  k = -delta_z / delta_r;
  double r0 = (i_n + 0.5) * dr;
  double r1 = radius_old;
  b = (k_n + 1.) * dz - longitude_old;
  double ssd = density_z(q, r0, dr, dz, delta_t);
  double lr = log((radius_old + delta_r) / radius_old);
  double wj = ssd * (r0 * k * delta_r + k / 2. * delta_r * (radius_old + delta_r / 2.) +
                     0.5 * delta_r * (b - k * (2 * r0 + r1)) +
                     delta_r * (b - k * r1) * (4 * r0 * r0 - dr * dr) / (8 * radius_old * (radius_old + delta_r)) +
                     (k * (r0 * r0 / 2. - dr * dr / 8.)) * lr);
  acc.template add<kCR>(i_n, k_n, wj);
  b = longitude_old - k_n * dz;
  wj = ssd * (-r0 * k * delta_r - k / 2. * delta_r * (radius_old + delta_r / 2.) +
              0.5 * delta_r * (b + k * (2 * r0 + r1)) +
              delta_r * (b + k * r1) * (4 * r0 * r0 - dr * dr) / (8 * radius_old * (radius_old + delta_r)) -
              (k * (r0 * r0 / 2. - dr * dr / 8.)) * lr);
  acc.template add<kCR>(i_n, k_n + 1, wj);
}

// strict_motion_weighting, src/particles.cpp:786-988
template <class S>
__device__ void strict_motion(const GridDesc &g, S &acc, double q, double radius_new,
                              double longitude_new, double radius_old, double longitude_old) {
  const double dr = g.dr, dz = g.dz, delta_t = g.dt;
  int i_n = cell_r(g, radius_new);
  int k_n = cell_z(g, longitude_new);
  int i_o = cell_r(g, radius_old);
  int k_o = cell_z(g, longitude_old);
  if (i_n < 0) i_n = 0;
  if (k_n < 0) k_n = 0;
  if (i_o < 0) i_o = 0;
  if (k_o < 0) k_o = 0;
  bool still_r = fabs(radius_new - radius_old) < kMnzl;
  bool still_z = fabs(longitude_new - longitude_old) < kMnzl;
  if (still_r && still_z) return;
  if (still_r) {
    // strictly axial motion (:810-887); the reference's two `+ 0.` adds to
    // J_r are no-ops and are omitted
    double value_part = 2. * kPi * radius_new * dr * dz;
    double r1 = radius_new - 0.5 * dr, r2 = (i_n + 0.5) * dr, r3 = radius_new + 0.5 * dr;
    double wl, wu;
    if (i_n == 0)
      wl = q / (delta_t * kPi * dr * dr / 4.) * kPi * (r2 * r2 - r1 * r1) / value_part;
    else
      wl = q / (delta_t * 2. * kPi * i_n * dr * dr) * kPi * (r2 * r2 - r1 * r1) / value_part;
    wu = q / (delta_t * 2 * kPi * (i_n + 1) * dr * dr) * kPi * (r3 * r3 - r2 * r2) / value_part;
    int res_k = k_n - k_o;
    if (res_k == 0) {
      double delta_z = longitude_new - longitude_old;
      acc.template add<kCZ>(i_n, k_n, wl * delta_z);
      acc.template add<kCZ>(i_n + 1, k_n, wu * delta_z);
    } else if (res_k == 1) {
      double delta_z = k_n * dz - longitude_old;
      acc.template add<kCZ>(i_n, k_n - 1, wl * delta_z);
      acc.template add<kCZ>(i_n + 1, k_n - 1, wu * delta_z);
      delta_z = longitude_new - k_n * dz;
      acc.template add<kCZ>(i_n, k_n, wl * delta_z);
      acc.template add<kCZ>(i_n + 1, k_n, wu * delta_z);
    } else if (res_k == -1) {
      double delta_z = (k_n + 1) * dz - longitude_old;
      acc.template add<kCZ>(i_n, k_n + 1, wl * delta_z);
      acc.template add<kCZ>(i_n + 1, k_n + 1, wu * delta_z);
      delta_z = longitude_new - (k_n + 1) * dz;
      acc.template add<kCZ>(i_n, k_n, wl * delta_z);
      acc.template add<kCZ>(i_n + 1, k_n, wu * delta_z);
    }
  } else if (still_z) {
    // strictly radial motion (:890-986)
    int res_i = i_n - i_o;
    double left_delta_z = (k_n + 1) * dz - longitude_new;
    double right_delta_z = longitude_new - k_n * dz;
    if (res_i == 0) {
      double delta_r = radius_new - radius_old;
      double r0 = (i_n + 0.5) * dr;
      double ssd = density_strict(q, r0, dr, dz, delta_t);
      double wj = ssd * (delta_r - r0 * r0 / (radius_old + delta_r) + r0 * r0 / radius_old +
                         dr * dr / (4. * (radius_old + delta_r)) - dr * dr / (4. * radius_old));
      acc.template add<kCR>(i_n, k_n, wj * left_delta_z);
      acc.template add<kCR>(i_n, k_n + 1, wj * right_delta_z);
    } else if (res_i == 1) {
      double delta_r = (i_n)*dr - radius_old;
      double r0 = (i_n - 0.5) * dr;
      double ssd = density_strict(q, r0, dr, dz, delta_t);
      double wj = ssd * (delta_r - r0 * r0 / (radius_old + delta_r) + r0 * r0 / radius_old +
                         dr * dr / (4. * (radius_old + delta_r)) - dr * dr / (4. * radius_old));
      acc.template add<kCR>(i_n - 1, k_n, wj * left_delta_z);
      acc.template add<kCR>(i_n - 1, k_n + 1, wj * right_delta_z);
      delta_r = radius_new - i_n * dr;
      r0 = (i_n + 0.5) * dr;
      ssd = density_strict(q, r0, dr, dz, delta_t);
      // `r0 * r0 / i_n * dr` is ((r0*r0)/i_n)*dr in the reference (:941)
      const double slip = g.strict_fix ? r0 * r0 / (i_n * dr) : r0 * r0 / i_n * dr;
      wj = ssd * (delta_r - r0 * r0 / (i_n * dr + delta_r) + slip +
                  dr * dr / (4. * (i_n * dr + delta_r)) - dr * dr / (4. * i_n * dr));
      acc.template add<kCR>(i_n, k_n, wj * left_delta_z);
      acc.template add<kCR>(i_n, k_n + 1, wj * right_delta_z);
    } else if (res_i == -1) {
      double delta_r = (i_n + 1) * dr - radius_old;
      double r0 = (i_n + 1.5) * dr;
      double ssd = density_strict(q, r0, dr, dz, delta_t);
      double wj = ssd * (delta_r - r0 * r0 / (radius_old + delta_r) + r0 * r0 / radius_old +
                         dr * dr / (4. * (radius_old + delta_r)) - dr * dr / (4. * radius_old));
      acc.template add<kCR>(i_n + 1, k_n, wj * left_delta_z);
      acc.template add<kCR>(i_n + 1, k_n + 1, wj * right_delta_z);
      delta_r = radius_new - (i_n + 1) * dr;
      r0 = (i_n + 0.5) * dr;
      ssd = density_strict(q, r0, dr, dz, delta_t);
      // same precedence as above (:976)
      const double slip1 = g.strict_fix ? r0 * r0 / ((i_n + 1) * dr) : r0 * r0 / (i_n + 1) * dr;
      wj = ssd * (delta_r - r0 * r0 / ((i_n + 1) * dr + delta_r) + slip1 +
                  dr * dr / (4. * ((i_n + 1) * dr + delta_r)) - dr * dr / (4. * (i_n + 1) * dr));
      acc.template add<kCR>(i_n, k_n, wj * left_delta_z);
      acc.template add<kCR>(i_n, k_n + 1, wj * right_delta_z);
    }
  }
}

// one straight piece of a path: segment(rn, zn, ro, zo) inside cell (i, k)
struct Seg {
  double rn, zn, ro, zo;
  int i, k;
};

// current_distribution for one particle, src/particles.cpp:480-685: pick the
// cells of both ends, split the straight path at the cell faces it crosses
// and deposit every piece inside its own cell.  The pieces are collected
// first (with the reference's boundary-point arithmetic) and then deposited
// by one loop, so that a single copy of segment() is generated.
template <class S>
__device__ void deposit_j(const GridDesc &g, S &acc, double q, double pr, double pz, double por,
                          double poz) {
  const double dr = g.dr, dz = g.dz;
  int i_n = cell_r(g, pr);
  int k_n = cell_z(g, pz);
  int i_o = cell_r(g, por);
  int k_o = cell_z(g, poz);
  if (i_n < 0) i_n = 0;
  if (k_n < 0) k_n = 0;
  if (i_o < 0) i_o = 0;
  if (k_o < 0) k_o = 0;
  if (por == (i_o + 1) * dr) i_o = i_n;
  if (poz == (k_o + 1) * dz) k_o = k_n;
  if (pr == (i_n + 1) * dr) i_n = i_o;
  if (pz == (k_n + 1) * dz) k_n = k_o;
  int res_cell = abs(i_n - i_o) + abs(k_n - k_o);
  if ((fabs(pr - por) < kMnzl) || (fabs(pz - poz) < kMnzl)) {
    strict_motion(g, acc, q, pr, pz, por, poz);
    return;
  }
  Seg s0{pr, pz, por, poz, i_n, k_n}, s1 = s0, s2 = s0;
  int ns = 0;
  if (res_cell == 0) {
    ns = 1;
  } else if (res_cell == 1) {
    if ((i_n != i_o) && (k_n == k_o)) {
      if (por > (i_n + 1) * dr) {
        double a = (por - pr) / (poz - pz);
        double r_b = (i_n + 1) * dr;
        double z_b = pz + (r_b - pr) / a;
        s0 = Seg{r_b, z_b, por, poz, i_n + 1, k_n};
        s1 = Seg{pr, pz, r_b, z_b, i_n, k_n};
      } else {
        double a = (pr - por) / (pz - poz);
        double r_b = (i_n)*dr;
        double z_b = poz + (r_b - por) / a;
        s0 = Seg{r_b, z_b, por, poz, i_n - 1, k_n};
        s1 = Seg{pr, pz, r_b, z_b, i_n, k_n};
      }
      ns = 2;
    } else if ((i_n == i_o) && (k_n != k_o)) {
      if (poz < k_n * dz) {
        double z_b = k_n * dz;
        double a = (pr - por) / (pz - poz);
        double r_b = por + a * (z_b - poz);
        s0 = Seg{r_b, z_b, por, poz, i_n, k_n - 1};
        s1 = Seg{pr, pz, r_b, z_b, i_n, k_n};
      } else {
        double z_b = (k_n + 1) * dz;
        double a = (por - pr) / (poz - pz);
        double r_b = pr + a * (z_b - pz);
        s0 = Seg{r_b, z_b, por, poz, i_n, k_n + 1};
        s1 = Seg{pr, pz, r_b, z_b, i_n, k_n};
      }
      ns = 2;
    }
  } else if (res_cell == 2) {
    if (i_o < i_n) {
      if (k_o < k_n) {
        double a = (pr - por) / (pz - poz);
        double r1 = i_n * dr;
        double z1 = poz + (r1 - por) / a;
        double z2 = k_n * dz;
        double r2 = por + (z2 - poz) * a;
        if (z1 < k_n * dz) {
          s0 = Seg{r1, z1, por, poz, i_n - 1, k_n - 1};
          s1 = Seg{r2, z2, r1, z1, i_n, k_n - 1};
          s2 = Seg{pr, pz, r2, z2, i_n, k_n};
          ns = 3;
        } else if (z1 > k_n * dz) {
          s0 = Seg{r2, z2, por, poz, i_n - 1, k_n - 1};
          s1 = Seg{r1, z1, r2, z2, i_n - 1, k_n};
          s2 = Seg{pr, pz, r1, z1, i_n, k_n};
          ns = 3;
        }
      } else {
        double a = (pr - por) / (pz - poz);
        double r1 = i_n * dr;
        double z1 = poz + (r1 - por) / a;
        double z2 = (k_n + 1) * dz;
        double r2 = por + -(poz - z2) * a;
        if (z1 > (k_n + 1) * dz) {
          s0 = Seg{r1, z1, por, poz, i_n - 1, k_n + 1};
          s1 = Seg{r2, z2, r1, z1, i_n, k_n + 1};
          s2 = Seg{pr, pz, r2, z2, i_n, k_n};
          ns = 3;
        } else if (z1 < (k_n + 1) * dz) {
          s0 = Seg{r2, z2, por, poz, i_n - 1, k_n + 1};
          s1 = Seg{r1, z1, r2, z2, i_n - 1, k_n};
          s2 = Seg{pr, pz, r1, z1, i_n, k_n};
          ns = 3;
        }
      }
    } else if (i_o > i_n) {
      if (k_o < k_n) {
        double a = (pr - por) / (pz - poz);
        double r1 = (i_n + 1) * dr;
        double z1 = poz + -(por - r1) / a;
        double z2 = k_n * dz;
        double r2 = por - -(z2 - poz) * a;
        if (z1 < (k_n)*dz) {
          s0 = Seg{r1, z1, por, poz, i_n + 1, k_n - 1};
          s1 = Seg{r2, z2, r1, z1, i_n, k_n - 1};
          s2 = Seg{pr, pz, r2, z2, i_n, k_n};
          ns = 3;
        } else if (z1 > (k_n)*dz) {
          s0 = Seg{r2, z2, por, poz, i_n + 1, k_n - 1};
          s1 = Seg{r1, z1, r2, z2, i_n + 1, k_n};
          s2 = Seg{pr, pz, r1, z1, i_n, k_n};
          ns = 3;
        }
      } else if (k_o > k_n) {
        double a = (por - pr) / (poz - pz);
        double r1 = (i_n + 1) * dr;
        double z1 = pz + (r1 - pr) / a;
        double z2 = (k_n + 1) * dz;
        double r2 = pr + (z2 - pz) * a;
        if (z1 > (k_n + 1) * dz) {
          s0 = Seg{r1, z1, por, poz, i_n + 1, k_n + 1};
          s1 = Seg{r2, z2, r1, z1, i_n, k_n + 1};
          s2 = Seg{pr, pz, r2, z2, i_n, k_n};
          ns = 3;
        } else if (z1 < (k_n + 1) * dz) {
          s0 = Seg{r2, z2, por, poz, i_n + 1, k_n + 1};
          s1 = Seg{r1, z1, r2, z2, i_n + 1, k_n};
          s2 = Seg{pr, pz, r1, z1, i_n, k_n};
          ns = 3;
        }
      }
    }
  }
#pragma nounroll
  for (int t = 0; t < ns; t++) {
    const Seg c = t == 0 ? s0 : (t == 1 ? s1 : s2);
    segment(g, acc, q, c.rn, c.zn, c.ro, c.zo, c.i, c.k);
  }
}

// reflection_single (src/particles.cpp:1318-1353) / Bunch::reflection_single
// (src/bunch.cpp:126-153) on the Cartesian half-step position
__device__ __forceinline__ void walls(const GridDesc &g, bool is_bunch, double &x, double &z, double &v0,
                                      double &v2, bool &alive) {
  const double dr = g.dr, dz = g.dz;
  double radius_wall = g.r_size - dr / 2.;
  double longitude_wall = g.z_size - dz / 2.;
  double half_dr = dr / 2., half_dz = dz / 2.;
  if (is_bunch) {
    if (x > radius_wall) alive = false;
    if (z > longitude_wall) alive = false;
  } else {
    if (x > radius_wall) {
      x = radius_wall * 2. - x;
      v0 = -v0;
    }
    if (z > longitude_wall) {
      z = longitude_wall * 2. - z;
      v2 = -v2;
    }
  }
  if (x < half_dr) {
    x = dr - x;
    v0 = -v0;
  }
  if (z < half_dz) {
    z = dz - z;
    v2 = -v2;
  }
}

__device__ __forceinline__ bool finite2(double a, double b) { return isfinite(a) && isfinite(b); }

// step_v part for one particle (src/particles.cpp:1296-1316): push with the
// gathered fields, then half step and rotation to r-z.  Returns false (and
// flags) on a non-finite position or a superluminal velocity.
__device__ bool push_half(const GridDesc &g, const Fields &f, int pusher, int *err, double q, double m,
                          double &r, double &z, double v[3], double &sin_t) {
  if (!finite2(r, z)) {
    atomicOr(err, kErrNonfinite);
    return false;
  }
  double e[3], h[3];
  gather_e(g, f, r, z, e, err);
  gather_h(g, f, r, z, h, err);
  bool ok = (pusher == 1) ? vay(q, m, g.dt, e, h, v) : boris(q, m, g.dt, e, h, v);
  if (!ok) {
    atomicOr(err, kErrSuperluminal);
    return false;
  }
  const double half_dt = g.dt / 2.;
  double x = r + v[0] * half_dt;
  double y = 0.0 + v[1] * half_dt;
  z = z + v[2] * half_dt;
  double rr = sqrt(x * x + y * y);
  sin_t = y / rr;
  r = rr;
  return true;
}

// full_current_distribution part for one particle (src/particles.cpp:1367-1373)
template <class S>
__device__ bool deposit_move(const GridDesc &g, S &sk, int *err, bool is_bunch, bool want_rho, double q,
                             double &r, double &z, double v[3], bool &alive, double r_old, double z_old) {
  deposit_jphi(g, sk, q, r, z, v[1]);
  if (!finite2(r, z)) {
    atomicOr(err, kErrNonfinite);
    return false;
  }
  const double half_dt = g.dt / 2.;
  double x = r + v[0] * half_dt;
  double y = 0.0 + v[1] * half_dt;
  z = z + v[2] * half_dt;
  walls(g, is_bunch, x, z, v[0], v[2], alive);
  double rr = sqrt(x * x + y * y);
  double st = y / rr;
  r = rr;
  if (alive) {
    deposit_j(g, sk, q, r, z, r_old, z_old);
    double ct = sqrt(1 - st * st);
    double u2 = ct * v[0] - st * v[1];
    double w2 = st * v[0] + ct * v[1];
    v[0] = u2;
    v[1] = w2;
    if (want_rho) deposit_rho(g, sk, q, r, z);
  }
  return true;
}

// the whole step of one particle in slot i (push, both half steps, deposits)
template <class S>
__device__ __forceinline__ void step_particle(const GridDesc &g, const Fields &f, S &sk, int *err,
                                              const SpeciesArgs &s, int pusher, bool want_rho, long long i) {
  if (s.alive[i] != 1) return;
  double r = s.r[i], z = s.z[i];
  double v[3] = {s.vr[i], s.vphi[i], s.vz[i]};
  const double w = s.w[i];
  const double q = s.charge * w, m = s.mass * w;
  const double r_old = r, z_old = z;
  // contributions of this particle are centred on its starting cell
  int ci = cell_r(g, r), ck = cell_z(g, z);
  sk.begin_cell(ci < 0 ? 0 : ci, ck < 0 ? 0 : ck);
  double sin_t;
  if (!push_half(g, f, pusher, err, q, m, r, z, v, sin_t)) return;
  bool alive = true;
  if (!deposit_move(g, sk, err, s.is_bunch, want_rho, q, r, z, v, alive, r_old, z_old)) return;
  s.r[i] = r;
  s.z[i] = z;
  s.vr[i] = v[0];
  s.vphi[i] = v[1];
  s.vz[i] = v[2];
  if (!alive) s.alive[i] = 0;
}

// chunked plasma species: 4 wavefronts per block, each walking one chunk of
// kChunk slots kPPT particles per lane, deposits through the tile sink
// with the divisions taken from shared reciprocals (div_cr) the compiler's own
// choice, 227 VGPRs at 2 waves per SIMD without spills, is the fastest: 25.9 ms
// at cfg3 against 26.7 ms for the IEEE divisions at 3 waves (168 VGPRs, 12
// spilled); forcing 3 waves with div_cr spills 146 VGPRs (43 ms)
__global__ void __launch_bounds__(256)
k_fused_tiled(GridDesc g, Fields f, GlobalSink gs, SpeciesArgs s, long long n_cap, int pusher) {
  __shared__ double box[TileSink::kSize];
  const long long base = (long long)blockIdx.x * kBlockParticles;
  // anchor samples: slot base + j is the logical particle 16 j of the block
  int ai, ak;
  {
    bool live[3];
    int ri[3], kk[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      live[j] = base + j < n_cap && s.alive[base + j] == 1;
      ri[j] = live[j] ? max(0, cell_r(g, s.r[base + j])) : 0;
      kk[j] = live[j] ? max(0, cell_z(g, s.z[base + j])) : 0;
    }
    tile_anchor(live, ri, kk, ai, ak);
  }
  TileSink sk;
  box_init(box, ai, ak, sk, gs);
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long cbase = base + (long long)wave * kChunk + lane;
  for (int j = 0; j < kPPT; j++) {
    long long p = cbase + (long long)j * 64;
    if (p < n_cap) step_particle(g, f, sk, gs.err, s, pusher, false, p);
  }
  sk.flush();
  __syncthreads();
  box_flush(box, sk);
}

// bunches (plain layout): the tiled walk with the box anchored at the
// smallest living cell of the block; PPT particles per lane (wave w of a
// block takes the 64 * PPT slots from w * 64 * PPT), few for a small bunch so
// that its blocks still spread over the GPU
template <int PPT>
__global__ void __launch_bounds__(256) k_fused_bunch(GridDesc g, Fields f, GlobalSink gs, SpeciesArgs s,
                                                     long long n_cap, int pusher) {
  __shared__ double box[BunchSink::kSize];
  const long long base = (long long)blockIdx.x * (kWavesPerBlock * 64 * PPT);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long cbase = base + (long long)wave * (64 * PPT) + lane;
  int ai = 0x7fffffff, ak = 0x7fffffff;
  for (int j = 0; j < PPT; j++) {
    long long p = cbase + (long long)j * 64;
    if (p < n_cap && s.alive[p] == 1) {
      ai = min(ai, max(0, cell_r(g, s.r[p])));
      ak = min(ak, max(0, cell_z(g, s.z[p])));
    }
  }
  block_min2(ai, ak);
  if (ai == 0x7fffffff) return;
  BunchSink sk;
  box_init(box, ai, ak, sk, gs);
  __syncthreads();
  for (int j = 0; j < PPT; j++) {
    long long p = cbase + (long long)j * 64;
    if (p < n_cap) step_particle(g, f, sk, gs.err, s, pusher, true, p);
  }
  sk.flush();
  __syncthreads();
  box_flush(box, sk);
}

// plain layout, one particle per thread, deposits straight to HBM
__global__ void __launch_bounds__(256) k_fused_plain(GridDesc g, Fields f, GlobalSink gs, SpeciesArgs s,
                                                     long long n_cap, int pusher, int want_rho) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cap) return;
  GlobalSink sk = gs;
  step_particle(g, f, sk, gs.err, s, pusher, want_rho != 0, i);
}

__global__ void __launch_bounds__(256) k_step_v(GridDesc g, Fields f, int *err, SpeciesArgs s, long long n_cap,
                                                int pusher) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cap) return;
  if (s.alive[i] != 1) return;
  double r = s.r[i], z = s.z[i];
  double v[3] = {s.vr[i], s.vphi[i], s.vz[i]};
  const double w = s.w[i];
  s.r_old[i] = r;
  s.z_old[i] = z;
  double sin_t;
  if (!push_half(g, f, pusher, err, s.charge * w, s.mass * w, r, z, v, sin_t)) return;
  s.r[i] = r;
  s.z[i] = z;
  s.vr[i] = v[0];
  s.vphi[i] = v[1];
  s.vz[i] = v[2];
  s.sin_t[i] = sin_t;
}

__global__ void __launch_bounds__(256) k_full_current(GridDesc g, GlobalSink gs, SpeciesArgs s, long long n_cap) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cap) return;
  if (s.alive[i] != 1) return;
  double r = s.r[i], z = s.z[i];
  double v[3] = {s.vr[i], s.vphi[i], s.vz[i]};
  const double q = s.charge * s.w[i];
  bool alive = true;
  GlobalSink sk = gs;
  if (!deposit_move(g, sk, gs.err, s.is_bunch, false, q, r, z, v, alive, s.r_old[i], s.z_old[i])) return;
  s.r[i] = r;
  s.z[i] = z;
  s.vr[i] = v[0];
  s.vphi[i] = v[1];
  s.vz[i] = v[2];
  if (!alive) s.alive[i] = 0;
}

__global__ void __launch_bounds__(256) k_charge(GridDesc g, GlobalSink gs, SpeciesArgs s, long long n_cap) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_cap) return;
  if (s.alive[i] != 1) return;
  double r = s.r[i], z = s.z[i];
  if (!finite2(r, z)) {
    atomicOr(gs.err, kErrNonfinite);
    return;
  }
  GlobalSink sk = gs;
  deposit_rho(g, sk, s.charge * s.w[i], r, z);
}

__global__ void k_gather(GridDesc g, Fields f, int which, const double *r, const double *z, int n, double *out,
                         int *err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v[3];
  if (which == 0)
    gather_e(g, f, r[i], z[i], v, err);
  else
    gather_h(g, f, r[i], z[i], v, err);
  out[3 * i + 0] = v[0];
  out[3 * i + 1] = v[1];
  out[3 * i + 2] = v[2];
}

// Bunch::bunch_inject body, src/bunch.cpp:104-121: particles start..start+count
__global__ void k_inject(double *z, double *vz, uint8_t *alive, const double *u, int n, unsigned start,
                         unsigned count, double dl, double half_dz, double velocity) {
  unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  unsigned j = i + start;
  if (j >= (unsigned)n) return;
  z[j] = dl * u[j] + half_dz;
  vz[j] = velocity;
  alive[j] = 1;
}

Fields fields_of(Ctx &c) {
  return Fields{c.a[kER], c.a[kEPhi], c.a[kEZ], c.a[kHHR], c.a[kHHPhi], c.a[kHHZ]};
}

GlobalSink sink_of(Ctx &c) {
  GlobalSink gs;
  gs.nr = c.g.nr;
  gs.nz = c.g.nz;
  gs.j = c.a[kJR];
  gs.rho = c.a[kRhoBunch];
  gs.err = c.err;
  return gs;
}

SpeciesArgs args_of(Species &sp) {
  SpeciesArgs s;
  s.r = sp.f64.r;
  s.z = sp.f64.z;
  s.vr = sp.f64.vr;
  s.vphi = sp.f64.vphi;
  s.vz = sp.f64.vz;
  s.w = sp.f64.w;
  s.alive = sp.f64.alive;
  s.r_old = sp.f64.r_old;
  s.z_old = sp.f64.z_old;
  s.sin_t = sp.f64.sin_t;
  s.charge = sp.charge;
  s.mass = sp.mass;
  s.n = sp.n;
  s.is_bunch = sp.kind == 1;
  return s;
}

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

void launch_fused_f64(Ctx &c, Species &sp) {
  if (sp.n_cap <= 0) return;
  if (sp.chunked) {
    unsigned blocks = (unsigned)((sp.n_cap + kBlockParticles - 1) / kBlockParticles);
    hipLaunchKernelGGL(k_fused_tiled, dim3(blocks), dim3(256), 0, c.stream, c.g, fields_of(c), sink_of(c),
                       args_of(sp), sp.n_cap, c.pusher);
  } else {
    // the most particles per lane that still give 1024 blocks
    const long long n = sp.n_cap, per = (long long)kWavesPerBlock * 64;
    if (n >= 1024 * per * kPPT)
      hipLaunchKernelGGL(k_fused_bunch<kPPT>, dim3((unsigned)((n + per * kPPT - 1) / (per * kPPT))), dim3(256), 0,
                         c.stream, c.g, fields_of(c), sink_of(c), args_of(sp), n, c.pusher);
    else if (n >= 1024 * per * 4)
      hipLaunchKernelGGL(k_fused_bunch<4>, dim3((unsigned)((n + per * 4 - 1) / (per * 4))), dim3(256), 0, c.stream,
                         c.g, fields_of(c), sink_of(c), args_of(sp), n, c.pusher);
    else
      hipLaunchKernelGGL(k_fused_bunch<1>, dim3((unsigned)((n + per - 1) / per)), dim3(256), 0, c.stream, c.g,
                         fields_of(c), sink_of(c), args_of(sp), n, c.pusher);
  }
}

void launch_step_v_f64(Ctx &c, Species &sp) {
  if (sp.n_cap <= 0) return;
  hipLaunchKernelGGL(k_step_v, dim3(blocks_for(sp.n_cap)), dim3(256), 0, c.stream, c.g, fields_of(c), c.err,
                     args_of(sp), sp.n_cap, c.pusher);
}

void launch_full_current_f64(Ctx &c, Species &sp) {
  if (sp.n_cap <= 0) return;
  hipLaunchKernelGGL(k_full_current, dim3(blocks_for(sp.n_cap)), dim3(256), 0, c.stream, c.g, sink_of(c),
                     args_of(sp), sp.n_cap);
}

void launch_charge_weighting_f64(Ctx &c, Species &sp, double *rho) {
  if (sp.n_cap <= 0) return;
  GlobalSink gs = sink_of(c);
  if (rho) gs.rho = rho;
  hipLaunchKernelGGL(k_charge, dim3(blocks_for(sp.n_cap)), dim3(256), 0, c.stream, c.g, gs, args_of(sp), sp.n_cap);
}

void launch_gather_f64(Ctx &c, int which, const double *r, const double *z, int n, double *out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather, dim3(blocks_for(n)), dim3(256), 0, c.stream, c.g, fields_of(c), which, r, z, n,
                     out, c.err);
}

void launch_bunch_inject_f64(Ctx &c, Species &sp, unsigned start, unsigned count, double dl) {
  if (count == 0) return;
  hipLaunchKernelGGL(k_inject, dim3((count + 255) / 256), dim3(256), 0, c.stream, sp.f64.z, sp.f64.vz,
                     sp.f64.alive, sp.inject_u, sp.n, start, count, dl, c.g.dz / 2.0, sp.velocity);
}

}  // namespace pdp3
