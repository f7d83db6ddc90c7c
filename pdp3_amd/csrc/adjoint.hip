// Hand-written adjoint of one loop iteration (src/loadInitParam.cpp:471-587)
// in the FP64 plain form's arithmetic, for gradients of a field objective with
// respect to the particle weights.
//
// Forward iteration n:  H^{n+1} = H^n + A E^n, Hh^n = H^n + A E^n / 2
//                       (x, v)^{n+1}, J^n = P((x, v)^n, E^n, Hh^n, w)
//                       E^{n+1} = Ke E^n - Kh (J^n - C H^{n+1})
// Backward, given the adjoints of E^{n+1}, H^{n+1} and the particle states:
//   1. lambda_J = -Kh lambda_E;  lambda_H += C^T (Kh lambda_E);  lambda_E <- Ke lambda_E
//   2. per particle, the vector-Jacobian product of P: the particle's step is
//      re-executed in dual numbers (value + N directional derivatives) over
//      its 12 inputs (r, z, v, the gathered E and H, w) in 12/N passes; every
//      output (new r, z, v and each grid deposit) contributes its adjoint
//      times its derivative; the adjoints of the gathered fields are scattered
//      back to the nodes with the gather weights, the weight adjoint is
//      accumulated per particle
//   3. lambda_alpha = lambda_H + lambda_Hh / 2;  lambda_E += A^T lambda_alpha;
//      lambda_H <- lambda_H + lambda_Hh
// The operators A (HField::calc_field, src/hField.cpp:93-157), C, Ke, Kh
// (EField::calc_field, src/eField.cpp:85-170, PML in Ke / Kh) are transposed
// node by node in pull form (no atomics).  The particle step follows
// particles_fp64.hip expression for expression.
#include "context.hpp"
#include "crmath.hpp"
#include "sink.hpp"

namespace pdp3 {

namespace {

constexpr int kDN = 4;       // directions per pass
static_assert(kDN == kTangentDirs, "the tangent-linear model runs one direction per dual slot");
constexpr int kInputs = 12;  // r, z, v0, v1, v2, e0, e1, e2, h0, h1, h2, w
constexpr int kPasses = kInputs / kDN;

struct D {
  double v;
  double d[kDN];
};

__device__ __forceinline__ D cst(double x) {
  D r;
  r.v = x;
  for (int j = 0; j < kDN; j++) r.d[j] = 0.0;
  return r;
}
// x as input `input` of pass `pass` (derivative 1 in its direction); selects
// on constant indices, so that no dual number is addressed through memory
__device__ __forceinline__ D seed(double x, int input, int pass) {
  D r;
  r.v = x;
  const int jin = input - pass * kDN;
#pragma unroll
  for (int j = 0; j < kDN; j++) r.d[j] = j == jin ? 1.0 : 0.0;
  return r;
}
__device__ __forceinline__ D operator+(D a, D b) {
  D r;
  r.v = a.v + b.v;
  for (int j = 0; j < kDN; j++) r.d[j] = a.d[j] + b.d[j];
  return r;
}
__device__ __forceinline__ D operator-(D a, D b) {
  D r;
  r.v = a.v - b.v;
  for (int j = 0; j < kDN; j++) r.d[j] = a.d[j] - b.d[j];
  return r;
}
__device__ __forceinline__ D operator-(D a) {
  D r;
  r.v = -a.v;
  for (int j = 0; j < kDN; j++) r.d[j] = -a.d[j];
  return r;
}
__device__ __forceinline__ D operator*(D a, D b) {
  D r;
  r.v = a.v * b.v;
  for (int j = 0; j < kDN; j++) r.d[j] = a.d[j] * b.v + a.v * b.d[j];
  return r;
}
__device__ __forceinline__ D operator/(D a, D b) {
  D r;
  r.v = a.v / b.v;
  const double ib = 1.0 / b.v;
  for (int j = 0; j < kDN; j++) r.d[j] = (a.d[j] - r.v * b.d[j]) * ib;
  return r;
}
__device__ __forceinline__ D operator+(D a, double b) { return a + cst(b); }
__device__ __forceinline__ D operator+(double a, D b) { return cst(a) + b; }
__device__ __forceinline__ D operator-(D a, double b) { return a - cst(b); }
__device__ __forceinline__ D operator-(double a, D b) { return cst(a) - b; }
__device__ __forceinline__ D operator*(D a, double b) {
  D r;
  r.v = a.v * b;
  for (int j = 0; j < kDN; j++) r.d[j] = a.d[j] * b;
  return r;
}
__device__ __forceinline__ D operator*(double a, D b) {
  D r;
  r.v = a * b.v;
  for (int j = 0; j < kDN; j++) r.d[j] = a * b.d[j];
  return r;
}
__device__ __forceinline__ D operator/(D a, double b) {
  D r;
  r.v = a.v / b;
  for (int j = 0; j < kDN; j++) r.d[j] = a.d[j] / b;
  return r;
}
__device__ __forceinline__ D operator/(double a, D b) { return cst(a) / b; }
__device__ __forceinline__ D dsqrt(D a) {
  D r;
  r.v = sqrt(a.v);
  const double h = 0.5 / r.v;
  for (int j = 0; j < kDN; j++) r.d[j] = a.d[j] * h;
  return r;
}
__device__ __forceinline__ D dlog(D a) {
  D r;
  r.v = log(a.v);
  for (int j = 0; j < kDN; j++) r.d[j] = a.d[j] / a.v;
  return r;
}
// pow(y, -0.5) of lib::get_gamma_inv with the correctly rounded value
__device__ __forceinline__ D drsqrt(D a) {
  D r;
  r.v = rsqrt_cr(a.v);
  const double h = -0.5 * r.v * r.v * r.v;
  for (int j = 0; j < kDN; j++) r.d[j] = a.d[j] * h;
  return r;
}

constexpr double kC2 = kLightSpeed * kLightSpeed;

// ------------------------------------------------------------ the particle

struct AdjFields {
  const double *er, *ephi, *ez, *hhr, *hhphi, *hhz;  // E^n, Hh^n
};
struct AdjGrid {
  const double *ljr, *ljphi, *ljz;             // adjoint of J^n
  double *ler, *lephi, *lez, *lhr, *lhphi, *lhz;  // adjoint of E^n, Hh^n (scatter, atomics)
};
struct AdjSpecies {
  const double *r, *z, *vr, *vphi, *vz, *w;  // state at step n
  const uint8_t *alive;
  double *lr, *lz, *lvr, *lvphi, *lvz;  // in: adjoint of the state after step n; out: before
  double *lw;                           // accumulated adjoint of the weight
  double charge, mass;
  long long n;
  int is_bunch;
};

__device__ __forceinline__ int cell_number(double x, double dx) { return (int)ceil(x / dx) - 1; }
__device__ __forceinline__ double cell_volume(int i, double dr, double dz) { return kPi * dz * dr * dr * 2.0 * i; }
__device__ __forceinline__ double cyl_vol(double z, double r) { return kPi * z * r * r / 4.; }

// the node weights of the four-node gather (src/eField.cpp:204-316) for one
// component family; weights[] multiply F at (i_r, k), (i_r+1, k), (i_r, k+1),
// (i_r+1, k+1).  Dual in (r, z).
struct GatherW {
  int ir, kz;
  D w[4];
};
__device__ GatherW gather_w(const GridDesc &g, D radius, D longitude, bool lo, bool zhalf) {
  const double dr = g.dr, dz = g.dz;
  D r1 = radius - 0.5 * dr;
  D r3 = radius + 0.5 * dr;
  int i_r = lo ? cell_number(radius.v - 0.5 * dr, dr) : cell_number(radius.v, dr);
  int k_z = zhalf ? cell_number(longitude.v - 0.5 * dz, dz) : cell_number(longitude.v, dz);
  if (i_r < 0) i_r = 0;
  if (k_z < 0) k_z = 0;
  i_r = min(i_r, g.nr - 2);
  k_z = min(k_z, g.nz - 2);
  double vol_1, vol_2, r2;
  if (lo) {
    vol_1 = cell_volume(i_r + 1, dr, dz);
    vol_2 = cell_volume(i_r + 3, dr, dz);
    r2 = (i_r + 1) * dr;
  } else {
    vol_1 = (radius.v > dr) ? cell_volume(i_r, dr, dz) : cyl_vol(dz, dr);
    r2 = (i_r + 0.5) * dr;
    vol_2 = cell_volume(i_r + 2, dr, dz);
  }
  D dz1, dz2;
  if (zhalf) {
    dz1 = (k_z + 1.5) * dz - longitude;
    dz2 = longitude - (k_z + 0.5) * dz;
  } else {
    dz1 = (k_z + 1) * dz - longitude;
    dz2 = longitude - k_z * dz;
  }
  GatherW out;
  out.ir = i_r;
  out.kz = k_z;
  // ring_term(F, dzx, ra, rb, vol) = F * pi * dzx * (rb^2 - ra^2) / vol
  D a_in = r2 * r2 - r1 * r1, a_out = r3 * r3 - r2 * r2;
  out.w[0] = kPi * dz1 * a_in / vol_1;
  out.w[1] = kPi * dz1 * a_out / vol_2;
  out.w[2] = kPi * dz2 * a_in / vol_1;
  out.w[3] = kPi * dz2 * a_out / vol_2;
  return out;
}

__device__ __forceinline__ D gather_d(const GridDesc &g, const double *F, D r, D z, bool lo, bool zhalf) {
  GatherW gw = gather_w(g, r, z, lo, zhalf);
  const double *row0 = F + (size_t)gw.ir * (size_t)g.nz + gw.kz;
  const double *row1 = row0 + g.nz;
  return gw.w[0] * row0[0] + gw.w[1] * row1[0] + gw.w[2] * row0[1] + gw.w[3] * row1[1];
}

// the adjoints of the gathered fields of one block's particles are summed in
// an LDS box of kABR x kABK nodes per component (anchored one row and two
// columns below the block's smallest cell) and added to the grid once per
// block; stencils outside the box go to the grid directly
constexpr int kABR = 4, kABK = 64, kABComp = 6;
struct AdjBox {
  double *box;  // LDS, kABComp x kABR x kABK
  int bi, bk;   // node of box[.][0][0]
};
__device__ __forceinline__ void box_scatter(const GridDesc &g, const AdjBox &b, int comp, double *LF, double r,
                                            double z, bool lo, bool zhalf, double lam) {
  if (lam == 0.0) return;
  GatherW gw = gather_w(g, cst(r), cst(z), lo, zhalf);
  // weights 0..3 multiply the nodes (ir, kz), (ir+1, kz), (ir, kz+1), (ir+1, kz+1)
  const double v00 = gw.w[0].v * lam, v10 = gw.w[1].v * lam, v01 = gw.w[2].v * lam, v11 = gw.w[3].v * lam;
  const int li = gw.ir - b.bi, lk = gw.kz - b.bk;
  if ((unsigned)li < (unsigned)(kABR - 1) && (unsigned)lk < (unsigned)(kABK - 1)) {
    double *q = b.box + (comp * kABR + li) * kABK + lk;
    atomicAdd(q, v00);
    atomicAdd(q + kABK, v10);
    atomicAdd(q + 1, v01);
    atomicAdd(q + kABK + 1, v11);
  } else {
    double *row0 = LF + (size_t)gw.ir * (size_t)g.nz + gw.kz;
    double *row1 = row0 + g.nz;
    unsafeAtomicAdd(row0, v00);
    unsafeAtomicAdd(row1, v10);
    unsafeAtomicAdd(row0 + 1, v01);
    unsafeAtomicAdd(row1 + 1, v11);
  }
}

// the deposit "sink" of the adjoint: every contribution v at node (i, k) of
// component C adds lambda_J[C](i, k) * dv to the accumulated input adjoints
struct AdjSink {
  const GridDesc *g;
  const double *ljr, *ljphi, *ljz;
  double acc[kDN];
  __device__ __forceinline__ void add(int comp, int i, int k, D v) {
    if ((unsigned)i >= (unsigned)g->nr || (unsigned)k >= (unsigned)g->nz) return;
    const double *L = comp == 0 ? ljr : comp == 1 ? ljphi : ljz;
    const double l = L[(size_t)i * (size_t)g->nz + (size_t)k];
    for (int j = 0; j < kDN; j++) acc[j] += l * v.d[j];
  }
};
constexpr int kAR = 0, kAPhi = 1, kAZ = 2;

__device__ __forceinline__ D dot3(const D *a, const D *b) {
  D res = cst(0.0);
  res = res + a[0] * b[0];
  res = res + a[1] * b[1];
  res = res + a[2] * b[2];
  return res;
}

// relativistic Boris (src/particles.cpp:1025-1150)
__device__ void boris_d(D q, D m, double dt, D e[3], D b[3], D v[3]) {
  D const1 = q * dt / (2 * m);
  for (int j = 0; j < 3; j++) e[j] = e[j] * const1;
  D bs = kMagnConst * const1;
  for (int j = 0; j < 3; j++) b[j] = b[j] * bs;
  D u[3], t[3];
  for (int j = 0; j < 3; j++) u[j] = (fabs(v[j].v) < kMnzl) ? cst(0.0) : v[j];
  D beta = dot3(u, u) / kC2;
  D gamma = 1 / dsqrt(1.0 - beta);
  for (int j = 0; j < 3; j++) u[j] = u[j] * gamma;
  for (int j = 0; j < 3; j++) u[j] = u[j] + e[j];
  gamma = drsqrt(1.0 + dot3(u, u) / kC2);
  for (int j = 0; j < 3; j++) b[j] = b[j] * gamma;
  D const2 = 2. / (1. + dot3(b, b));
  for (int j = 0; j < 3; j++) t[j] = u[j];
  u[0] = t[0] + const2 * ((t[1] - t[0] * b[2] + t[2] * b[0]) * b[2] - (t[2] + t[0] * b[1] - t[1] * b[0]) * b[1]);
  u[1] = t[1] + const2 * (-(t[0] + t[1] * b[2] - t[2] * b[1]) * b[2] + (t[2] + t[0] * b[1] - t[1] * b[0]) * b[0]);
  u[2] = t[2] + const2 * ((t[0] + t[1] * b[2] - t[2] * b[1]) * b[1] - (t[1] - t[0] * b[2] + t[2] * b[0]) * b[0]);
  for (int j = 0; j < 3; j++) u[j] = u[j] + e[j];
  gamma = drsqrt(1.0 + dot3(u, u) / kC2);
  for (int j = 0; j < 3; j++) v[j] = u[j] * gamma;
}

// Vay with the reference's scaling (src/particles.cpp:1152-1254)
__device__ void vay_d(D q, D m, double dt, D e[3], D b[3], D v[3]) {
  D const1 = q * dt / (2 * m);
  D s2 = 2 * const1;
  for (int j = 0; j < 3; j++) e[j] = e[j] * s2;
  D bs = kMagnConst * const1;
  for (int j = 0; j < 3; j++) b[j] = b[j] * bs;
  D u[3], t[3];
  for (int j = 0; j < 3; j++) u[j] = (fabs(v[j].v) < kMnzl) ? cst(0.0) : v[j];
  D beta = dot3(u, u) / kC2;
  D gamma = 1 / dsqrt(1.0 - beta);
  for (int j = 0; j < 3; j++) u[j] = u[j] * gamma;
  for (int j = 0; j < 3; j++) t[j] = u[j];
  for (int j = 0; j < 3; j++) b[j] = b[j] * s2;
  for (int j = 0; j < 3; j++) u[j] = u[j] + e[j];
  for (int j = 0; j < 3; j++) b[j] = b[j] * const1;
  gamma = drsqrt(1.0 + dot3(t, t) / kC2);
  u[0] = u[0] + gamma * (t[1] * b[2] - t[2] * b[1]);
  u[1] = u[1] + gamma * (t[2] * b[0] - t[0] * b[2]);
  u[2] = u[2] + gamma * (t[0] * b[1] - t[1] * b[0]);
  gamma = drsqrt(1.0 + dot3(u, u) / kC2);
  D alpha = 1. / gamma / gamma;
  D B2 = dot3(b, b);
  D sg = alpha - B2;
  D ub = dot3(u, b);
  D us2 = ub * ub;
  alpha = 1. / dsqrt(0.5 * (sg + dsqrt(sg * sg + 4. * (B2 + us2))));
  for (int j = 0; j < 3; j++) b[j] = b[j] * alpha;
  sg = 1. / (1. + dot3(b, b));
  alpha = dot3(u, b);
  t[0] = sg * (u[0] + alpha * b[0] + b[2] * u[1] - b[1] * u[2]);
  t[1] = sg * (u[1] + alpha * b[1] + b[0] * u[2] - b[2] * u[0]);
  t[2] = sg * (u[2] + alpha * b[2] + b[1] * u[0] - b[0] * u[1]);
  gamma = drsqrt(1.0 + dot3(t, t) / kC2);
  for (int j = 0; j < 3; j++) v[j] = t[j] * gamma;
}

__device__ __forceinline__ D ring(D F, D dzx, D ra, D rb, double vol) { return F * kPi * dzx * (rb * rb - ra * ra) / vol; }

// azimuthal_current_distribution, src/particles.cpp:688-784
template <class S>
__device__ void jphi_d(const GridDesc &g, S &sk, D q, D r, D z, D vphi) {
  const double dr = g.dr, dz = g.dz;
  int r_i = cell_number(r.v, dr), z_k = cell_number(z.v, dz);
  if (r_i < 0) r_i = 0;
  if (z_k < 0) z_k = 0;
  D r1 = r - 0.5 * dr;
  D r2 = cst((r_i + 0.5) * dr);
  D r3 = r + 0.5 * dr;
  D ro_v = q / (2. * kPi * dz * dr * r);
  double v_1 = (r.v > dr) ? cell_volume(r_i, dr, dz) : cyl_vol(dz, dr);
  double v_2 = cell_volume(r_i + 1, dr, dz);
  D dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
  D dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
  sk.add(kAPhi, r_i, z_k, ring(ro_v, dz1, r1, r2, v_1) * vphi);
  sk.add(kAPhi, r_i + 1, z_k, ring(ro_v, dz1, r2, r3, v_2) * vphi);
  sk.add(kAPhi, r_i, z_k + 1, ring(ro_v, dz2, r1, r2, v_1) * vphi);
  sk.add(kAPhi, r_i + 1, z_k + 1, ring(ro_v, dz2, r2, r3, v_2) * vphi);
}

// (log(1+x) - x) / x without cancellation (series below |x| = 0.1)
__device__ __forceinline__ D m_of_d(D x) {
  if (fabs(x.v) < 0.1) {
    D s = cst(-1.0 / 12.0);
    const double c[] = {1.0 / 11.0, -0.1, 1.0 / 9.0, -0.125, 1.0 / 7.0, -1.0 / 6.0, 0.2, -0.25, 1.0 / 3.0, -0.5};
    for (double ck : c) s = ck + x * s;
    return x * s;
  }
  D l;
  l.v = log1p(x.v);
  for (int j = 0; j < kDN; j++) l.d[j] = x.d[j] / (1.0 + x.v);
  return (l - x) / x;
}

// simple_current_distribution, src/particles.cpp:333-474.  Values: the
// reference's expressions.  Derivatives: those of the same integrals written
// without cancellation in cell units (r = (i + a) dr, rho = i + a,
// x = drho / rho, M = (ln(1+x) - x) / x, N = -x / (1+x) - M):
//   J_z[i]   = q dzeta / (4 pi dr^2 dt) [ (1-a)(2i+a)/rho - drho/2 + i(i+1)/rho M ] / i
//   J_z[i+1] = q dzeta / (4 pi dr^2 dt) [ a(2i+1+a)/rho + drho/2 - i(i+1)/rho M ] / (i+1)
//   J_r[k]   = q / (2 pi dr dz dt) [ drho((1-b)/2 - dzeta/4) + i(i+1)/(2 rho) (drho (1-b)/rho_e + dzeta N) ] / (i+1/2)
//   J_r[k+1] = q / (2 pi dr dz dt) [ drho(b/2 + dzeta/4) + i(i+1)/(2 rho) (drho b/rho_e - dzeta N) ] / (i+1/2)
// (and the polynomial forms on the axis cell).  The reference's forms are
// differences of O(i) terms: their derivative loses ~1/x digits for nearly
// axial paths, which a relativistic beam's particles are.
template <class S>
__device__ void segment_d(const GridDesc &g, S &acc, D q, D radius_new, D longitude_new, D radius_old,
                          D longitude_old, int i_n, int k_n) {
  const double dr = g.dr, dz = g.dz, delta_t = g.dt;
  const double delta_r = radius_new.v - radius_old.v;
  const double delta_z = longitude_new.v - longitude_old.v;
  if ((fabs(delta_r) < kMnzl) || (fabs(delta_z) < kMnzl)) return;
  // values, as the reference computes them
  double vz0, vz1, vr0, vr1;
  {
    const double qv = q.v;
    double k = delta_r / delta_z;
    double b = radius_old.v;
    double lg = log((k * delta_z + b) / b);
    if (i_n >= 1) {
      double ssd = qv / (2 * kPi * (i_n * dr) * dr * dz * delta_t * 2 * dr);
      vz0 = ssd * (dr * delta_z - k * delta_z * delta_z / 2. - delta_z * b +
                   dr * dr / k * ((i_n + 0.5) * (i_n + 0.5) - 0.25) * lg);
      ssd = qv / (2 * kPi * ((i_n + 1) * dr) * dr * dz * delta_t * 2 * dr);
      vz1 = ssd * (k * delta_z * delta_z / 2. + delta_z * b + delta_z * dr +
                   dr * dr / k * (0.25 - (i_n + 0.5) * (i_n + 0.5)) * lg);
    } else {
      double ssd = qv / (2. * kPi * dr / 4. * dr * dz * delta_t * dr);
      vz0 = ssd * (dr * delta_z - k * delta_z * delta_z / 2. - delta_z * b);
      ssd = qv / (2 * kPi * dr * dr * dz * delta_t * 2 * dr);
      vz1 = ssd * (k * delta_z * delta_z / 2. + delta_z * dr + delta_z * b);
    }
    k = -delta_z / delta_r;
    const double r0 = (i_n + 0.5) * dr;
    const double r1 = radius_old.v;
    b = (k_n + 1.) * dz - longitude_old.v;
    const double ssd = qv / (2 * kPi * r0 * dr * dz * delta_t * dz);
    const double lr = log((radius_old.v + delta_r) / radius_old.v);
    vr0 = ssd * (r0 * k * delta_r + k / 2. * delta_r * (radius_old.v + delta_r / 2.) +
                 0.5 * delta_r * (b - k * (2 * r0 + r1)) +
                 delta_r * (b - k * r1) * (4 * r0 * r0 - dr * dr) / (8 * radius_old.v * (radius_old.v + delta_r)) +
                 (k * (r0 * r0 / 2. - dr * dr / 8.)) * lr);
    b = longitude_old.v - k_n * dz;
    vr1 = ssd * (-r0 * k * delta_r - k / 2. * delta_r * (radius_old.v + delta_r / 2.) +
                 0.5 * delta_r * (b + k * (2 * r0 + r1)) +
                 delta_r * (b + k * r1) * (4 * r0 * r0 - dr * dr) / (8 * radius_old.v * (radius_old.v + delta_r)) -
                 (k * (r0 * r0 / 2. - dr * dr / 8.)) * lr);
  }
  // derivatives, from the cancellation-free forms
  const double fi = (double)i_n;
  D a = radius_old / dr - fi;
  D bo = longitude_old / dz - (double)k_n;
  D drho = (radius_new - radius_old) / dr;
  D dzeta = (longitude_new - longitude_old) / dz;
  D jz0, jz1, jr0, jr1;
  if (i_n == 0) {
    jz0 = 8.0 * (1.0 - 0.5 * drho - a);
    jz1 = 0.5 * drho + 1.0 + a;
    jr0 = drho * (0.5 * (1.0 - bo) - 0.25 * dzeta) * 2.0;
    jr1 = drho * (0.5 * bo + 0.25 * dzeta) * 2.0;
  } else {
    D rs = fi + a;
    D re = rs + drho;
    D x = drho / rs;
    D M = m_of_d(x);
    D N = -(x / (1.0 + x)) - M;
    const double ii = fi * (fi + 1.0);
    jz0 = ((1.0 - a) * (2.0 * fi + a) / rs - 0.5 * drho + ii / rs * M) / fi;
    jz1 = (a * (2.0 * fi + 1.0 + a) / rs + 0.5 * drho - ii / rs * M) / (fi + 1.0);
    D cc = 0.5 * ii / rs;
    jr0 = (drho * (0.5 * (1.0 - bo) - 0.25 * dzeta) + cc * (drho * (1.0 - bo) / re + dzeta * N)) / (fi + 0.5);
    jr1 = (drho * (0.5 * bo + 0.25 * dzeta) + cc * (drho * bo / re - dzeta * N)) / (fi + 0.5);
  }
  D sz = q * dzeta / (4.0 * kPi * dr * dr * delta_t);
  D sr = q / (2.0 * kPi * dr * dz * delta_t);
  D out[4] = {sz * jz0, sz * jz1, sr * jr0, sr * jr1};
  out[0].v = vz0;
  out[1].v = vz1;
  out[2].v = vr0;
  out[3].v = vr1;
  acc.add(kAZ, i_n, k_n, out[0]);
  acc.add(kAZ, i_n + 1, k_n, out[1]);
  acc.add(kAR, i_n, k_n, out[2]);
  acc.add(kAR, i_n, k_n + 1, out[3]);
}

// strict_motion_weighting, src/particles.cpp:786-988
template <class S>
__device__ void strict_d(const GridDesc &g, S &acc, D q, D radius_new, D longitude_new, D radius_old,
                         D longitude_old) {
  const double dr = g.dr, dz = g.dz, delta_t = g.dt;
  int i_n = cell_number(radius_new.v, dr), k_n = cell_number(longitude_new.v, dz);
  int i_o = cell_number(radius_old.v, dr), k_o = cell_number(longitude_old.v, dz);
  if (i_n < 0) i_n = 0;
  if (k_n < 0) k_n = 0;
  if (i_o < 0) i_o = 0;
  if (k_o < 0) k_o = 0;
  bool still_r = fabs(radius_new.v - radius_old.v) < kMnzl;
  bool still_z = fabs(longitude_new.v - longitude_old.v) < kMnzl;
  if (still_r && still_z) return;
  if (still_r) {
    D value_part = 2. * kPi * radius_new * dr * dz;
    D r1 = radius_new - 0.5 * dr, r3 = radius_new + 0.5 * dr;
    D r2 = cst((i_n + 0.5) * dr);
    D wl, wu;
    if (i_n == 0)
      wl = q / (delta_t * kPi * dr * dr / 4.) * kPi * (r2 * r2 - r1 * r1) / value_part;
    else
      wl = q / (delta_t * 2. * kPi * i_n * dr * dr) * kPi * (r2 * r2 - r1 * r1) / value_part;
    wu = q / (delta_t * 2 * kPi * (i_n + 1) * dr * dr) * kPi * (r3 * r3 - r2 * r2) / value_part;
    int res_k = k_n - k_o;
    if (res_k == 0) {
      D delta_z = longitude_new - longitude_old;
      acc.add(kAZ, i_n, k_n, wl * delta_z);
      acc.add(kAZ, i_n + 1, k_n, wu * delta_z);
    } else if (res_k == 1) {
      D delta_z = k_n * dz - longitude_old;
      acc.add(kAZ, i_n, k_n - 1, wl * delta_z);
      acc.add(kAZ, i_n + 1, k_n - 1, wu * delta_z);
      delta_z = longitude_new - k_n * dz;
      acc.add(kAZ, i_n, k_n, wl * delta_z);
      acc.add(kAZ, i_n + 1, k_n, wu * delta_z);
    } else if (res_k == -1) {
      D delta_z = (k_n + 1) * dz - longitude_old;
      acc.add(kAZ, i_n, k_n + 1, wl * delta_z);
      acc.add(kAZ, i_n + 1, k_n + 1, wu * delta_z);
      delta_z = longitude_new - (k_n + 1) * dz;
      acc.add(kAZ, i_n, k_n, wl * delta_z);
      acc.add(kAZ, i_n + 1, k_n, wu * delta_z);
    }
  } else if (still_z) {
    int res_i = i_n - i_o;
    D left_delta_z = (k_n + 1) * dz - longitude_new;
    D right_delta_z = longitude_new - k_n * dz;
    auto part = [&](D delta_r, double r0, D rad_old) {
      D ssd = q / (kPi * 4. * r0 * dz * dz * dr * delta_t);
      return ssd * (delta_r - r0 * r0 / (rad_old + delta_r) + r0 * r0 / rad_old + dr * dr / (4. * (rad_old + delta_r)) -
                    dr * dr / (4. * rad_old));
    };
    if (res_i == 0) {
      D wj = part(radius_new - radius_old, (i_n + 0.5) * dr, radius_old);
      acc.add(kAR, i_n, k_n, wj * left_delta_z);
      acc.add(kAR, i_n, k_n + 1, wj * right_delta_z);
    } else if (res_i == 1) {
      D wj = part(i_n * dr - radius_old, (i_n - 0.5) * dr, radius_old);
      acc.add(kAR, i_n - 1, k_n, wj * left_delta_z);
      acc.add(kAR, i_n - 1, k_n + 1, wj * right_delta_z);
      // the reference's `r0 * r0 / i_n * dr` (src/particles.cpp:941)
      D delta_r = radius_new - i_n * dr;
      const double r0 = (i_n + 0.5) * dr;
      D ssd = q / (kPi * 4. * r0 * dz * dz * dr * delta_t);
      wj = ssd * (delta_r - r0 * r0 / (i_n * dr + delta_r) + r0 * r0 / i_n * dr + dr * dr / (4. * (i_n * dr + delta_r)) -
                  dr * dr / (4. * i_n * dr));
      acc.add(kAR, i_n, k_n, wj * left_delta_z);
      acc.add(kAR, i_n, k_n + 1, wj * right_delta_z);
    } else if (res_i == -1) {
      D wj = part((i_n + 1) * dr - radius_old, (i_n + 1.5) * dr, radius_old);
      acc.add(kAR, i_n + 1, k_n, wj * left_delta_z);
      acc.add(kAR, i_n + 1, k_n + 1, wj * right_delta_z);
      D delta_r = radius_new - (i_n + 1) * dr;
      const double r0 = (i_n + 0.5) * dr;
      D ssd = q / (kPi * 4. * r0 * dz * dz * dr * delta_t);
      wj = ssd * (delta_r - r0 * r0 / ((i_n + 1) * dr + delta_r) + r0 * r0 / (i_n + 1) * dr +
                  dr * dr / (4. * ((i_n + 1) * dr + delta_r)) - dr * dr / (4. * (i_n + 1) * dr));
      acc.add(kAR, i_n, k_n, wj * left_delta_z);
      acc.add(kAR, i_n, k_n + 1, wj * right_delta_z);
    }
  }
}

struct SegD {
  D rn, zn, ro, zo;
  int i, k;
};

// current_distribution, src/particles.cpp:480-685
template <class S>
__device__ void deposit_j_d(const GridDesc &g, S &acc, D q, D pr, D pz, D por, D poz) {
  const double dr = g.dr, dz = g.dz;
  int i_n = cell_number(pr.v, dr), k_n = cell_number(pz.v, dz);
  int i_o = cell_number(por.v, dr), k_o = cell_number(poz.v, dz);
  if (i_n < 0) i_n = 0;
  if (k_n < 0) k_n = 0;
  if (i_o < 0) i_o = 0;
  if (k_o < 0) k_o = 0;
  if (por.v == (i_o + 1) * dr) i_o = i_n;
  if (poz.v == (k_o + 1) * dz) k_o = k_n;
  if (pr.v == (i_n + 1) * dr) i_n = i_o;
  if (pz.v == (k_n + 1) * dz) k_n = k_o;
  int res_cell = abs(i_n - i_o) + abs(k_n - k_o);
  if ((fabs(pr.v - por.v) < kMnzl) || (fabs(pz.v - poz.v) < kMnzl)) {
    strict_d(g, acc, q, pr, pz, por, poz);
    return;
  }
  if (res_cell == 0) {  // the common case, without the segment list (which lives in scratch memory)
    segment_d(g, acc, q, pr, pz, por, poz, i_n, k_n);
    return;
  }
  SegD s[3];
  int ns = 0;
  if (res_cell == 1) {
    if ((i_n != i_o) && (k_n == k_o)) {
      if (por.v > (i_n + 1) * dr) {
        D a = (por - pr) / (poz - pz);
        D r_b = cst((i_n + 1) * dr);
        D z_b = pz + (r_b - pr) / a;
        s[0] = SegD{r_b, z_b, por, poz, i_n + 1, k_n};
        s[1] = SegD{pr, pz, r_b, z_b, i_n, k_n};
      } else {
        D a = (pr - por) / (pz - poz);
        D r_b = cst(i_n * dr);
        D z_b = poz + (r_b - por) / a;
        s[0] = SegD{r_b, z_b, por, poz, i_n - 1, k_n};
        s[1] = SegD{pr, pz, r_b, z_b, i_n, k_n};
      }
      ns = 2;
    } else if ((i_n == i_o) && (k_n != k_o)) {
      if (poz.v < k_n * dz) {
        D z_b = cst(k_n * dz);
        D a = (pr - por) / (pz - poz);
        D r_b = por + a * (z_b - poz);
        s[0] = SegD{r_b, z_b, por, poz, i_n, k_n - 1};
        s[1] = SegD{pr, pz, r_b, z_b, i_n, k_n};
      } else {
        D z_b = cst((k_n + 1) * dz);
        D a = (por - pr) / (poz - pz);
        D r_b = pr + a * (z_b - pz);
        s[0] = SegD{r_b, z_b, por, poz, i_n, k_n + 1};
        s[1] = SegD{pr, pz, r_b, z_b, i_n, k_n};
      }
      ns = 2;
    }
  } else if (res_cell == 2) {
    if (i_o < i_n) {
      if (k_o < k_n) {
        D a = (pr - por) / (pz - poz);
        D r1 = cst(i_n * dr);
        D z1 = poz + (r1 - por) / a;
        D z2 = cst(k_n * dz);
        D r2 = por + (z2 - poz) * a;
        if (z1.v < k_n * dz) {
          s[0] = SegD{r1, z1, por, poz, i_n - 1, k_n - 1};
          s[1] = SegD{r2, z2, r1, z1, i_n, k_n - 1};
          s[2] = SegD{pr, pz, r2, z2, i_n, k_n};
          ns = 3;
        } else if (z1.v > k_n * dz) {
          s[0] = SegD{r2, z2, por, poz, i_n - 1, k_n - 1};
          s[1] = SegD{r1, z1, r2, z2, i_n - 1, k_n};
          s[2] = SegD{pr, pz, r1, z1, i_n, k_n};
          ns = 3;
        }
      } else {
        D a = (pr - por) / (pz - poz);
        D r1 = cst(i_n * dr);
        D z1 = poz + (r1 - por) / a;
        D z2 = cst((k_n + 1) * dz);
        D r2 = por + -(poz - z2) * a;
        if (z1.v > (k_n + 1) * dz) {
          s[0] = SegD{r1, z1, por, poz, i_n - 1, k_n + 1};
          s[1] = SegD{r2, z2, r1, z1, i_n, k_n + 1};
          s[2] = SegD{pr, pz, r2, z2, i_n, k_n};
          ns = 3;
        } else if (z1.v < (k_n + 1) * dz) {
          s[0] = SegD{r2, z2, por, poz, i_n - 1, k_n + 1};
          s[1] = SegD{r1, z1, r2, z2, i_n - 1, k_n};
          s[2] = SegD{pr, pz, r1, z1, i_n, k_n};
          ns = 3;
        }
      }
    } else if (i_o > i_n) {
      if (k_o < k_n) {
        D a = (pr - por) / (pz - poz);
        D r1 = cst((i_n + 1) * dr);
        D z1 = poz + -(por - r1) / a;
        D z2 = cst(k_n * dz);
        D r2 = por - -(z2 - poz) * a;
        if (z1.v < (k_n)*dz) {
          s[0] = SegD{r1, z1, por, poz, i_n + 1, k_n - 1};
          s[1] = SegD{r2, z2, r1, z1, i_n, k_n - 1};
          s[2] = SegD{pr, pz, r2, z2, i_n, k_n};
          ns = 3;
        } else if (z1.v > (k_n)*dz) {
          s[0] = SegD{r2, z2, por, poz, i_n + 1, k_n - 1};
          s[1] = SegD{r1, z1, r2, z2, i_n + 1, k_n};
          s[2] = SegD{pr, pz, r1, z1, i_n, k_n};
          ns = 3;
        }
      } else if (k_o > k_n) {
        D a = (por - pr) / (poz - pz);
        D r1 = cst((i_n + 1) * dr);
        D z1 = pz + (r1 - pr) / a;
        D z2 = cst((k_n + 1) * dz);
        D r2 = pr + (z2 - pz) * a;
        if (z1.v > (k_n + 1) * dz) {
          s[0] = SegD{r1, z1, por, poz, i_n + 1, k_n + 1};
          s[1] = SegD{r2, z2, r1, z1, i_n, k_n + 1};
          s[2] = SegD{pr, pz, r2, z2, i_n, k_n};
          ns = 3;
        } else if (z1.v < (k_n + 1) * dz) {
          s[0] = SegD{r2, z2, por, poz, i_n + 1, k_n + 1};
          s[1] = SegD{r1, z1, r2, z2, i_n + 1, k_n};
          s[2] = SegD{pr, pz, r1, z1, i_n, k_n};
          ns = 3;
        }
      }
    }
  }
#pragma nounroll
  for (int t = 0; t < ns; t++) segment_d(g, acc, q, s[t].rn, s[t].zn, s[t].ro, s[t].zo, s[t].i, s[t].k);
}

// reflection_single / Bunch::reflection_single on the half-step point
__device__ __forceinline__ void walls_d(const GridDesc &g, bool is_bunch, D &x, D &z, D &v0, D &v2, bool &alive) {
  const double dr = g.dr, dz = g.dz;
  const double radius_wall = g.r_size - dr / 2., longitude_wall = g.z_size - dz / 2.;
  const double half_dr = dr / 2., half_dz = dz / 2.;
  if (is_bunch) {
    if (x.v > radius_wall) alive = false;
    if (z.v > longitude_wall) alive = false;
  } else {
    if (x.v > radius_wall) {
      x = radius_wall * 2. - x;
      v0 = -v0;
    }
    if (z.v > longitude_wall) {
      z = longitude_wall * 2. - z;
      v2 = -v2;
    }
  }
  if (x.v < half_dr) {
    x = dr - x;
    v0 = -v0;
  }
  if (z.v < half_dz) {
    z = dz - z;
    v2 = -v2;
  }
}

// vector-Jacobian product of one particle's step
// the vector-Jacobian product of one living particle's step
__device__ __forceinline__ void adj_particle(const GridDesc &g, const AdjFields &f, const AdjGrid &lg,
                                             const AdjSpecies &s, int pusher, long long i, const AdjBox &b) {
  const double r0v = s.r[i], z0v = s.z[i];
  const double lam_r = s.lr[i], lam_z = s.lz[i], lam_v0 = s.lvr[i], lam_v1 = s.lvphi[i], lam_v2 = s.lvz[i];
  double lin[kInputs] = {};
  for (int pass = 0; pass < kPasses; pass++) {
    AdjSink sk;
    sk.g = &g;
    sk.ljr = lg.ljr;
    sk.ljphi = lg.ljphi;
    sk.ljz = lg.ljz;
    for (int j = 0; j < kDN; j++) sk.acc[j] = 0.0;
    D r = seed(r0v, 0, pass), z = seed(z0v, 1, pass);
    D v[3] = {seed(s.vr[i], 2, pass), seed(s.vphi[i], 3, pass), seed(s.vz[i], 4, pass)};
    D w = seed(s.w[i], 11, pass);
    D q = s.charge * w, m = s.mass * w;
    const D r_old = r, z_old = z;
    D e[3], h[3];
    e[0] = gather_d(g, f.er, r, z, true, false) + seed(0.0, 5, pass);
    e[1] = gather_d(g, f.ephi, r, z, false, false) + seed(0.0, 6, pass);
    e[2] = gather_d(g, f.ez, r, z, false, true) + seed(0.0, 7, pass);
    h[0] = gather_d(g, f.hhr, r, z, false, true) + seed(0.0, 8, pass);
    h[1] = gather_d(g, f.hhphi, r, z, true, true) + seed(0.0, 9, pass);
    h[2] = gather_d(g, f.hhz, r, z, true, false) + seed(0.0, 10, pass);
    if (pusher == 1)
      vay_d(q, m, g.dt, e, h, v);
    else
      boris_d(q, m, g.dt, e, h, v);
    const double half_dt = g.dt / 2.;
    // first half step and rotation (step_v)
    D x = r + v[0] * half_dt;
    D y = 0.0 + v[1] * half_dt;
    z = z + v[2] * half_dt;
    r = dsqrt(x * x + y * y);
    // full_current_distribution
    jphi_d(g, sk, q, r, z, v[1]);
    x = r + v[0] * half_dt;
    y = 0.0 + v[1] * half_dt;
    z = z + v[2] * half_dt;
    bool alive = true;
    walls_d(g, s.is_bunch != 0, x, z, v[0], v[2], alive);
    D rr = dsqrt(x * x + y * y);
    D st = y / rr;
    r = rr;
    if (alive) {
      deposit_j_d(g, sk, q, r, z, r_old, z_old);
      D ct = dsqrt(1 - st * st);
      D u2 = ct * v[0] - st * v[1];
      D w2 = st * v[0] + ct * v[1];
      v[0] = u2;
      v[1] = w2;
      for (int j = 0; j < kDN; j++)
        sk.acc[j] += lam_r * r.d[j] + lam_z * z.d[j] + lam_v0 * v[0].d[j] + lam_v1 * v[1].d[j] + lam_v2 * v[2].d[j];
    }
#pragma unroll
    for (int o = 0; o < kInputs; o++) lin[o] = o / kDN == pass ? sk.acc[o % kDN] : lin[o];
  }
  s.lr[i] = lin[0];
  s.lz[i] = lin[1];
  s.lvr[i] = lin[2];
  s.lvphi[i] = lin[3];
  s.lvz[i] = lin[4];
  s.lw[i] += lin[11];
  box_scatter(g, b, 0, lg.ler, r0v, z0v, true, false, lin[5]);
  box_scatter(g, b, 1, lg.lephi, r0v, z0v, false, false, lin[6]);
  box_scatter(g, b, 2, lg.lez, r0v, z0v, false, true, lin[7]);
  box_scatter(g, b, 3, lg.lhr, r0v, z0v, false, true, lin[8]);
  box_scatter(g, b, 4, lg.lhphi, r0v, z0v, true, true, lin[9]);
  box_scatter(g, b, 5, lg.lhz, r0v, z0v, true, false, lin[10]);
}

__global__ void __launch_bounds__(256) k_adj_particles(GridDesc g, AdjFields f, AdjGrid lg, AdjSpecies s,
                                                       int pusher) {
  __shared__ double box[kABComp * kABR * kABK];
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < s.n && s.alive[i] == 1;
  // the box anchor: the smallest cell of the block's living particles
  int ci = INT_MAX, ck = INT_MAX;
  if (live) {
    ci = max(cell_number(s.r[i], g.dr), 0);
    ck = max(cell_number(s.z[i], g.dz), 0);
  }
  block_min2(ci, ck);
  if (ci == INT_MAX) ci = ck = 0;
  for (int t = threadIdx.x; t < kABComp * kABR * kABK; t += blockDim.x) box[t] = 0.0;
  __syncthreads();
  const AdjBox b{box, ci - 1, ck - 2};
  if (live) adj_particle(g, f, lg, s, pusher, i, b);
  else if (i < s.n) s.lr[i] = s.lz[i] = s.lvr[i] = s.lvphi[i] = s.lvz[i] = 0.0;
  __syncthreads();
  double *const lf[kABComp] = {lg.ler, lg.lephi, lg.lez, lg.lhr, lg.lhphi, lg.lhz};
  for (int t = threadIdx.x; t < kABComp * kABR * kABK; t += blockDim.x) {
    const double v = box[t];
    if (v == 0.0) continue;
    const int comp = t / (kABR * kABK), rem = t - comp * kABR * kABK;
    const int gi = b.bi + rem / kABK, gk = b.bk + rem % kABK;
    // nodes off the grid never receive a stencil (gather_w clamps its cell)
    if ((unsigned)gi < (unsigned)g.nr && (unsigned)gk < (unsigned)g.nz)
      unsafeAtomicAdd(lf[comp] + (size_t)gi * (size_t)g.nz + gk, v);
  }
}

// ------------------------------------------------ tangent-linear (forward mode)
// The same particle step in dual numbers, seeded with kDN tangent directions
// of the state, of the gathered fields and of the weight instead of unit
// vectors: the derivative parts of the outputs are the tangents of the new
// state and of every deposit (J-dot, scattered with atomics).  Together with
// the forward field updates applied to tangent grids this is the
// tangent-linear model of the loop, the exact transpose of what the adjoint
// above computes; <dphi/dw, dw> from the adjoint must equal the tangent of phi
// along dw to round-off (the dot-product test).

struct TlmFields {  // tangents of E^n and Hh^n, one set per direction
  const double *er[kDN], *ephi[kDN], *ez[kDN], *hhr[kDN], *hhphi[kDN], *hhz[kDN];
};
struct TlmJ {  // tangents of J^n, one set per direction
  double *jr[kDN], *jphi[kDN], *jz[kDN];
};
struct TlmSpecies {
  const double *r, *z, *vr, *vphi, *vz, *w;  // state at step n (values)
  const uint8_t *alive;
  double *tr, *tz, *tvr, *tvphi, *tvz;  // kDN x n tangents: in at step n, out at n+1
  const double *tw;                     // kDN x n tangents of the weight
  double charge, mass;
  long long n;
  int is_bunch;
};

// deposit tangents: every contribution v at node (i, k) adds its derivative
// part (one per direction) to the tangent J grids
struct TlmSink {
  const GridDesc *g;
  TlmJ j;
  __device__ __forceinline__ void add(int comp, int i, int k, D v) {
    if ((unsigned)i >= (unsigned)g->nr || (unsigned)k >= (unsigned)g->nz) return;
    const size_t p = (size_t)i * (size_t)g->nz + (size_t)k;
#pragma unroll
    for (int d = 0; d < kDN; d++) {
      double *J = comp == kAR ? j.jr[d] : comp == kAPhi ? j.jphi[d] : j.jz[d];
      if (v.d[d] != 0.0) unsafeAtomicAdd(J + p, v.d[d]);
    }
  }
};

// the gather of the tangent fields at the particle's (value) position: the
// gather is linear in the field, so its tangent along the field is the same
// weights applied to the tangent grid
__device__ __forceinline__ double gather_value(const GridDesc &g, const double *F, double r, double z, bool lo,
                                               bool zhalf) {
  GatherW gw = gather_w(g, cst(r), cst(z), lo, zhalf);
  const double *row0 = F + (size_t)gw.ir * (size_t)g.nz + gw.kz;
  const double *row1 = row0 + g.nz;
  return gw.w[0].v * row0[0] + gw.w[1].v * row1[0] + gw.w[2].v * row0[1] + gw.w[3].v * row1[1];
}

__global__ void __launch_bounds__(256) k_tlm_particles(GridDesc g, AdjFields f, TlmFields tf, TlmJ tj,
                                                       TlmSpecies s, int pusher) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= s.n) return;
  const long long n = s.n;
  if (s.alive[i] != 1) {
    // dead (or padding) slots keep zero tangents; a particle injected later
    // starts with zero tangents, as its injected state does not depend on w
#pragma unroll
    for (int d = 0; d < kDN; d++) s.tr[d * n + i] = s.tz[d * n + i] = s.tvr[d * n + i] = s.tvphi[d * n + i] =
        s.tvz[d * n + i] = 0.0;
    return;
  }
  auto dual = [&](double v, const double *t) {
    D x;
    x.v = v;
#pragma unroll
    for (int d = 0; d < kDN; d++) x.d[d] = t[d * n + i];
    return x;
  };
  TlmSink sk{&g, tj};
  D r = dual(s.r[i], s.tr), z = dual(s.z[i], s.tz);
  D v[3] = {dual(s.vr[i], s.tvr), dual(s.vphi[i], s.tvphi), dual(s.vz[i], s.tvz)};
  D w = dual(s.w[i], s.tw);
  D q = s.charge * w, m = s.mass * w;
  const D r_old = r, z_old = z;
  const double r0v = r.v, z0v = z.v;
  D e[3], h[3];
  e[0] = gather_d(g, f.er, r, z, true, false);
  e[1] = gather_d(g, f.ephi, r, z, false, false);
  e[2] = gather_d(g, f.ez, r, z, false, true);
  h[0] = gather_d(g, f.hhr, r, z, false, true);
  h[1] = gather_d(g, f.hhphi, r, z, true, true);
  h[2] = gather_d(g, f.hhz, r, z, true, false);
#pragma unroll
  for (int d = 0; d < kDN; d++) {
    e[0].d[d] += gather_value(g, tf.er[d], r0v, z0v, true, false);
    e[1].d[d] += gather_value(g, tf.ephi[d], r0v, z0v, false, false);
    e[2].d[d] += gather_value(g, tf.ez[d], r0v, z0v, false, true);
    h[0].d[d] += gather_value(g, tf.hhr[d], r0v, z0v, false, true);
    h[1].d[d] += gather_value(g, tf.hhphi[d], r0v, z0v, true, true);
    h[2].d[d] += gather_value(g, tf.hhz[d], r0v, z0v, true, false);
  }
  if (pusher == 1)
    vay_d(q, m, g.dt, e, h, v);
  else
    boris_d(q, m, g.dt, e, h, v);
  const double half_dt = g.dt / 2.;
  D x = r + v[0] * half_dt;
  D y = 0.0 + v[1] * half_dt;
  z = z + v[2] * half_dt;
  r = dsqrt(x * x + y * y);
  jphi_d(g, sk, q, r, z, v[1]);
  x = r + v[0] * half_dt;
  y = 0.0 + v[1] * half_dt;
  z = z + v[2] * half_dt;
  bool alive = true;
  walls_d(g, s.is_bunch != 0, x, z, v[0], v[2], alive);
  D rr = dsqrt(x * x + y * y);
  D st = y / rr;
  r = rr;
  if (alive) {
    deposit_j_d(g, sk, q, r, z, r_old, z_old);
    D ct = dsqrt(1 - st * st);
    D u2 = ct * v[0] - st * v[1];
    D w2 = st * v[0] + ct * v[1];
    v[0] = u2;
    v[1] = w2;
  }
#pragma unroll
  for (int d = 0; d < kDN; d++) {
    const bool keep = alive;
    s.tr[d * n + i] = keep ? r.d[d] : 0.0;
    s.tz[d * n + i] = keep ? z.d[d] : 0.0;
    s.tvr[d * n + i] = keep ? v[0].d[d] : 0.0;
    s.tvphi[d * n + i] = keep ? v[1].d[d] : 0.0;
    s.tvz[d * n + i] = keep ? v[2].d[d] : 0.0;
  }
}

// tangent of the objective: phi_dot[d] += scale * sum 2 E_z(0, k) E_z-dot_d(0, k)
// over k0..k1, and phi += scale * sum E_z(0,k)^2 (one block)
__global__ void __launch_bounds__(256) k_tlm_objective(GridDesc g, const double *ez, const double *ezd0,
                                                       const double *ezd1, const double *ezd2, const double *ezd3,
                                                       int k0, int k1, double scale, double *phi) {
  __shared__ double part[5][256];
  double acc[5] = {0, 0, 0, 0, 0};
  const double *ezd[4] = {ezd0, ezd1, ezd2, ezd3};
  for (int k = k0 + threadIdx.x; k <= k1; k += blockDim.x) {
    const double e = ez[k];
    acc[0] += scale * e * e;
#pragma unroll
    for (int d = 0; d < 4; d++) acc[1 + d] += scale * 2.0 * e * ezd[d][k];
  }
#pragma unroll
  for (int a = 0; a < 5; a++) part[a][threadIdx.x] = acc[a];
  __syncthreads();
  for (int s2 = 128; s2 > 0; s2 >>= 1) {
    if ((int)threadIdx.x < s2)
#pragma unroll
      for (int a = 0; a < 5; a++) part[a][threadIdx.x] += part[a][threadIdx.x + s2];
    __syncthreads();
  }
  if (threadIdx.x < 5) phi[threadIdx.x] += part[threadIdx.x][0];
}

// ------------------------------------------------------------- the fields

struct AdjE {  // adjoints of E, H (after step), Hh, J and the coefficients
  double *ler, *lephi, *lez, *lhr, *lhphi, *lhz, *lhhr, *lhhphi, *lhhz, *ljr, *ljphi, *ljz;
  const double *ke, *kh;
};

__device__ __forceinline__ size_t at(const GridDesc &g, int i, int k) { return (size_t)i * (size_t)g.nz + (size_t)k; }

// which E entries EField::calc_field updates (src/eField.cpp:99-168)
__device__ __forceinline__ bool upd_er(const GridDesc &g, int i, int k) { return i <= g.nr - 2 && k >= 1 && k <= g.nz - 2; }
__device__ __forceinline__ bool upd_ephi(const GridDesc &g, int i, int k) {
  return i >= 1 && i <= g.nr - 2 && k >= 1 && k <= g.nz - 2;
}
__device__ __forceinline__ bool upd_ez(const GridDesc &g, int i, int k) { return i <= g.nr - 2 && k <= g.nz - 2; }

// step 1a: lambda_H += C^T (Kh lambda_E), lambda_J = -Kh lambda_E (pull form:
// each H node sums the E nodes whose update reads it)
__global__ void __launch_bounds__(256) k_adj_e_curl(GridDesc g, AdjE a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (k >= g.nz) return;
  const double dr = g.dr, dz = g.dz;
  const size_t p = at(g, i, k);
  // J adjoints
  a.ljr[p] = upd_er(g, i, k) ? -a.kh[p] * a.ler[p] : 0.0;
  a.ljphi[p] = upd_ephi(g, i, k) ? -a.kh[p] * a.lephi[p] : 0.0;
  a.ljz[p] = upd_ez(g, i, k) ? -a.kh[p] * a.lez[p] : 0.0;
  // H_phi(i, k): E_r(i, k) has -Kh (H_phi(i,k) - H_phi(i,k-1)) / dz, E_r(i, k+1)
  // the same with k+1; E_z(i, k) has +Kh ((H_phi(i,k) - H_phi(i-1,k)) / dr +
  // (H_phi(i,k) + H_phi(i-1,k)) / (2 dr i)) for i >= 1 and +Kh 4/dr H_phi(0,k)
  // on the axis; E_z(i+1, k) the same with i+1
  double lhphi = 0.0;
  if (upd_er(g, i, k)) lhphi += -a.kh[p] / dz * a.ler[p];
  if (k + 1 < g.nz && upd_er(g, i, k + 1)) lhphi += a.kh[p + 1] / dz * a.ler[p + 1];
  if (upd_ez(g, i, k)) {
    if (i == 0)
      lhphi += a.kh[p] * (4.0 / dr) * a.lez[p];
    else
      lhphi += a.kh[p] * (1.0 / dr + 1.0 / (2.0 * dr * i)) * a.lez[p];
  }
  if (i + 1 < g.nr && upd_ez(g, i + 1, k)) {
    const size_t pu = p + g.nz;
    lhphi += a.kh[pu] * (-1.0 / dr + 1.0 / (2.0 * dr * (i + 1))) * a.lez[pu];
  }
  a.lhphi[p] += lhphi;
  // H_r(i, k): E_phi(i, k) has +Kh (H_r(i,k) - H_r(i,k-1)) / dz, E_phi(i, k+1) -
  double lhr = 0.0;
  if (upd_ephi(g, i, k)) lhr += a.kh[p] / dz * a.lephi[p];
  if (k + 1 < g.nz && upd_ephi(g, i, k + 1)) lhr += -a.kh[p + 1] / dz * a.lephi[p + 1];
  a.lhr[p] += lhr;
  // H_z(i, k): E_phi(i, k) has -Kh (H_z(i,k) - H_z(i-1,k)) / dr, E_phi(i+1, k) +
  double lhz = 0.0;
  if (upd_ephi(g, i, k)) lhz += -a.kh[p] / dr * a.lephi[p];
  if (i + 1 < g.nr && upd_ephi(g, i + 1, k)) lhz += a.kh[p + g.nz] / dr * a.lephi[p + g.nz];
  a.lhz[p] += lhz;
}

// step 1b: lambda_E <- Ke lambda_E on the updated entries
__global__ void __launch_bounds__(256) k_adj_e_scale(GridDesc g, AdjE a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (k >= g.nz) return;
  const size_t p = at(g, i, k);
  if (upd_er(g, i, k)) a.ler[p] *= a.ke[p];
  if (upd_ephi(g, i, k)) a.lephi[p] *= a.ke[p];
  if (upd_ez(g, i, k)) a.lez[p] *= a.ke[p];
}

// step 3a: lambda_alpha = lambda_H + lambda_Hh / 2 (H_z: -(...)), pulled into
// lambda_E through A^T (HField::calc_field, src/hField.cpp:93-157):
//   H_r(i,k),   i <= nr-1, k <= nz-2: alpha = dt (E_phi(i,k+1) - E_phi(i,k)) / (dz mu0)
//   H_phi(i,k), i <= nr-2, k <= nz-2: alpha = dt ((E_z(i+1,k) - E_z(i,k)) / dr - (E_r(i,k+1) - E_r(i,k)) / dz) / mu0
//   H_z(i,k),   i <= nr-2, k <= nz-2: H_z -= alpha, alpha = dt ((E_phi(i+1,k) + E_phi(i,k)) / (2 dr (i+1/2))
//                                                   + (E_phi(i+1,k) - E_phi(i,k)) / dr) / mu0
__global__ void __launch_bounds__(256) k_adj_h(GridDesc g, AdjE a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (k >= g.nz) return;
  const double dt = g.dt, dr = g.dr, dz = g.dz, mu = kMagnConst;
  const size_t p = at(g, i, k);
  auto la_r = [&](int ii, int kk) {  // lambda_alpha of H_r at (ii, kk), 0 outside its update domain
    if (ii < 0 || ii > g.nr - 1 || kk < 0 || kk > g.nz - 2) return 0.0;
    const size_t q = at(g, ii, kk);
    return a.lhr[q] + 0.5 * a.lhhr[q];
  };
  auto la_phi = [&](int ii, int kk) {
    if (ii < 0 || ii > g.nr - 2 || kk < 0 || kk > g.nz - 2) return 0.0;
    const size_t q = at(g, ii, kk);
    return a.lhphi[q] + 0.5 * a.lhhphi[q];
  };
  auto la_z = [&](int ii, int kk) {
    if (ii < 0 || ii > g.nr - 2 || kk < 0 || kk > g.nz - 2) return 0.0;
    const size_t q = at(g, ii, kk);
    return -(a.lhz[q] + 0.5 * a.lhhz[q]);
  };
  // E_phi(i,k): H_r(i,k-1) +dt/(dz mu), H_r(i,k) -dt/(dz mu);
  // H_z(i-1,k): dt/mu (1/(2 dr (i-1/2)) + 1/dr); H_z(i,k): dt/mu (1/(2 dr (i+1/2)) - 1/dr)
  double le_phi = dt / (dz * mu) * (la_r(i, k - 1) - la_r(i, k));
  if (i >= 1) le_phi += dt / mu * (1.0 / (2.0 * dr * (i - 0.5)) + 1.0 / dr) * la_z(i - 1, k);
  le_phi += dt / mu * (1.0 / (2.0 * dr * (i + 0.5)) - 1.0 / dr) * la_z(i, k);
  // E_z(i,k): H_phi(i-1,k) +dt/(dr mu), H_phi(i,k) -dt/(dr mu)
  double le_z = dt / (dr * mu) * (la_phi(i - 1, k) - la_phi(i, k));
  // E_r(i,k): H_phi(i,k-1) -dt/(dz mu), H_phi(i,k) +dt/(dz mu)
  double le_r = dt / (dz * mu) * (la_phi(i, k) - la_phi(i, k - 1));
  a.ler[p] += le_r;
  a.lephi[p] += le_phi;
  a.lez[p] += le_z;
}

// step 3b: lambda_H <- lambda_H + lambda_Hh, lambda_Hh <- 0
__global__ void __launch_bounds__(256) k_adj_h_fin(GridDesc g, AdjE a) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (k >= g.nz) return;
  const size_t p = at(g, i, k);
  a.lhr[p] += a.lhhr[p];
  a.lhphi[p] += a.lhhphi[p];
  a.lhz[p] += a.lhhz[p];
  a.lhhr[p] = 0.0;
  a.lhhphi[p] = 0.0;
  a.lhhz[p] = 0.0;
}

// objective seed: lambda_Ez(0, k) += scale * 2 E_z(0, k) on k0..k1, and the
// objective's sum scale * sum E_z(0,k)^2 into *phi (one block)
__global__ void __launch_bounds__(256) k_adj_seed(GridDesc g, const double *ez, double *lez, int k0, int k1, double scale,
                                                  double *phi) {
  __shared__ double part[256];
  double acc = 0.0;
  for (int k = k0 + threadIdx.x; k <= k1; k += blockDim.x) {
    const double e = ez[k];  // row 0
    lez[k] += scale * 2.0 * e;
    acc += scale * e * e;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) *phi += part[0];
}

}  // namespace

// ------------------------------------------------------------- launchers

struct AdjLaunch {
  GridDesc g;
  AdjE e;
};

void launch_adj_fields_pre(const GridDesc &g, double *const *lam, const double *ke, const double *kh,
                           hipStream_t st) {
  AdjE a{lam[0], lam[1], lam[2], lam[3], lam[4], lam[5], lam[6], lam[7], lam[8], lam[9], lam[10], lam[11], ke, kh};
  dim3 grid((g.nz + 255) / 256, g.nr);
  hipLaunchKernelGGL(k_adj_e_curl, grid, dim3(256), 0, st, g, a);
  hipLaunchKernelGGL(k_adj_e_scale, grid, dim3(256), 0, st, g, a);
}

void launch_adj_fields_post(const GridDesc &g, double *const *lam, const double *ke, const double *kh,
                            hipStream_t st) {
  AdjE a{lam[0], lam[1], lam[2], lam[3], lam[4], lam[5], lam[6], lam[7], lam[8], lam[9], lam[10], lam[11], ke, kh};
  dim3 grid((g.nz + 255) / 256, g.nr);
  hipLaunchKernelGGL(k_adj_h, grid, dim3(256), 0, st, g, a);
  hipLaunchKernelGGL(k_adj_h_fin, grid, dim3(256), 0, st, g, a);
}

void launch_adj_particles(const GridDesc &g, const double *const *fields6, double *const *lam, const double *r,
                          const double *z, const double *vr, const double *vphi, const double *vz, const double *w,
                          const uint8_t *alive, double *const *lstate, double *lw, double charge, double mass,
                          long long n, int is_bunch, int pusher, hipStream_t st) {
  if (n <= 0) return;
  AdjFields f{fields6[0], fields6[1], fields6[2], fields6[3], fields6[4], fields6[5]};
  AdjGrid lg{lam[9], lam[10], lam[11], lam[0], lam[1], lam[2], lam[6], lam[7], lam[8]};
  AdjSpecies s{r, z, vr, vphi, vz, w, alive, lstate[0], lstate[1], lstate[2], lstate[3], lstate[4], lw,
               charge, mass, n, is_bunch};
  hipLaunchKernelGGL(k_adj_particles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, f, lg, s, pusher);
}

void launch_adj_seed(const GridDesc &g, const double *ez, double *lez, int k0, int k1, double scale, double *phi,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_adj_seed, dim3(1), dim3(256), 0, st, g, ez, lez, k0, k1, scale, phi);
}

void launch_tlm_particles(const GridDesc &g, const double *const *fields6, double *const *const *tgrid,
                          const double *r, const double *z, const double *vr, const double *vphi, const double *vz,
                          const double *w, const uint8_t *alive, double *const *tstate, const double *tw,
                          double charge, double mass, long long n, int is_bunch, int pusher, hipStream_t st) {
  if (n <= 0) return;
  AdjFields f{fields6[0], fields6[1], fields6[2], fields6[3], fields6[4], fields6[5]};
  TlmFields tf;
  TlmJ tj;
  for (int d = 0; d < kDN; d++) {
    double *const *a = tgrid[d];
    tf.er[d] = a[kER];
    tf.ephi[d] = a[kEPhi];
    tf.ez[d] = a[kEZ];
    tf.hhr[d] = a[kHHR];
    tf.hhphi[d] = a[kHHPhi];
    tf.hhz[d] = a[kHHZ];
    tj.jr[d] = a[kJR];
    tj.jphi[d] = a[kJPhi];
    tj.jz[d] = a[kJZ];
  }
  TlmSpecies s{r, z, vr, vphi, vz, w, alive, tstate[0], tstate[1], tstate[2], tstate[3], tstate[4], tw,
               charge, mass, n, is_bunch};
  hipLaunchKernelGGL(k_tlm_particles, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, g, f, tf, tj, s, pusher);
}

void launch_tlm_objective(const GridDesc &g, const double *ez, double *const *const *tgrid, int k0, int k1,
                          double scale, double *phi5, hipStream_t st) {
  hipLaunchKernelGGL(k_tlm_objective, dim3(1), dim3(256), 0, st, g, ez, tgrid[0][kEZ], tgrid[1][kEZ],
                     tgrid[2][kEZ], tgrid[3][kEZ], k0, k1, scale, phi5);
}

}  // namespace pdp3
