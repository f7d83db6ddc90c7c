// FP32 cell-relative form of the per-particle work of one step.
//
// State per particle: packed cell (bit 31 = dead, bits 16..30 = radial cell
// i, bits 0..15 = axial cell k), offsets a, b in (0, 1] with r = (i + a) dr,
// z = (k + b) dz (the reference's half-open cells (i dx, (i+1) dx],
// include/geometry.h:44), momentum per mass u = gamma v (FP32), weight w.
// Storing offsets instead of absolute coordinates keeps every bit of
// resolution however large the grid is; storing u keeps gamma well
// conditioned near c.
//
// Arithmetic: push and deposit in FP32; every formula of the reference is
// rewritten in cell units without catastrophic cancellation (ring-volume
// weights as products (1-a)(2i+a), the segment integrals of
// simple_current_distribution through log1p-type series, the r-z rotation
// through dy^2/(R+X)); the near-axis cell (1/r, the CYL_VOL annulus weights)
// is evaluated in FP64; the grid accumulation is FP64 (hardware FP64
// atomics); fields and the PML stay FP64.
#include <algorithm>
#include <vector>

#include "context.hpp"
#include "sink.hpp"

namespace pdp3 {

namespace {

constexpr uint32_t kDead = 0x80000000u;

// a / b, 1/sqrt(x) and sqrt(x) through the hardware instructions (about 1
// ulp each; no denormal arguments reach them here: the Lorentz factors are
// >= 1, the other arguments are O(1) in cell units or exactly 0)
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ float frsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ int cell_i(uint32_t c) { return (int)((c >> 16) & 0x7FFF); }
__device__ __forceinline__ int cell_k(uint32_t c) { return (int)(c & 0xFFFF); }
__device__ __forceinline__ uint32_t pack(int i, int k) { return ((uint32_t)i << 16) | (uint32_t)k; }

struct F32Args {
  int nr, nz;
  float hr, hz;      // dt / (2 dr), dt / (2 dz): half-step displacement per unit speed, in cells
  float inv_c2;      // 1 / c^2
  float c1;          // q dt / (2 m) of the species (w cancels, src/particles.cpp:1060)
  float bscale;      // Boris: mu0 c1; Vay: mu0 c1 * 2 c1^2 (the reference's Vay scaling, :1187-1210)
  float escale;      // Boris: c1; Vay: 2 c1
  float kz, kr, kp;  // charge/(4 pi dr^2 dt), charge/(2 pi dr dz dt), charge/(4 pi dz dr^2)
  int pusher;
  int is_bunch;
  int want_rho;
  int n;
  // the context's grid: component c (kER, kEPhi, kEZ, kHHR, kHHPhi, kHHZ) at
  // grid + c * nr * nz (one base pointer instead of six keeps the kernels'
  // scalar registers from spilling)
  const double *grid;
  // per radial cell index i, two entries: {1/(2i), 1/(2(i+1)), 1/(2(i+2)),
  // 1/(2(i+3))} and {1/i, 1/(i+1), 1/(i+1/2), i(i+1)} (1/0 read as 0)
  const float4 *rows;
  int *err;
  uint32_t *cell;
  float *fa, *fb, *ur, *uphi, *uz;
  const float *w;
};

// component c of the grid
__device__ __forceinline__ const double *field(const F32Args &p, int c) {
  return p.grid + (size_t)c * ((size_t)p.nr * (size_t)p.nz);
}

// radial weights of the four-node gathers (src/eField.cpp:204-316,
// src/hField.cpp:159-267) in cell units.  "lo" family (E_r, H_z, H_phi):
// staggered index from r - dr/2; otherwise (E_phi, E_z, H_r) the particle's
// own cell with the CYL_VOL volume on the axis cell.  Weights A (node i_r),
// B (node i_r+1), exactly the reference's ring volume / cell volume ratios
// (which do not sum to one).
struct RadW {
  int ir;
  float A, B;
};
// both radial weight pairs of a particle in cell i at offset a; t is the
// first table entry of row i.  The lo family's denominators 2(l+1), 2(l+3)
// (l = its staggered index, i or i-1) and the own family's 2i, 2(i+2) are all
// entries of that row.
__device__ __forceinline__ void rad_weights(int i, float a, const float4 &t, RadW &lo, RadW &own) {
  // staggered index from r - dr/2: i-1 with offset a+1/2 when a <= 1/2, except
  // on the axis cell, where the reference clamps the index at 0 and keeps the
  // offset a-1/2
  const bool down = a <= 0.5f && i > 0;
  const int ir = down ? i - 1 : i;
  const float al = down ? a + 0.5f : a - 0.5f;
  const float fl = (float)ir, fi = (float)i;
  lo.ir = ir;
  lo.A = (1.f - al) * (2.f * fl + 1.f + al) * (down ? t.x : t.y);
  lo.B = al * (2.f * fl + 2.f + al) * (down ? t.z : t.w);
  own.ir = i;
  if (i == 0) {
    double ad = a;  // axis cell in FP64: CYL_VOL annulus weight
    own.A = (float)((1.0 - ad) * ad * 4.0);
    own.B = (float)(ad * (1.0 + ad) / 4.0);
  } else {
    own.A = (1.f - a) * (2.f * fi + a) * t.x;
    own.B = a * (2.f * fi + 1.f + a) * t.z;
  }
}
struct AxW {
  int kz;
  float b;  // weight of node kz+1; node kz gets 1-b
};
__device__ __forceinline__ AxW ax_own(int k, float b) { return AxW{k, b}; }
// staggered axial index from z - dz/2 (clamped at 0 as the reference does)
__device__ __forceinline__ AxW ax_half(int k, float b) {
  const bool down = b <= 0.5f && k > 0;
  return AxW{down ? k - 1 : k, down ? b + 0.5f : b - 0.5f};
}

// node values are rounded to FP32 and multiplied by the push's scale s
// (escale for E, bscale for H) before the interpolation
__device__ __forceinline__ float gather(const F32Args &p, const double *__restrict__ F, float s, RadW rw, AxW aw) {
  int ir = rw.ir, kz = aw.kz;
  if (ir > p.nr - 2 || kz > p.nz - 2) {
    atomicOr(p.err, kErrDomain);
    ir = min(ir, p.nr - 2);
    kz = min(kz, p.nz - 2);
  }
  const double *row0 = F + (size_t)ir * (size_t)p.nz + kz;
  const double *row1 = row0 + p.nz;
  float f00 = (float)row0[0] * s, f01 = (float)row0[1] * s, f10 = (float)row1[0] * s, f11 = (float)row1[1] * s;
  float b = aw.b, ob = 1.f - aw.b;
  return (f00 * ob + f01 * b) * rw.A + (f10 * ob + f11 * b) * rw.B;
}

// field sources of the push.  GlobalFields reads the six FP64 components
// from HBM / L2 per particle; BoxFields serves the particles of a cell-sorted
// block from an LDS copy (converted to FP32 once per node) of the nodes their
// cells can touch, and falls back to HBM for the rest.  Both evaluate the same
// FP32 expression on the same FP32-rounded node values, so the result does not
// depend on the source.
// the two table entries of row i (the host builds p.rows with the same FP64
// expressions, so both give the same FP32 values)
__device__ __forceinline__ void row_entries(int i, float4 &t0, float4 &t1) {
  const double fi = i;
  t0 = make_float4(i ? (float)(1.0 / (2.0 * fi)) : 0.f, (float)(1.0 / (2.0 * (fi + 1))), (float)(1.0 / (2.0 * (fi + 2))),
                   (float)(1.0 / (2.0 * (fi + 3))));
  t1 = make_float4(i ? (float)(1.0 / fi) : 0.f, (float)(1.0 / (fi + 1)), (float)(1.0 / (fi + 0.5)), (float)(fi * (fi + 1)));
}

struct GlobalFields {
  // the two table entries of row i (clamped only for memory safety: a corrupt
  // cell word is caught by the gathers' domain check)
  __device__ __forceinline__ void row(const F32Args &p, int i, float4 &t0, float4 &t1) const {
    const int r = min(max(i, 0), p.nr - 1);
    t0 = p.rows[2 * r];
    t1 = p.rows[2 * r + 1];
  }
  __device__ __forceinline__ void get(const F32Args &p, int, int, RadW rl, RadW ro, AxW zo, AxW zh, float e[3],
                                      float h[3]) const {
    e[0] = gather(p, field(p, kER), p.escale, rl, zo);
    e[1] = gather(p, field(p, kEPhi), p.escale, ro, zo);
    e[2] = gather(p, field(p, kEZ), p.escale, ro, zh);
    h[0] = gather(p, field(p, kHHR), p.bscale, ro, zh);
    h[1] = gather(p, field(p, kHHPhi), p.bscale, rl, zh);
    h[2] = gather(p, field(p, kHHZ), p.bscale, rl, zo);
  }
};

// rows bi..bi+kFR-1, columns bk..bk+kFK-1 of the six components, anchored two
// rows and two columns below the tile's anchor cell (ai, ak): a particle in
// cell (i, k) reads rows i-1..i+1 and columns k-1..k+1, so every cell with
// i in [ai-1, ai+1] and k in [ak-1, ak+kFK-4] is served from the box (a
// sorted tile lies along row ai; between sorts some of its particles have
// moved to the rows next to it, or to the column before its first).  Each
// entry is a pair (node k, node k+1) of its row, so that the four nodes of
// one component's stencil are two 8-byte entries kFK apart (one LDS
// instruction) and the six components differ only by a constant offset.
constexpr int kFR = 5, kFK = 64, kFComp = 6, kFSize = kFComp * kFR * kFK;
constexpr int kFBelow = 2;  // box rows and columns below the anchor cell
// the LDS array (floats): the field box (two floats per entry), then two
// float4 row entries for each box row
constexpr int kFBoxFloats = 2 * kFSize + 8 * kFR;

__device__ __forceinline__ float gather_box(const float2 *e, RadW rw, AxW aw) {
  const float2 lo = e[0], hi = e[kFK];
  float b = aw.b, ob = 1.f - aw.b;
  return (lo.x * ob + lo.y * b) * rw.A + (hi.x * ob + hi.y * b) * rw.B;
}

struct BoxFields {
  const float *box;  // LDS, kFComp x kFR x kFK pairs, then the row entries of rows ai-2 .. ai+2
  int ai, ak;        // anchor cell of the tile
  // the box's row entries (an LDS read for every particle), replaced by the
  // entries evaluated in place for a row outside the box.  A global load here
  // would make the read a flat load, whose wait also waits for the next
  // particle's state loads already in flight.
  __device__ __forceinline__ void row(const F32Args &p, int i, float4 &t0, float4 &t1) const {
    const int li = i - ai + kFBelow;
    const int lc = min(max(li, 0), kFR - 1);
    const float4 *t = reinterpret_cast<const float4 *>(box + 2 * kFSize);
    t0 = t[2 * lc];
    t1 = t[2 * lc + 1];
    if ((unsigned)li >= (unsigned)kFR) row_entries(min(max(i, 0), p.nr - 1), t0, t1);
  }
  __device__ __forceinline__ void get(const F32Args &p, int i, int k, RadW rl, RadW ro, AxW zo, AxW zh, float e[3],
                                      float h[3]) const {
    const bool inbox = (unsigned)(i - ai + kFBelow - 1) < (unsigned)(kFR - 2) &&
                       (unsigned)(k - ak + kFBelow - 1) < (unsigned)(kFK - 2) && i <= p.nr - 2 && k <= p.nz - 2;
    if (inbox) {
      const int bi = ai - kFBelow, bk = ak - kFBelow;
      constexpr int C = kFR * kFK;
      const float2 *f = reinterpret_cast<const float2 *>(box);
      // the stencil's first entry for each (radial, axial) weight pair
      const int rlo = (rl.ir - bi) * kFK, roo = (ro.ir - bi) * kFK, zoo = zo.kz - bk, zho = zh.kz - bk;
      const float2 *lo_o = f + (rlo + zoo), *oo_o = f + (roo + zoo), *oo_h = f + (roo + zho), *lo_h = f + (rlo + zho);
      const float e0 = gather_box(lo_o + 0 * C, rl, zo);
      const float e1 = gather_box(oo_o + 1 * C, ro, zo);
      const float e2 = gather_box(oo_h + 2 * C, ro, zh);
      const float h0 = gather_box(oo_h + 3 * C, ro, zh);
      const float h1 = gather_box(lo_h + 4 * C, rl, zh);
      const float h2 = gather_box(lo_o + 5 * C, rl, zo);
      e[0] = e0; e[1] = e1; e[2] = e2;
      h[0] = h0; h[1] = h1; h[2] = h2;
    } else {
      const float e0 = gather(p, field(p, kER), p.escale, rl, zo);
      const float e1 = gather(p, field(p, kEPhi), p.escale, ro, zo);
      const float e2 = gather(p, field(p, kEZ), p.escale, ro, zh);
      const float h0 = gather(p, field(p, kHHR), p.bscale, ro, zh);
      const float h1 = gather(p, field(p, kHHPhi), p.bscale, rl, zh);
      const float h2 = gather(p, field(p, kHHZ), p.bscale, rl, zo);
      e[0] = e0; e[1] = e1; e[2] = e2;
      h[0] = h0; h[1] = h1; h[2] = h2;
    }
  }
};

// block prologue: fill the field box (nodes outside the grid read as 0; no
// particle that the box serves reads them, nor the second half of a row's
// last entry)
__device__ __forceinline__ void field_box_fill(const F32Args &p, float *box, int ai, int ak) {
  const int bi = ai - kFBelow, bk = ak - kFBelow;
  if (threadIdx.x < 2 * kFR) {
    const int r = min(max(bi + (int)(threadIdx.x >> 1), 0), p.nr - 1);
    reinterpret_cast<float4 *>(box + 2 * kFSize)[threadIdx.x] = p.rows[2 * r + (threadIdx.x & 1)];
  }
  const double *src[kFComp] = {field(p, kER), field(p, kEPhi), field(p, kEZ), field(p, kHHR), field(p, kHHPhi),
                               field(p, kHHZ)};
  float2 *f = reinterpret_cast<float2 *>(box);
#pragma unroll
  for (int comp = 0; comp < kFComp; comp++) {
    const double *F = src[comp];
    const float sc = comp < 3 ? p.escale : p.bscale;
    // one wavefront per box row (kFK = 64 columns): each node is loaded once
    // and the next column's value comes from the next lane
    for (int t = threadIdx.x; t < kFR * kFK; t += blockDim.x) {
      const int r = t / kFK, col = t - r * kFK;
      const int gi = bi + r, gk = bk + col;
      float v0 = 0.f;
      if ((unsigned)gi < (unsigned)p.nr && (unsigned)gk < (unsigned)p.nz) v0 = (float)F[(size_t)gi * p.nz + gk] * sc;
      const float v1 = __shfl_down(v0, 1);  // lane 63 (the row's last entry) gets its own value: never read
      f[comp * kFR * kFK + t] = make_float2(v0, v1);
    }
  }
}

__device__ __forceinline__ float dot3f(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// relativistic Boris on u = gamma v, src/particles.cpp:1025-1150; e and h
// arrive multiplied by escale and bscale (the field sources apply them per node)
__device__ __forceinline__ void boris_u(const F32Args &p, const float e[3], const float h[3], float u[3]) {
  const float ee[3] = {e[0], e[1], e[2]};
  float b[3] = {h[0], h[1], h[2]};
  for (int j = 0; j < 3; j++) u[j] += ee[j];
  float g = frsq(1.f + dot3f(u, u) * p.inv_c2);
  for (int j = 0; j < 3; j++) b[j] *= g;
  float c2 = fdiv(2.f, 1.f + dot3f(b, b));
  float t0 = u[0], t1 = u[1], t2 = u[2];
  u[0] = t0 + c2 * ((t1 - t0 * b[2] + t2 * b[0]) * b[2] - (t2 + t0 * b[1] - t1 * b[0]) * b[1]);
  u[1] = t1 + c2 * (-(t0 + t1 * b[2] - t2 * b[1]) * b[2] + (t2 + t0 * b[1] - t1 * b[0]) * b[0]);
  u[2] = t2 + c2 * ((t0 + t1 * b[2] - t2 * b[1]) * b[1] - (t1 - t0 * b[2] + t2 * b[0]) * b[0]);
  for (int j = 0; j < 3; j++) u[j] += ee[j];
}

// Vay on u, src/particles.cpp:1152-1254 (with the reference's b scaling; e
// and h arrive multiplied by escale and bscale)
__device__ __forceinline__ void vay_u(const F32Args &p, const float e[3], const float h[3], float u[3]) {
  const float ee[3] = {e[0], e[1], e[2]};
  float b[3] = {h[0], h[1], h[2]};
  float t[3] = {u[0], u[1], u[2]};
  float g = frsq(1.f + dot3f(t, t) * p.inv_c2);
  float up[3];
  up[0] = t[0] + ee[0] + g * (t[1] * b[2] - t[2] * b[1]);
  up[1] = t[1] + ee[1] + g * (t[2] * b[0] - t[0] * b[2]);
  up[2] = t[2] + ee[2] + g * (t[0] * b[1] - t[1] * b[0]);
  float alpha = 1.f + dot3f(up, up) * p.inv_c2;  // gamma'^2
  float B2 = dot3f(b, b);
  float s = alpha - B2;
  float ub = dot3f(up, b);
  float us2 = ub * ub;
  float ig = frsq(0.5f * (s + fsqrt(s * s + 4.f * (B2 + us2))));  // 1/gamma^{n+1}
  for (int j = 0; j < 3; j++) b[j] *= ig;
  s = fdiv(1.f, 1.f + dot3f(b, b));
  alpha = dot3f(up, b);
  u[0] = s * (up[0] + alpha * b[0] + b[2] * up[1] - b[1] * up[2]);
  u[1] = s * (up[1] + alpha * b[1] + b[0] * up[2] - b[2] * up[0]);
  u[2] = s * (up[2] + alpha * b[2] + b[1] * up[0] - b[0] * up[1]);
}

// bring an offset back into (0, 1]: by ceil(o) - 1 cells (one cell at most
// after a half step; exact in FP32)
__device__ __forceinline__ void renorm(int &c, float &o) {
  const float m = ceilf(o) - 1.f;
  o -= m;
  c += (int)m;
}

// r of the Cartesian point (X, dy), X = i + xa, as an offset update that
// never forms X^2 + dy^2 - X^2: R - X = dy^2 / (R + X) for X > 0 (and
// R + |X| when the point lies behind the axis).  Returns sin(theta) = dy / R;
// ct = |X| / R is the reference's sqrt(1 - sin^2) (src/particles.cpp:1010).
__device__ __forceinline__ float rotate_back(int i, float &xa, float dy, float &ct) {
  const float X = (float)i + xa;
  const float s = fmaxf(X * X + dy * dy, 1e-30f);
  const float q = frsq(s);  // 1 / R
  const float R = s * q;
  xa += X > 0.f ? fdiv(dy * dy, R + X) : R - X;
  ct = fabsf(X) * q;
  return dy * q;
}

// J_phi of one particle at the half-step position,
// azimuthal_current_distribution (src/particles.cpp:688-784) in cell units
template <class S>
__device__ void deposit_jphi(const F32Args &p, S &sk, float w, int i, float a, int k, float b, float vphi) {
  double c0, c1;  // node (i,k) and (i+1,k) weights before the axial split
  if (i == 0) {
    double ad = a;  // axis cell in FP64; rho = a here
    c0 = 8.0 * (1.0 - ad);
    c1 = 1.0 + ad;
  } else {
    // 1 / (rho i) and 1 / (rho (i+1)) from one reciprocal
    const float fi = (float)i, rho = fi + a, fi1 = fi + 1.f;
    const float inv = __builtin_amdgcn_rcpf(rho * fi * fi1);
    c0 = (1.f - a) * (2.f * fi + a) * (fi1 * inv);
    c1 = a * (2.f * fi + 1.f + a) * (fi * inv);
  }
  float s = p.kp * w * vphi;
  const double s0 = (double)(s * (1.f - b)), s1 = (double)(s * b);
  sk.template add4<kCPhi>(i, k, s0 * c0, s0 * c1, s1 * c0, s1 * c1);
}

// Particles::charge_weighting of one particle (src/particles.cpp:101-226)
template <class S>
__device__ void deposit_rho(const F32Args &p, S &sk, float w, int i, float a, int k, float b) {
  double c0, c1;
  if (i == 0) {
    double rho = a;
    if (rho <= 0.5) {  // r <= dr/2 branch (:155-186)
      c0 = 16.0 * (0.5 - rho + rho * rho) / (2.0 * rho * rho + 0.5);
      c1 = 2.0 * rho * (rho + 1.0) / (2.0 * rho * rho + 0.5);
    } else {  // dr/2 < r <= dr (:188-218)
      c0 = 8.0 * (1.0 - rho);
      c1 = 1.0 + rho;
    }
  } else {
    const float fi = (float)i, rho = fi + a, fi1 = fi + 1.f;
    const float inv = __builtin_amdgcn_rcpf(rho * fi * fi1);
    c0 = (1.f - a) * (2.f * fi + a) * (fi1 * inv);
    c1 = a * (2.f * fi + 1.f + a) * (fi * inv);
  }
  float s = p.kp * w;
  const double s0 = (double)(s * (1.f - b)), s1 = (double)(s * b);
  sk.template add4<kCRho>(i, k, s0 * c0, s0 * c1, s1 * c0, s1 * c1);
}

// (log(1+x) - x) / x without cancellation
__device__ __forceinline__ float m_of(float x) {
  if (fabsf(x) < 0.1f)
    return x * (-0.5f + x * (1.f / 3.f + x * (-0.25f + x * (0.2f + x * (-1.f / 6.f + x * (1.f / 7.f - x * 0.125f))))));
  return fdiv(log1pf(x) - x, x);
}

// one straight piece of the path inside cell (i, k): simple_current_distribution
// (src/particles.cpp:333-474) rewritten in cell units.  (a, b): start offsets
// in the cell, (dr_, dz_): displacement in cells.  The closed forms
//   J_z[i]   ~ dz/i     [ (1-a)(2i+a)/rho - dr/2 + i(i+1)/rho M(x) ]
//   J_z[i+1] ~ dz/(i+1) [ a(2i+1+a)/rho  + dr/2 - i(i+1)/rho M(x) ]
//   J_r[k]   ~ [dr((1-b)/2 - dz/4) + i(i+1)/(2rho)(dr(1-b)/rho_e + dz N(x))] / (i+1/2)
//   J_r[k+1] ~ [dr(b/2 + dz/4)     + i(i+1)/(2rho)(dr b/rho_e     - dz N(x))] / (i+1/2)
// with rho = i + a, x = dr/rho, M = (ln(1+x)-x)/x, N = 1/(1+x) - ln(1+x)/x are
// algebraically identical to the reference's expressions (which evaluate them
// as differences of O(i) terms); they also reduce to strict_motion_weighting
// (:786-988) for purely axial or radial motion.
template <class S>
__device__ void segment(const F32Args &p, S &sk, float w, int i, int k, float a, float b, float drho, float dzeta,
                        const float4 &t) {
  float jz0, jz1, jr0, jr1;
  float f0 = 1.f, f1 = 1.f, fr = 1.f;  // factors of the J_z pair and of J_r
  if (i == 0) {
    double rs = a;  // axis cell in FP64
    jz0 = (float)(8.0 * (1.0 - 0.5 * drho - rs));
    jz1 = (float)(0.5 * drho + 1.0 + rs);
    jr0 = drho * (0.5f * (1.f - b) - 0.25f * dzeta) * 2.f;
    jr1 = drho * (0.5f * b + 0.25f * dzeta) * 2.f;
  } else {
    // t: second table entry of row i (1/i, 1/(i+1), 1/(i+1/2), i(i+1)); the
    // closed forms above, with 1/rho factored out of the J_z brackets
    const float fi = (float)i;
    const float rs = fi + a, re = rs + drho;
    // 1/rs, 1/re from one reciprocal
    const float q = __builtin_amdgcn_rcpf(rs * re);
    const float irs = re * q, ire = rs * q;
    const float ii = t.w;
    const float M = m_of(drho * irs);
    const float iiM = ii * M;
    const float N = -drho * ire - M;  // x / (1 + x) = drho / re
    const float hdr = 0.5f * drho;
    jz0 = fmaf(irs, fmaf(1.f - a, 2.f * fi + a, iiM), -hdr);
    jz1 = fmaf(irs, fmaf(a, 2.f * fi + 1.f + a, -iiM), hdr);
    const float cc = 0.5f * ii * irs, qdz = 0.25f * dzeta, dri = drho * ire, dN = dzeta * N, ub = 1.f - b;
    jr0 = fmaf(drho, fmaf(0.5f, ub, -qdz), cc * fmaf(dri, ub, dN));
    jr1 = fmaf(drho, fmaf(0.5f, b, qdz), cc * fmaf(dri, b, -dN));
    // the divisions by i, i+1 and i+1/2 go into the scale factors
    f0 = t.x;
    f1 = t.y;
    fr = t.z;
  }
  const float sz = p.kz * w * dzeta, sr = p.kr * w * fr;
  sk.template add2<kCZ>(i, k, (double)(jz0 * (sz * f0)), (double)(jz1 * (sz * f1)));
  sk.template add2<kCR>(i, k, (double)(sr * jr0), (double)(sr * jr1));
}

// the path parameter in [0, 1] at which a move of `d` cells reaches a face
// `f` cells away.  A move across a face can round to d = 0 in cell units: a
// particle 4e-9 cells above a face that ends 1e-9 cells below it has the new
// offset 1 - 1e-9, which is 1.0f, in the cell below, and the displacement
// (-1) + (1.0f - 4e-9) is 0.  The face is then taken at the start of the path
// (the deposited current is that of the part of the move that FP32 keeps).
__device__ __forceinline__ float cross_param(float f, float d) {
  return d != 0.f ? fminf(fmaxf(fdiv(f, d), 0.f), 1.f) : 0.f;
}

// current_distribution (src/particles.cpp:476-686): the straight path from
// the old to the new position, split at the cell faces it crosses (at most
// one radial and one axial face: |displacement| <= c dt <= dr/2), each piece
// deposited in its own cell
template <class S, class FS>
__device__ void deposit_j(const F32Args &p, S &sk, const FS &fs, float w, int io, float ao, int ko, float bo, int in_,
                          float an, int kn, float bn, float4 rt) {
  int di = in_ - io, dk = kn - ko;
  if (di < -1 || di > 1 || dk < -1 || dk > 1) {
    atomicOr(p.err, kErrDomain);
    return;
  }
  float drho = (float)di + (an - ao);
  float dzeta = (float)dk + (bn - bo);
  // pieces before the faces the path crosses, at the path parameters tr, tz
  // where it meets them (in old-cell units the radial face is at offset 1
  // (di=+1) or 0 (di=-1))
  int ci = io, ck = ko;
  float ca = ao, cb = bo;  // current start point, offsets in cell (ci, ck)
  if (di != 0 || dk != 0) {
    float tr = 2.f, tz = 2.f;
    if (di != 0) tr = cross_param((di > 0 ? 1.f : 0.f) - ao, drho);
    if (dk != 0) tz = cross_param((dk > 0 ? 1.f : 0.f) - bo, dzeta);
    float tprev = 0.f;
    for (int pass = 0; pass < 2; pass++) {
      bool rfirst = tr <= tz;
      float t = rfirst ? tr : tz;
      if (t > 1.f) break;
      float sd_r = (t - tprev) * drho, sd_z = (t - tprev) * dzeta;
      if (rfirst) {
        float fr = di > 0 ? 1.f : 0.f;
        segment(p, sk, w, ci, ck, ca, cb, fr - ca, sd_z, rt);
        ci += di;
        float4 rt0;
        fs.row(p, ci, rt0, rt);  // the field source's copy of the table (LDS in a tile)
        ca = di > 0 ? 0.f : 1.f;
        cb += sd_z;
        tr = 2.f;
      } else {
        float fz = dk > 0 ? 1.f : 0.f;
        segment(p, sk, w, ci, ck, ca, cb, sd_r, fz - cb, rt);
        ck += dk;
        cb = dk > 0 ? 0.f : 1.f;
        ca += sd_r;
        tz = 2.f;
      }
      tprev = t;
    }
  }
  // last piece to the new position
  segment(p, sk, w, ci, ck, ca, cb, (an + (float)(in_ - ci)) - ca, (bn + (float)(kn - ck)) - cb, rt);
}

// wall handling of move_half_reflect on the Cartesian half-step point
// (src/particles.cpp:1318-1353, src/bunch.cpp:126-153), in cell units: the
// walls sit at offset 1/2 of the first and last cells
__device__ __forceinline__ void walls(const F32Args &p, int &i, float &xa, int &k, float &b, float u[3],
                                      bool &alive) {
  // a point at least two cells from every wall cannot reach one (|offset
  // change| <= 1/4 per half step): nothing to test
  if (i >= 2 && i <= p.nr - 4 && k >= 2 && k <= p.nz - 4) return;
  const int il = p.nr - 2, kl = p.nz - 2;
  float off = (float)(i - il) + xa;
  if (i >= il - 1 && off > 0.5f) {
    if (p.is_bunch) {
      alive = false;
    } else {
      i = il;
      xa = 1.f - off;
      u[0] = -u[0];
    }
  }
  float offz = (float)(k - kl) + b;
  if (k >= kl - 1 && offz > 0.5f) {
    if (p.is_bunch) {
      alive = false;
    } else {
      k = kl;
      b = 1.f - offz;
      u[2] = -u[2];
    }
  }
  if (i <= 1) {
    float o0 = (float)i + xa;
    if (o0 < 0.5f) {
      i = 0;
      xa = 1.f - o0;
      u[0] = -u[0];
    }
  }
  if (k <= 1) {
    float o0 = (float)k + b;
    if (o0 < 0.5f) {
      k = 0;
      b = 1.f - o0;
      u[2] = -u[2];
    }
  }
}

// the state of one particle as loaded from HBM
struct PState {
  uint32_t c;
  float a, b, u0, u1, u2, w;
};

// the 4-byte element at byte offset `off` from p: a 32-bit offset from a
// uniform base is one VGPR and scalar-base addressing, where an element index
// would be widened to a 64-bit address per array
template <class T>
__device__ __forceinline__ T &at(T *p, unsigned off) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + off);
}
template <class T>
__device__ __forceinline__ const T &at(const T *p, unsigned off) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + off);
}

__device__ __forceinline__ PState load_state(const F32Args &p, unsigned off) {
  PState s;
  // streaming (non-temporal) loads and stores: each slot is touched once per
  // step, so the state should not displace the fields and tables from the caches
  s.c = __builtin_nontemporal_load(&at(p.cell, off));
  s.a = __builtin_nontemporal_load(&at(p.fa, off));
  s.b = __builtin_nontemporal_load(&at(p.fb, off));
  s.u0 = __builtin_nontemporal_load(&at(p.ur, off));
  s.u1 = __builtin_nontemporal_load(&at(p.uphi, off));
  s.u2 = __builtin_nontemporal_load(&at(p.uz, off));
  s.w = __builtin_nontemporal_load(&at(p.w, off));
  // all seven loads are issued together, ahead of the dead-particle branch
  // (instead of the six that the branch guards waiting for the cell word)
  asm volatile("" ::"v"(s.c), "v"(s.a), "v"(s.b), "v"(s.u0), "v"(s.u1), "v"(s.u2), "v"(s.w));
  return s;
}

// the state for the lane's next iteration: plain loads (they are waited for
// where the compiler needs them, or by wait_loads)
__device__ __forceinline__ PState load_state_next(const F32Args &p, unsigned off) {
  PState s;
  s.c = __builtin_nontemporal_load(&at(p.cell, off));
  s.a = __builtin_nontemporal_load(&at(p.fa, off));
  s.b = __builtin_nontemporal_load(&at(p.fb, off));
  s.u0 = __builtin_nontemporal_load(&at(p.ur, off));
  s.u1 = __builtin_nontemporal_load(&at(p.uphi, off));
  s.u2 = __builtin_nontemporal_load(&at(p.uz, off));
  s.w = __builtin_nontemporal_load(&at(p.w, off));
  return s;
}
// s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima): every vector
// memory access of the wave has completed
__device__ __forceinline__ void wait_loads() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// the new state of one particle
struct POut {
  uint32_t c;
  float a, b, u0, u1, u2;
};

// the whole step of one particle (state ps) without its stores: false for a
// dead slot (nothing to store)
template <class S, class FS>
__device__ __forceinline__ bool step_state32(const F32Args &p, S &sk, const FS &fs, const PState &ps, POut &o) {
  const uint32_t c = ps.c;
  if (c & kDead) return false;
  int i = cell_i(c), k = cell_k(c);
  float a = ps.a, b = ps.b;
  float u[3] = {ps.u0, ps.u1, ps.u2};
  const float w = ps.w;
  sk.begin_cell(i, k);
  // gathers (E at the full step, H at the half-time copy)
  RadW rl, ro;
  float4 t0, t1;
  fs.row(p, i, t0, t1);
  rad_weights(i, a, t0, rl, ro);
  AxW zo = ax_own(k, b), zh = ax_half(k, b);
  float e[3], h[3];
  fs.get(p, i, k, rl, ro, zo, zh, e, h);
  if (p.pusher == 1)
    vay_u(p, e, h, u);
  else
    boris_u(p, e, h, u);
  float g = frsq(1.f + dot3f(u, u) * p.inv_c2);
  float dx = u[0] * g * p.hr, dy = u[1] * g * p.hr, dz = u[2] * g * p.hz;
  // first half step and rotation back to r-z (step_v, :1296-1316)
  int i1 = i, k1 = k;
  float xa = a + dx, b1 = b + dz;
  float ct;
  rotate_back(i1, xa, dy, ct);
  renorm(i1, xa);
  renorm(k1, b1);
  deposit_jphi(p, sk, w, i1, xa, k1, b1, u[1] * g);
  // second half step with the walls (move_half_reflect, :1355-1365)
  float xa2 = xa + dx, b2 = b1 + dz;
  bool alive = true;
  walls(p, i1, xa2, k1, b2, u, alive);
  const float st = rotate_back(i1, xa2, dy, ct);
  renorm(i1, xa2);
  renorm(k1, b2);
  if (alive) {
    deposit_j(p, sk, fs, w, i, a, k, b, i1, xa2, k1, b2, t1);
    // back_velocity_to_rz with the second half step's angle (:1005-1023)
    float u0 = ct * u[0] - st * u[1];
    float u1 = st * u[0] + ct * u[1];
    u[0] = u0;
    u[1] = u1;
    if (p.want_rho) deposit_rho(p, sk, w, i1, xa2, k1, b2);
  }
  o.c = pack(i1, k1) | (alive ? 0u : kDead);
  o.a = xa2;
  o.b = b2;
  o.u0 = u[0];
  o.u1 = u[1];
  o.u2 = u[2];
  return true;
}

__device__ __forceinline__ void store_state(const F32Args &p, const POut &o, unsigned off) {
  // a fresh copy of the offset: the stores then address from the scalar bases
  // too, instead of reusing 64-bit addresses formed for the loads
  unsigned so = off;
  asm("" : "+v"(so));
  __builtin_nontemporal_store(o.c, &at(p.cell, so));
  __builtin_nontemporal_store(o.a, &at(p.fa, so));
  __builtin_nontemporal_store(o.b, &at(p.fb, so));
  __builtin_nontemporal_store(o.u0, &at(p.ur, so));
  __builtin_nontemporal_store(o.u1, &at(p.uphi, so));
  __builtin_nontemporal_store(o.u2, &at(p.uz, so));
}

// the whole step of the particle in slot `slot` of the arrays of p (which
// start at most 2^30 slots before it)
template <class S, class FS>
__device__ __forceinline__ void step_particle32(const F32Args &p, S &sk, const FS &fs, unsigned slot) {
  POut o;
  if (step_state32(p, sk, fs, load_state(p, slot * 4u), o)) store_state(p, o, slot * 4u);
}

// the arguments with the particle arrays starting at slot `base` (uniform over
// the block): slots are then addressed by a 32-bit offset from a scalar base
__device__ __forceinline__ F32Args from_slot(const F32Args &p, long long base) {
  F32Args q = p;
  q.cell += base;
  q.fa += base;
  q.fb += base;
  q.ur += base;
  q.uphi += base;
  q.uz += base;
  q.w += base;
  return q;
}

// one tile of a cell-sorted (chunked) species: CB = WAVES * PPT / kPPT chunks
// starting at slot `base`.  Wave w takes chunk w / (kPPT / PPT) and, in it,
// the slots of iterations j0 .. j0+PPT-1, j0 = (w % (kPPT / PPT)) * PPT: its
// lanes read contiguous slots (coalesced) and lane L walks the consecutive
// logical particles L*kPPT+j0 .. L*kPPT+j0+PPT-1, mostly of one cell.
template <int PPT, int WAVES>
__device__ __forceinline__ void fused32_tile(const F32Args &p, const GlobalSink &gs, long long base, long long n_cap,
                                             double *box, float *fbox) {
  static_assert(kFK == 64, "one wavefront per field box row");
  static_assert(kPPT % PPT == 0 && (WAVES * PPT) % kPPT == 0, "tile shape");
  // anchor samples: slot base + j is the logical particle 16 j of a plasma
  // tile (the chunk-transposed layout), particle j of an ensemble's bunch block
  int ai, ak;
  {
    bool live[3];
    int ri[3], kk[3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const uint32_t c = base + j < n_cap ? p.cell[base + j] : kDead;
      live[j] = !(c & kDead);
      ri[j] = cell_i(c);
      kk[j] = cell_k(c);
    }
    tile_anchor(live, ri, kk, ai, ak);
  }
  TileSink sk;
  box_init(box, ai, ak, sk, gs);
  field_box_fill(p, fbox, ai, ak);
  __syncthreads();
  const BoxFields fs{fbox, ai, ak};
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int kSplit = kPPT / PPT;
  const int chunk = wave / kSplit, j0 = (wave % kSplit) * PPT;
  constexpr long long kTile = (long long)WAVES * PPT / kPPT * kChunk;
  const F32Args q = from_slot(p, base);
  const unsigned lim = (unsigned)min(n_cap - base, kTile);
  const unsigned off0 = (unsigned)(chunk * kChunk + j0 * 64 + lane);
  // software pipeline: the state of the lane's next particle is loaded before
  // the current one is pushed; all its loads are waited for before the
  // current particle's stores, so the wait never includes stores in flight
  PState nx{kDead, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  if (off0 < lim) nx = load_state_next(q, off0 * 4u);
  for (int j = 0; j < PPT; j++) {
    const PState cur = nx;
    unsigned n = off0 + (unsigned)j * 64u, n1 = n + 64u;
    asm("" : "+v"(n), "+v"(n1));  // keeps the slots 32-bit offsets from the scalar bases (no 64-bit pointer per array)
    if (j + 1 < PPT && n1 < lim) nx = load_state_next(q, n1 * 4u);
    POut o;
    const bool st = n < lim && step_state32(q, sk, fs, cur, o);
    wait_loads();
    if (st) store_state(q, o, n * 4u);
  }
  sk.flush();
  __syncthreads();
  box_flush(box, sk);
}

template <int PPT, int WAVES>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) k_fused32_tiled(F32Args p, GlobalSink gs, long long n_cap) {
  __shared__ double box[TileSink::kSize];
  __shared__ __align__(16) float fbox[kFBoxFloats];
  constexpr long long kTile = (long long)WAVES * PPT / kPPT * kChunk;
  // launched for the chunked layout only, i.e. plasma species: no bunch walls
  // and no charge density (constants here drop their code and registers)
  F32Args q = p;
  q.is_bunch = 0;
  q.want_rho = 0;
  // the sink's grid size and error word are the arguments' (one copy each in
  // scalar registers)
  GlobalSink g = gs;
  g.nr = q.nr;
  g.nz = q.nz;
  g.err = q.err;
  fused32_tile<PPT, WAVES>(q, g, (long long)blockIdx.x * kTile, n_cap, box, fbox);
}

// bunches (plain layout): the same walk, the box anchored at the smallest
// living cell of the block
// one block of a bunch (plain layout): WAVES chunks of kChunk slots from
// `base`; the register / LDS sink is anchored at the smallest cell of the
// block's living particles (a bunch block is a thin disc injected in one step)
template <int WAVES, int PPT>
__device__ __forceinline__ void fused32_bunch_tile(const F32Args &p, const GlobalSink &gs, long long base,
                                                   long long n_cap, double *box) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long long cbase = base + (long long)wave * (64 * PPT) + lane;
  int ai = 0x7fffffff, ak = 0x7fffffff;
  for (int j = 0; j < PPT; j++) {
    long long n = cbase + (long long)j * 64;
    if (n < n_cap) {
      uint32_t c = p.cell[n];
      if (!(c & kDead)) {
        ai = min(ai, cell_i(c));
        ak = min(ak, cell_k(c));
      }
    }
  }
  block_min2(ai, ak);
  if (ai == 0x7fffffff) return;  // no living particle in the block (uniform over the block)
  BunchSink sk;
  box_init(box, ai, ak, sk, gs);
  __syncthreads();
  const F32Args q = from_slot(p, base);
  const unsigned lim = (unsigned)min(n_cap - base, (long long)WAVES * 64 * PPT);
  const unsigned off0 = (unsigned)(wave * 64 * PPT + lane);
  for (int j = 0; j < PPT; j++) {
    const unsigned n = off0 + (unsigned)j * 64u;
    if (n < lim) step_particle32(q, sk, GlobalFields(), n);
  }
  sk.flush();
  __syncthreads();
  box_flush(box, sk);
}

// bunches: PPT particles per lane (wave w of a block takes the 64 * PPT slots
// from w * 64 * PPT); a small bunch takes few per lane so that its blocks
// still spread over the GPU
template <int PPT>
__global__ void __launch_bounds__(256) k_fused32_bunch(F32Args p, GlobalSink gs, long long n_cap) {
  __shared__ double box[BunchSink::kSize];
  fused32_bunch_tile<kWavesPerBlock, PPT>(p, gs, (long long)blockIdx.x * (kWavesPerBlock * 64 * PPT), n_cap, box);
}

// ensembles: the plasma tiles and bunch blocks of every instance in one
// launch, one block per item of the work list (a persistent grid striding
// over the list kept the loop's state in scalar registers, which spilled).
// Bunch blocks take the plasma tile path too (their particles are not cell
// sorted, so most of their gathers and deposits take the boxes' HBM
// fallbacks), which keeps the kernel at the plasma path's register count.
struct EnsItem {
  int arg;   // index into the per-(instance, species) argument tables
  int kind;  // 0 plasma tile, 1 bunch block
  long long base, n_cap;
};

// A pointer read from device memory is generic (flat) to the compiler: its
// loads, stores and atomics become flat instructions on 64-bit per-lane
// addresses, which also hold up the kernel's LDS waits.  The ensemble's
// argument tables are therefore read through mirrors of F32Args and
// GlobalSink whose pointers are in the global address space (same layout);
// converted to the generic structs, every access derived from them compiles
// to global instructions on scalar bases, as for kernel arguments.
template <class T>
using gptr = __attribute__((address_space(1))) T *;
struct F32ArgsG {
  int nr, nz;
  float hr, hz, inv_c2, c1, bscale, escale, kz, kr, kp;
  int pusher, is_bunch, want_rho, n;
  gptr<const double> grid;
  gptr<const float4> rows;
  gptr<int> err;
  gptr<uint32_t> cell;
  gptr<float> fa, fb, ur, uphi, uz;
  gptr<const float> w;
};
static_assert(sizeof(F32ArgsG) == sizeof(F32Args) && offsetof(F32ArgsG, grid) == offsetof(F32Args, grid) &&
                  offsetof(F32ArgsG, w) == offsetof(F32Args, w) && offsetof(F32ArgsG, n) == offsetof(F32Args, n),
              "F32ArgsG mirrors F32Args");
struct GlobalSinkG {
  int nr, nz;
  gptr<double> j, rho;
  gptr<int> err;
};
static_assert(sizeof(GlobalSinkG) == sizeof(GlobalSink) && offsetof(GlobalSinkG, err) == offsetof(GlobalSink, err),
              "GlobalSinkG mirrors GlobalSink");

__device__ __forceinline__ F32Args from_global(const F32ArgsG &g) {
  F32Args a;
  a.nr = g.nr;
  a.nz = g.nz;
  a.hr = g.hr;
  a.hz = g.hz;
  a.inv_c2 = g.inv_c2;
  a.c1 = g.c1;
  a.bscale = g.bscale;
  a.escale = g.escale;
  a.kz = g.kz;
  a.kr = g.kr;
  a.kp = g.kp;
  a.pusher = g.pusher;
  a.is_bunch = g.is_bunch;
  a.want_rho = g.want_rho;
  a.n = g.n;
  a.grid = (const double *)g.grid;
  a.rows = (const float4 *)g.rows;
  a.err = (int *)g.err;
  a.cell = (uint32_t *)g.cell;
  a.fa = (float *)g.fa;
  a.fb = (float *)g.fb;
  a.ur = (float *)g.ur;
  a.uphi = (float *)g.uphi;
  a.uz = (float *)g.uz;
  a.w = (const float *)g.w;
  return a;
}
__device__ __forceinline__ GlobalSink from_global(const GlobalSinkG &g) {
  GlobalSink s;
  s.nr = g.nr;
  s.nz = g.nz;
  s.j = (double *)g.j;
  s.rho = (double *)g.rho;
  s.err = (int *)g.err;
  return s;
}

template <int PPT, int WAVES>
// 5 waves per SIMD (96 VGPRs, 6 scalar registers spilled into vector lanes)
// once the table pointers are global; with flat pointers it needed 121 VGPRs
// and ran at 4
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5)))
k_ens_fused32(const F32Args *__restrict__ args, const GlobalSink *__restrict__ sinks,
              const EnsItem *__restrict__ items, int n_items) {
  __shared__ double box[TileSink::kSize];
  __shared__ __align__(16) float fbox[kFBoxFloats];
  const int it = blockIdx.x;
  if (it >= n_items) return;
  const EnsItem item = items[it];
  const F32ArgsG *ag = reinterpret_cast<const F32ArgsG *>(args) + item.arg;
  const GlobalSinkG *sg = reinterpret_cast<const GlobalSinkG *>(sinks) + item.arg;
  const F32Args a = from_global(*ag);
  GlobalSink g = from_global(*sg);
  g.nr = a.nr;  // one copy of the grid size and error word (scalar registers)
  g.nz = a.nz;
  g.err = a.err;
  fused32_tile<PPT, WAVES>(a, g, item.base, item.n_cap, box, fbox);
}

// plain layout (bunches): one particle per thread, deposits straight to HBM
__global__ void __launch_bounds__(256) k_fused32_plain(F32Args p, GlobalSink gs, long long n_cap) {
  const long long base = (long long)blockIdx.x * blockDim.x;
  if (base + threadIdx.x >= n_cap) return;
  GlobalSink sk = gs;
  step_particle32(from_slot(p, base), sk, GlobalFields(), threadIdx.x);
}

__global__ void k_charge32(F32Args p, GlobalSink gs, long long n_cap) {
  const long long n = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= n_cap) return;
  uint32_t c = p.cell[n];
  if (c & kDead) return;
  GlobalSink sk = gs;
  deposit_rho(p, sk, p.w[n], cell_i(c), p.fa[n], cell_k(c), p.fb[n]);
}

// absolute FP64 state in logical order -> cell form in the species' layout
// (cell index by CELL_NUMBER clamped at 0; chunk padding slots marked)
__global__ void k_pack(GridDesc g, const double *r, const double *z, const double *vr, const double *vphi,
                       const double *vz, const double *w, const uint8_t *alive, int n, long long n_cap, int chunked,
                       uint32_t *cell, float *fa, float *fb, float *ur, float *uphi, float *uz, float *fw) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_cap) return;
  const long long j = chunked ? logical_of(p) : p;
  if (j >= n) {
    cell[p] = kPadCell;
    fa[p] = fb[p] = 0.5f;
    ur[p] = uphi[p] = uz[p] = fw[p] = 0.f;
    return;
  }
  double xr = r[j] / g.dr, xz = z[j] / g.dz;
  int i = (int)ceil(xr) - 1, k = (int)ceil(xz) - 1;
  if (i < 0) i = 0;
  if (k < 0) k = 0;
  if (i > g.nr - 2) i = g.nr - 2;
  if (k > g.nz - 2) k = g.nz - 2;
  float oa = (float)(xr - i), ob = (float)(xz - k);
  double v2 = vr[j] * vr[j] + vphi[j] * vphi[j] + vz[j] * vz[j];
  double gam = 1.0 / sqrt(1.0 - v2 / (kLightSpeed * kLightSpeed));
  cell[p] = pack(i, k) | (alive[j] == 1 ? 0u : kDead);
  fa[p] = oa;
  fb[p] = ob;
  ur[p] = (float)(gam * vr[j]);
  uphi[p] = (float)(gam * vphi[j]);
  uz[p] = (float)(gam * vz[j]);
  fw[p] = (float)w[j];
}

__global__ void k_unpack(GridDesc g, double *r, double *z, double *vr, double *vphi, double *vz, double *w,
                         uint8_t *alive, int n, int chunked, const uint32_t *cell, const float *fa, const float *fb,
                         const float *ur, const float *uphi, const float *uz, const float *fw) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const long long p = chunked ? phys_of(j) : j;
  uint32_t c = cell[p];
  r[j] = ((double)cell_i(c) + (double)fa[p]) * g.dr;
  z[j] = ((double)cell_k(c) + (double)fb[p]) * g.dz;
  double u0 = ur[p], u1 = uphi[p], u2 = uz[p];
  double ig = 1.0 / sqrt(1.0 + (u0 * u0 + u1 * u1 + u2 * u2) / (kLightSpeed * kLightSpeed));
  vr[j] = u0 * ig;
  vphi[j] = u1 * ig;
  vz[j] = u2 * ig;
  w[j] = fw[p];
  alive[j] = (c & kDead) ? 0 : 1;
}

// Bunch::bunch_inject (src/bunch.cpp:104-121) on the cell form: z from the
// draw, v_z set to the bunch speed, the radial state kept
__device__ __forceinline__ void inject_one32(const GridDesc &g, uint32_t *cell, float *fb, float *ur, float *uphi,
                                             float *uz, const double *u, unsigned j, double dl, double velocity) {
  double zabs = dl * u[j] + g.dz / 2.0;
  double xz = zabs / g.dz;
  int k = (int)ceil(xz) - 1;
  if (k < 0) k = 0;
  uint32_t c = cell[j];
  double u0 = ur[j], u1 = uphi[j], u2 = uz[j];
  double ig = 1.0 / sqrt(1.0 + (u0 * u0 + u1 * u1 + u2 * u2) / (kLightSpeed * kLightSpeed));
  double v0 = u0 * ig, v1 = u1 * ig, v2 = velocity;
  double gam = 1.0 / sqrt(1.0 - (v0 * v0 + v1 * v1 + v2 * v2) / (kLightSpeed * kLightSpeed));
  cell[j] = pack(cell_i(c), k);
  fb[j] = (float)(xz - k);
  ur[j] = (float)(gam * v0);
  uphi[j] = (float)(gam * v1);
  uz[j] = (float)(gam * v2);
}
__global__ void k_inject32(GridDesc g, uint32_t *cell, float *fb, float *ur, float *uphi, float *uz,
                           const double *u, int n, unsigned start, unsigned count, double dl, double velocity) {
  unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  unsigned j = t + start;
  if (j >= (unsigned)n) return;
  inject_one32(g, cell, fb, ur, uphi, uz, u, j, dl, velocity);
}

// ensembles: the injections of every (instance, bunch) pair for the
// ensemble's current step, read from the step counter on the device (so the
// launch can sit in the step's graph); item y of the grid, particle x
__global__ void k_ens_inject(const EnsInjectItem *__restrict__ items, const long long *__restrict__ step) {
  const EnsInjectItem &it = items[blockIdx.y];
  const long long s = *step;
  if (s >= it.table_len) return;
  const int start = it.start[s];
  if (start < 0) return;
  for (unsigned t = blockIdx.x * blockDim.x + threadIdx.x; t < (unsigned)it.count; t += gridDim.x * blockDim.x) {
    const unsigned j = t + (unsigned)start;
    if (j >= (unsigned)it.n) break;
    inject_one32(it.g, it.cell, it.fb, it.ur, it.uphi, it.uz, it.u, j, it.dl, it.velocity);
  }
}
__global__ void k_ens_count(long long *step) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *step += 1;
}

F32Args args32(Ctx &c, Species &sp) {
  F32Args p;
  const GridDesc &g = c.g;
  p.nr = g.nr;
  p.nz = g.nz;
  p.hr = (float)(g.dt / (2.0 * g.dr));
  p.hz = (float)(g.dt / (2.0 * g.dz));
  p.inv_c2 = (float)(1.0 / (kLightSpeed * kLightSpeed));
  double c1 = sp.charge * g.dt / (2.0 * sp.mass);
  p.c1 = (float)c1;
  if (c.pusher == 1) {
    p.escale = (float)(2.0 * c1);
    p.bscale = (float)(kMagnConst * c1 * 2.0 * c1 * c1);
  } else {
    p.escale = (float)c1;
    p.bscale = (float)(kMagnConst * c1);
  }
  p.kz = (float)(sp.charge / (4.0 * kPi * g.dr * g.dr * g.dt));
  p.kr = (float)(sp.charge / (2.0 * kPi * g.dr * g.dz * g.dt));
  p.kp = (float)(sp.charge / (4.0 * kPi * g.dz * g.dr * g.dr));
  p.pusher = c.pusher;
  p.is_bunch = sp.kind == 1;
  p.want_rho = sp.kind == 1;
  p.n = sp.n;
  p.grid = c.grid;
  p.rows = c.rows;
  p.err = c.err;
  p.cell = (uint32_t *)sp.f32.cell;
  p.fa = sp.f32.fr;
  p.fb = sp.f32.fz;
  p.ur = sp.f32.vr;
  p.uphi = sp.f32.vphi;
  p.uz = sp.f32.vz;
  p.w = sp.f32.w;
  return p;
}

GlobalSink sink32(Ctx &c) {
  GlobalSink gs;
  gs.nr = c.g.nr;
  gs.nz = c.g.nz;
  gs.j = c.a[kJR];
  gs.rho = c.a[kRhoBunch];
  gs.err = c.err;
  return gs;
}

bool ensure_f32(Species &sp) {
  if (sp.f32.cell) return true;
  size_t n = sp.n_cap > 0 ? (size_t)sp.n_cap : 1;
  bool ok = hipMalloc((void **)&sp.f32.cell, n * sizeof(int32_t)) == hipSuccess;
  ok = ok && hipMalloc((void **)&sp.f32.fr, n * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc((void **)&sp.f32.fz, n * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc((void **)&sp.f32.vr, n * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc((void **)&sp.f32.vphi, n * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc((void **)&sp.f32.vz, n * sizeof(float)) == hipSuccess;
  ok = ok && hipMalloc((void **)&sp.f32.w, n * sizeof(float)) == hipSuccess;
  return ok;
}

unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

}  // namespace

template <int PPT, int WAVES>
void launch_tile_t(hipStream_t st, const F32Args &p, const GlobalSink &gs, long long n_cap) {
  constexpr long long kTile = (long long)WAVES * PPT / kPPT * kChunk;
  unsigned blocks = (unsigned)((n_cap + kTile - 1) / kTile);
  hipLaunchKernelGGL((k_fused32_tiled<PPT, WAVES>), dim3(blocks), dim3(WAVES * 64), 0, st, p, gs, n_cap);
}

void launch_tile(int tile, hipStream_t st, const F32Args &p, const GlobalSink &gs, long long n_cap) {
  switch (tile) {
    case 0: launch_tile_t<16, 4>(st, p, gs, n_cap); break;
    case 1: launch_tile_t<16, 2>(st, p, gs, n_cap); break;
    case 2: launch_tile_t<16, 1>(st, p, gs, n_cap); break;
    case 3: launch_tile_t<8, 4>(st, p, gs, n_cap); break;
    case 4: launch_tile_t<8, 2>(st, p, gs, n_cap); break;
    default: launch_tile_t<4, 4>(st, p, gs, n_cap); break;
  }
}

void launch_fused_f32(Ctx &c, Species &sp) {
  if (sp.n_cap <= 0 || !sp.f32.cell) return;
  if (sp.chunked) {
    launch_tile(c.tile, c.stream, args32(c, sp), sink32(c), sp.n_cap);
  } else {
    // the most particles per lane that still give 1024 blocks
    const long long n = sp.n_cap, per = (long long)kWavesPerBlock * 64;
    const F32Args a = args32(c, sp);
    const GlobalSink g = sink32(c);
    if (n >= 1024 * per * kPPT)
      hipLaunchKernelGGL(k_fused32_bunch<kPPT>, dim3((unsigned)((n + per * kPPT - 1) / (per * kPPT))), dim3(256), 0,
                         c.stream, a, g, n);
    else if (n >= 1024 * per * 4)
      hipLaunchKernelGGL(k_fused32_bunch<4>, dim3((unsigned)((n + per * 4 - 1) / (per * 4))), dim3(256), 0, c.stream,
                         a, g, n);
    else
      hipLaunchKernelGGL(k_fused32_bunch<1>, dim3((unsigned)((n + per - 1) / per)), dim3(256), 0, c.stream, a, g, n);
  }
}

void launch_charge_weighting_f32(Ctx &c, Species &sp, double *rho) {
  if (sp.n_cap <= 0 || !sp.f32.cell) return;
  GlobalSink gs = sink32(c);
  if (rho) gs.rho = rho;
  hipLaunchKernelGGL(k_charge32, dim3(blocks_for(sp.n_cap)), dim3(256), 0, c.stream, args32(c, sp), gs, sp.n_cap);
}

void launch_bunch_inject_f32(Ctx &c, Species &sp, unsigned start, unsigned count, double dl) {
  if (count == 0 || !sp.f32.cell) return;
  hipLaunchKernelGGL(k_inject32, dim3((count + 255) / 256), dim3(256), 0, c.stream, c.g,
                     (uint32_t *)sp.f32.cell, sp.f32.fz, sp.f32.vr, sp.f32.vphi, sp.f32.vz, sp.inject_u, sp.n,
                     start, count, dl, sp.velocity);
}

bool launch_f32_from_f64(Ctx &c, Species &sp) {
  if (sp.n_cap <= 0) return true;
  if (!ensure_f32(sp)) return false;
  hipLaunchKernelGGL(k_pack, dim3(blocks_for(sp.n_cap)), dim3(256), 0, c.stream, c.g, sp.f64.r, sp.f64.z, sp.f64.vr,
                     sp.f64.vphi, sp.f64.vz, sp.f64.w, sp.f64.alive, sp.n, sp.n_cap, sp.chunked ? 1 : 0,
                     (uint32_t *)sp.f32.cell, sp.f32.fr, sp.f32.fz, sp.f32.vr, sp.f32.vphi, sp.f32.vz, sp.f32.w);
  return true;
}

void launch_f64_from_f32(Ctx &c, Species &sp) {
  if (sp.n <= 0 || !sp.f32.cell) return;
  hipLaunchKernelGGL(k_unpack, dim3(blocks_for(sp.n)), dim3(256), 0, c.stream, c.g, sp.f64.r, sp.f64.z,
                     sp.f64.vr, sp.f64.vphi, sp.f64.vz, sp.f64.w, sp.f64.alive, sp.n, sp.chunked ? 1 : 0,
                     (const uint32_t *)sp.f32.cell, sp.f32.fr, sp.f32.fz, sp.f32.vr, sp.f32.vphi, sp.f32.vz,
                     sp.f32.w);
}

struct EnsParticles {
  int n_args = 0;
  F32Args *args = nullptr;
  GlobalSink *sinks = nullptr;
  // work list per tile shape (plasma tiles and bunch blocks), and at index
  // kNumTiles the one-chunk tiles of the sparse plasma species (below
  // kSparsePpc particles per cell), which the lists of the larger shapes leave
  // out: a tile spans tile / ppc cells, and one chunk is the smallest tile
  EnsItem *items[kNumTiles + 1] = {};
  int n_items[kNumTiles + 1] = {};
};

namespace {
constexpr int kSparsePpc = 16;
constexpr int kChunkTile = kNumTiles - 1;  // 4 x 4: one chunk per tile
static_assert(kTileShapes[kChunkTile][0] * kTileShapes[kChunkTile][1] == kPPT, "one-chunk tile shape");
// the argument table and the work lists of the ensemble's species
void ens_lists(const std::vector<Ctx *> &ctxs, std::vector<F32Args> *args,
               std::vector<EnsItem> items[kNumTiles + 1]) {
  int a = 0;
  for (Ctx *cp : ctxs) {
    Ctx &c = *cp;
    // every species by index, a bunch whether or not it has been created
    // (until its injection its particles are dead and do nothing)
    for (int j = 0; j < (int)c.species.size(); j++) {
      Species &sp = c.species[j];
      if (sp.n_cap <= 0 || !sp.f32.cell) continue;
      args->push_back(args32(c, sp));
      const bool sparse = sp.chunked && sp.n < (long long)kSparsePpc * (c.g.nr - 1) * (c.g.nz - 1);
      for (int t = 0; t <= kNumTiles; t++) {
        // the one-chunk shape's own list holds every species; the others
        // leave the sparse ones to the extra list
        if (t < kNumTiles && t != kChunkTile && sparse) continue;
        if (t == kNumTiles && !sparse) continue;
        const int ts = t == kNumTiles ? kChunkTile : t;
        const long long tile = (long long)kTileShapes[ts][1] * kTileShapes[ts][0] / kPPT * kChunk;
        for (long long b = 0; b < sp.n_cap; b += tile) items[t].push_back(EnsItem{a, sp.chunked ? 0 : 1, b, sp.n_cap});
      }
      a++;
    }
  }
}
}  // namespace

// the argument tables (a sort swaps the particle arrays); the work lists
// depend only on the species sizes and stay as built
int ens_particles_update(EnsParticles *e, const std::vector<Ctx *> &ctxs) {
  std::vector<F32Args> args;
  std::vector<GlobalSink> sinks;
  for (Ctx *cp : ctxs)
    for (int j = 0; j < (int)cp->species.size(); j++) {
      Species &sp = cp->species[j];
      if (sp.n_cap <= 0 || !sp.f32.cell) continue;
      args.push_back(args32(*cp, sp));
      sinks.push_back(sink32(*cp));
    }
  if ((int)args.size() != e->n_args) return -1;  // the species may not change
  if (hipMemcpy(e->args, args.data(), args.size() * sizeof(F32Args), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->sinks, sinks.data(), sinks.size() * sizeof(GlobalSink), hipMemcpyHostToDevice) != hipSuccess)
    return -1;
  return 0;
}

EnsParticles *ens_particles_build(const std::vector<Ctx *> &ctxs) {
  auto *e = new EnsParticles;
  std::vector<F32Args> args;
  std::vector<EnsItem> items[kNumTiles + 1];
  ens_lists(ctxs, &args, items);
  e->n_args = (int)args.size();
  bool ok = hipMalloc((void **)&e->args, std::max(1, e->n_args) * sizeof(F32Args)) == hipSuccess &&
            hipMalloc((void **)&e->sinks, std::max(1, e->n_args) * sizeof(GlobalSink)) == hipSuccess;
  for (int t = 0; ok && t <= kNumTiles; t++) {
    e->n_items[t] = (int)items[t].size();
    ok = hipMalloc((void **)&e->items[t], std::max<size_t>(1, items[t].size()) * sizeof(EnsItem)) == hipSuccess &&
         (items[t].empty() || hipMemcpy(e->items[t], items[t].data(), items[t].size() * sizeof(EnsItem),
                                        hipMemcpyHostToDevice) == hipSuccess);
  }
  if (!ok || ens_particles_update(e, ctxs) != 0) {
    ens_particles_free(e);
    return nullptr;
  }
  return e;
}

void ens_particles_free(EnsParticles *e) {
  if (!e) return;
  if (e->args) (void)hipFree(e->args);
  if (e->sinks) (void)hipFree(e->sinks);
  for (auto *p : e->items)
    if (p) (void)hipFree(p);
  delete e;
}

template <int PPT, int WAVES>
void ens_launch_t(EnsParticles *e, int t, hipStream_t st) {
  const int blocks = e->n_items[t];
  if (blocks <= 0) return;
  hipLaunchKernelGGL((k_ens_fused32<PPT, WAVES>), dim3(blocks), dim3(WAVES * 64), 0, st, e->args, e->sinks,
                     e->items[t], e->n_items[t]);
}

void ens_launch_inject(const EnsInjectItem *items, int n_items, int max_count, const long long *step,
                       hipStream_t st) {
  if (n_items <= 0 || max_count <= 0) return;
  const unsigned bx = (unsigned)std::min<long long>(((long long)max_count + 255) / 256, 4096);
  hipLaunchKernelGGL(k_ens_inject, dim3(bx, (unsigned)n_items), dim3(256), 0, st, items, step);
}

void ens_launch_count(long long *step, hipStream_t st) {
  hipLaunchKernelGGL(k_ens_count, dim3(1), dim3(64), 0, st, step);
}

void ens_launch_particles(EnsParticles *e, int tile, hipStream_t st) {
  switch (tile) {
    case 0: ens_launch_t<16, 4>(e, 0, st); break;
    case 1: ens_launch_t<16, 2>(e, 1, st); break;
    case 2: ens_launch_t<16, 1>(e, 2, st); break;
    case 3: ens_launch_t<8, 4>(e, 3, st); break;
    case 4: ens_launch_t<8, 2>(e, 4, st); break;
    default: ens_launch_t<4, 4>(e, 5, st); break;
  }
  // the sparse species on one-chunk tiles (their list is empty when the
  // shape is the one-chunk shape itself, whose list holds them)
  if (tile != kChunkTile) ens_launch_t<4, 4>(e, kNumTiles, st);
}

}  // namespace pdp3
