// Synthetic plasma loader: ppc particles in every cell, positions uniform
// inside the cell (clipped to the reflecting walls at dr/2, R-dr/2, dz/2,
// L-dz/2), weights proportional to r as the reference's cylindrical loader
// does (src/particles.cpp:316-330), isotropic Maxwellian velocities.  All
// draws are counter-based (Philox-4x32-10, Salmon et al., SC'11) keyed on
// (seed, species, cell, slot), so the state does not depend on the launch
// shape and any device reproduces it.
#include "context.hpp"

namespace pdp3 {

namespace {

struct U4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

// Philox-4x32 with 10 rounds (multipliers and Weyl constants of the paper)
__device__ U4 philox(U4 ctr, uint32_t k0, uint32_t k1) {
  for (int round = 0; round < 10; round++) {
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, ctr.x, hi0, lo0);
    mulhilo(0xCD9E8D57u, ctr.z, hi1, lo1);
    ctr = U4{hi1 ^ ctr.y ^ k0, lo1, hi0 ^ ctr.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return ctr;
}

// 53-bit uniform in [0, 1) from two words
__device__ __forceinline__ double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

__global__ void k_init_uniform(GridDesc g, double *r, double *z, double *vr, double *vphi, double *vz, double *w,
                               uint8_t *alive, long long n, int ppc, uint32_t seed_lo, uint32_t seed_hi,
                               uint32_t species, double density, int ramp_cells, double sigma_v, int to_phys) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const long long dst = to_phys ? phys_of(p) : p;
  const int ncz = g.nz - 1;
  long long cell = p / ppc;
  uint32_t slot = (uint32_t)(p - cell * ppc);
  int i = (int)(cell / ncz);
  int k = (int)(cell - (long long)i * ncz);
  const double dr = g.dr, dz = g.dz;
  double r_lo = fmax(i * dr, dr / 2.), r_hi = fmin((i + 1) * dr, g.r_size - dr / 2.);
  double z_lo = fmax(k * dz, dz / 2.), z_hi = fmin((k + 1) * dz, g.z_size - dz / 2.);
  U4 a = philox(U4{slot, (uint32_t)cell, species, 0u}, seed_lo, seed_hi);
  U4 b = philox(U4{slot, (uint32_t)cell, species, 1u}, seed_lo, seed_hi);
  U4 c = philox(U4{slot, (uint32_t)cell, species, 2u}, seed_lo, seed_hi);
  double ur = u53(a.x, a.y), uz = u53(a.z, a.w);
  double rr = r_lo + ur * (r_hi - r_lo);
  double zz = z_lo + uz * (z_hi - z_lo);
  // Box-Muller on two pairs of 53-bit uniforms (1 - u keeps the log finite)
  double u1 = 1.0 - u53(b.x, b.y), u2 = u53(b.z, b.w);
  double u3 = 1.0 - u53(c.x, c.y), u4 = u53(c.z, c.w);
  double m1 = sqrt(-2.0 * log(u1)), m2 = sqrt(-2.0 * log(u3));
  double n0 = m1 * cos(2.0 * kPi * u2), n1 = m1 * sin(2.0 * kPi * u2);
  double n2 = m2 * cos(2.0 * kPi * u4);
  double dens = density;
  if (ramp_cells > 0) {
    double zr = zz / (ramp_cells * dz);
    if (zr < 1.0) dens = density * zr;
  }
  r[dst] = rr;
  z[dst] = zz;
  vr[dst] = sigma_v * n0;
  vphi[dst] = sigma_v * n1;
  vz[dst] = sigma_v * n2;
  // real particles represented: n(z) * 2 pi r (r_hi - r_lo)(z_hi - z_lo) / ppc
  w[dst] = dens * 2.0 * kPi * rr * (r_hi - r_lo) * (z_hi - z_lo) / ppc;
  alive[dst] = 1;
}

__global__ void k_mark_padding(uint8_t *alive, long long n, long long n_cap) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_cap) return;
  if (logical_of(p) >= n) alive[p] = kPadSlot;
}

__global__ void k_to_phys(const double *src, double *dst, long long n, long long n_cap) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_cap) return;
  long long s = logical_of(p);
  dst[p] = s < n ? src[s] : 0.0;
}
__global__ void k_to_logical(const double *src, double *dst, long long n) {
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  dst[s] = src[phys_of(s)];
}
__global__ void k_alive_to_phys(const uint8_t *src, uint8_t *dst, long long n, long long n_cap) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_cap) return;
  long long s = logical_of(p);
  dst[p] = s < n ? src[s] : kPadSlot;
}
__global__ void k_alive_to_logical(const uint8_t *src, uint8_t *dst, long long n) {
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  dst[s] = src[phys_of(s)];
}

}  // namespace

void launch_mark_padding(Ctx &c, uint8_t *alive, long long n, long long n_cap) {
  if (n_cap > n)
    hipLaunchKernelGGL(k_mark_padding, dim3((unsigned)((n_cap + 255) / 256)), dim3(256), 0, c.stream, alive, n,
                       n_cap);
}
void launch_to_phys(Ctx &c, const double *logical, double *phys, long long n, long long n_cap) {
  if (n_cap > 0)
    hipLaunchKernelGGL(k_to_phys, dim3((unsigned)((n_cap + 255) / 256)), dim3(256), 0, c.stream, logical, phys, n,
                       n_cap);
}
void launch_to_logical(Ctx &c, const double *phys, double *logical, long long n) {
  if (n > 0)
    hipLaunchKernelGGL(k_to_logical, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, phys, logical, n);
}
void launch_alive_to_phys(Ctx &c, const uint8_t *logical, uint8_t *phys, long long n, long long n_cap) {
  if (n_cap > 0)
    hipLaunchKernelGGL(k_alive_to_phys, dim3((unsigned)((n_cap + 255) / 256)), dim3(256), 0, c.stream, logical,
                       phys, n, n_cap);
}
void launch_alive_to_logical(Ctx &c, const uint8_t *phys, uint8_t *logical, long long n) {
  if (n > 0)
    hipLaunchKernelGGL(k_alive_to_logical, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, phys,
                       logical, n);
}

void launch_init_uniform(Ctx &c, Species &sp, int species_index, uint64_t seed, int ppc, double density,
                         int ramp_cells, double temperature_ev, bool to_phys) {
  long long n = sp.n;
  if (n <= 0) return;
  double sigma_v = sqrt(temperature_ev * kElCharge / sp.mass);
  unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_init_uniform, dim3(blocks), dim3(256), 0, c.stream, c.g, sp.f64.r, sp.f64.z, sp.f64.vr,
                     sp.f64.vphi, sp.f64.vz, sp.f64.w, sp.f64.alive, n, ppc, (uint32_t)seed,
                     (uint32_t)(seed >> 32), (uint32_t)species_index, density, ramp_cells, sigma_v, to_phys ? 1 : 0);
}

}  // namespace pdp3
