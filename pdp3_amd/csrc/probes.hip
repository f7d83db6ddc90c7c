// Probe diagnostics of one species on the device, from the resident particle
// state (no download of particles):
//   * TemperatureClassic::calc_t + normalize (src/temperatureClassic.cpp:57-158)
//   * ParticlesDensity::calc_density (src/particlesDensity.cpp:40-225)
// Both forms are read in place: the FP64 plain arrays, or the FP32 cell form
// (cell word + offsets + u), whose particle is taken at r = (i + a) dr,
// z = (k + b) dz with v = u / sqrt(1 + u^2/c^2) in FP64, exactly as its
// download.  Per-cell sums use FP64 atomics, so they match the reference's
// serial sums to FP64 round-off of the summation order.
#include "context.hpp"

namespace pdp3 {

namespace {

constexpr uint32_t kDeadBit = 0x80000000u;

// one particle of a species as the probes see it
struct ProbeParticle {
  double r, z, vr, vphi, vz, w;
  int i, k;  // CELL_NUMBER of r and z (include/geometry.h:44)
};

struct StateF64 {
  const double *r, *z, *vr, *vphi, *vz, *w;
  const uint8_t *alive;
  GridDesc g;
  __device__ __forceinline__ bool get(long long p, ProbeParticle &q) const {
    if (alive[p] != 1) return false;
    q.r = r[p];
    q.z = z[p];
    q.vr = vr[p];
    q.vphi = vphi[p];
    q.vz = vz[p];
    q.w = w[p];
    q.i = (int)ceil(q.r / g.dr) - 1;
    q.k = (int)ceil(q.z / g.dz) - 1;
    return true;
  }
};

struct StateF32 {
  const uint32_t *cell;
  const float *fa, *fb, *ur, *uphi, *uz, *w;
  GridDesc g;
  __device__ __forceinline__ bool get(long long p, ProbeParticle &q) const {
    const uint32_t c = cell[p];
    if (c & kDeadBit) return false;  // dead particles and padding slots
    q.i = (int)((c >> 16) & 0x7FFF);
    q.k = (int)(c & 0xFFFF);
    q.r = ((double)q.i + (double)fa[p]) * g.dr;
    q.z = ((double)q.k + (double)fb[p]) * g.dz;
    const double u0 = ur[p], u1 = uphi[p], u2 = uz[p];
    const double ig = 1.0 / sqrt(1.0 + (u0 * u0 + u1 * u1 + u2 * u2) / (kLightSpeed * kLightSpeed));
    q.vr = u0 * ig;
    q.vphi = u1 * ig;
    q.vz = u2 * ig;
    q.w = w[p];
    return true;
  }
};

// per cell: sum |v|, sum v_r, sum v_phi, sum v_z, count (src/temperatureClassic.cpp:83-102)
template <class S>
__global__ void __launch_bounds__(256) k_temperature_sums(S st, long long n_cap, double *sums, int *err) {
  const GridDesc &g = st.g;
  const size_t cells = (size_t)g.nr * (size_t)g.nz;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n_cap;
       p += (long long)gridDim.x * blockDim.x) {
    ProbeParticle q;
    if (!st.get(p, q)) continue;
    if ((unsigned)q.i >= (unsigned)g.nr || (unsigned)q.k >= (unsigned)g.nz) {
      atomicOr(err, kErrDomain);
      continue;
    }
    const size_t c = (size_t)q.i * (size_t)g.nz + (size_t)q.k;
    unsafeAtomicAdd(sums + c, sqrt(q.vr * q.vr + q.vphi * q.vphi + q.vz * q.vz));
    unsafeAtomicAdd(sums + cells + c, q.vr);
    unsafeAtomicAdd(sums + 2 * cells + c, q.vphi);
    unsafeAtomicAdd(sums + 3 * cells + c, q.vz);
    unsafeAtomicAdd(sums + 4 * cells + c, 1.0);
  }
}

// TemperatureClassic::normalize (src/temperatureClassic.cpp:127-158):
// v^2 = (S^2 - |V|^2) / N^2; T = m v^2 / 2 below REL_LIMIT^2, else
// m c^2 (gamma - 1) (lib::get_gamma, src/lib.cpp:45-70), divided by |charge|.
// Row nr-1 and column nz-1 are 0 (cleared, never normalised); an empty cell
// is NaN (0/0), as in the reference.
__global__ void __launch_bounds__(256) k_temperature_normalize(GridDesc g, const double *sums, double mass,
                                                               double abs_charge, double *out, int *err) {
  const size_t cells = (size_t)g.nr * (size_t)g.nz;
  const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cells) return;
  const int i = (int)(c / (size_t)g.nz), k = (int)(c % (size_t)g.nz);
  if (i == g.nr - 1 || k == g.nz - 1) {
    out[c] = 0.0;
    return;
  }
  const double s = sums[c], vr = sums[cells + c], vphi = sums[2 * cells + c], vz = sums[3 * cells + c];
  const double n = sums[4 * cells + c];
  const double vec2 = vr * vr + vphi * vphi + vz * vz;
  const double v2 = (s * s - vec2) / (n * n);
  const double rel_limit = 5e7;  // REL_LIMIT, src/temperatureClassic.cpp:4-6
  double t;
  if (v2 < rel_limit * rel_limit) {
    t = mass * v2 / 2;
  } else {
    const double beta = v2 / (kLightSpeed * kLightSpeed);
    if (beta > 1) atomicOr(err, kErrSuperluminal);
    double gamma = 1 / sqrt(1.0 - beta);
    if (isinf(gamma)) gamma = 1e100;
    t = mass * (kLightSpeed * kLightSpeed) * (gamma - 1);
  }
  out[c] = t / abs_charge;
}

// ParticlesDensity::calc_density (src/particlesDensity.cpp:40-225), the
// reference's expressions in its order (CYL_RNG_VOL, CELL_VOLUME, CYL_VOL of
// include/geometry.h:33-39); mass_array = mass * w (src/particles.cpp:329)
template <class S>
__global__ void __launch_bounds__(256) k_density(S st, long long n_cap, double mass, double *out, int *err) {
  const GridDesc &g = st.g;
  const double dr = g.dr, dz = g.dz;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n_cap;
       p += (long long)gridDim.x * blockDim.x) {
    ProbeParticle q;
    if (!st.get(p, q)) continue;
    int r_i = q.i;
    const int z_k = q.k;
    if (r_i < 0 || z_k < 0 || r_i > g.nr - 2 || z_k > g.nz - 2) {
      atomicOr(err, kErrDomain);
      continue;
    }
    const double r = q.r, z = q.z;
    const double m = mass * q.w;
    double r1, r2, r3, v_0, v_1, v_2, w00, w10, w01, w11;
    const double dz1 = (z_k + 0.5) * dz - (z - 0.5 * dz);
    const double dz2 = (z + 0.5 * dz) - (z_k + 0.5) * dz;
    if (r > dr) {
      r1 = r - 0.5 * dr;
      r2 = (r_i + 0.5) * dr;
      r3 = r + 0.5 * dr;
      v_0 = 2. * kPi * dz * dr * r;
      v_1 = kPi * dz * dr * dr * 2.0 * r_i;
      v_2 = kPi * dz * dr * dr * 2.0 * (r_i + 1);
      w00 = kPi * dz1 * (r2 * r2 - r1 * r1) / v_1;
      w10 = kPi * dz1 * (r3 * r3 - r2 * r2) / v_2;
      w01 = kPi * dz2 * (r2 * r2 - r1 * r1) / v_1;
      w11 = kPi * dz2 * (r3 * r3 - r2 * r2) / v_2;
    } else if (r <= dr / 2.) {
      r_i = 0;
      r2 = (r_i + 0.5) * dr;
      r3 = r + 0.5 * dr;
      v_0 = kPi * dz * (2. * r * r + dr * dr / 2.);
      v_1 = kPi * dz * dr * dr / 4.;
      v_2 = kPi * dz * dr * dr * 2.0 * (r_i + 1);
      w00 = kPi * dz1 * (dr * dr / 2. - r * dr + r * r) / v_1;
      w10 = kPi * dz1 * (r3 * r3 - r2 * r2) / v_2;
      w01 = kPi * dz2 * (dr * dr / 2. - r * dr + r * r) / v_1;
      w11 = kPi * dz2 * (r3 * r3 - r2 * r2) / v_2;
    } else {
      r1 = r - 0.5 * dr;
      r2 = (r_i + 0.5) * dr;
      r3 = r + 0.5 * dr;
      v_0 = 2. * kPi * dz * dr * r;
      v_1 = kPi * dz * dr * dr / 4.;
      v_2 = kPi * dz * dr * dr * 2.0 * (r_i + 1);
      w00 = kPi * dz1 * (r2 * r2 - r1 * r1) / v_1;
      w10 = kPi * dz1 * (r3 * r3 - r2 * r2) / v_2;
      w01 = kPi * dz2 * (r2 * r2 - r1 * r1) / v_1;
      w11 = kPi * dz2 * (r3 * r3 - r2 * r2) / v_2;
    }
    const double ro_p = m / kElMass / v_0;
    const size_t c = (size_t)r_i * (size_t)g.nz + (size_t)z_k;
    unsafeAtomicAdd(out + c, ro_p * w00);
    unsafeAtomicAdd(out + c + g.nz, ro_p * w10);
    unsafeAtomicAdd(out + c + 1, ro_p * w01);
    unsafeAtomicAdd(out + c + g.nz + 1, ro_p * w11);
  }
}

unsigned probe_blocks(long long n) {
  long long b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

StateF64 state64(const Ctx &c, const Species &sp) {
  return StateF64{sp.f64.r, sp.f64.z, sp.f64.vr, sp.f64.vphi, sp.f64.vz, sp.f64.w, sp.f64.alive, c.g};
}
StateF32 state32(const Ctx &c, const Species &sp) {
  return StateF32{(const uint32_t *)sp.f32.cell, sp.f32.fr, sp.f32.fz, sp.f32.vr, sp.f32.vphi, sp.f32.vz,
                  sp.f32.w, c.g};
}

}  // namespace

bool launch_temperature(Ctx &c, Species &sp, double *sums, double *out) {
  const size_t cells = c.cells();
  launch_fill(sums, 5 * cells, 0.0, c.stream);
  if (sp.n_cap > 0) {
    if (c.precision == 1) {
      if (!sp.f32.cell) return false;
      hipLaunchKernelGGL(k_temperature_sums<StateF32>, dim3(probe_blocks(sp.n_cap)), dim3(256), 0, c.stream,
                         state32(c, sp), sp.n_cap, sums, c.err);
    } else {
      hipLaunchKernelGGL(k_temperature_sums<StateF64>, dim3(probe_blocks(sp.n_cap)), dim3(256), 0, c.stream,
                         state64(c, sp), sp.n_cap, sums, c.err);
    }
  }
  hipLaunchKernelGGL(k_temperature_normalize, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, c.stream, c.g,
                     sums, sp.mass, fabs(sp.charge), out, c.err);
  return true;
}

bool launch_density(Ctx &c, Species &sp, double *out) {
  launch_fill(out, c.cells(), 0.0, c.stream);
  if (sp.n_cap <= 0) return true;
  if (c.precision == 1) {
    if (!sp.f32.cell) return false;
    hipLaunchKernelGGL(k_density<StateF32>, dim3(probe_blocks(sp.n_cap)), dim3(256), 0, c.stream, state32(c, sp),
                       sp.n_cap, sp.mass, out, c.err);
  } else {
    hipLaunchKernelGGL(k_density<StateF64>, dim3(probe_blocks(sp.n_cap)), dim3(256), 0, c.stream, state64(c, sp),
                       sp.n_cap, sp.mass, out, c.err);
  }
  return true;
}

}  // namespace pdp3
