// Yee FDTD updates of H and E on the r-z grid (FP64 in both precision forms).
// Built with -ffp-contract=off so every expression rounds exactly as the
// reference's IEEE build does.
//
// Data layout: each component is a row-major nr x nz FP64 array, z contiguous;
// one thread per node, a wave covers 64 consecutive z nodes of one row, so
// every load and store is a coalesced 512 B line; the i-1 / i+1 neighbours of
// the radial differences are re-read from L2.
#include "context.hpp"

namespace pdp3 {

namespace {

__device__ __forceinline__ size_t at(const GridDesc &g, int i, int k) {
  return (size_t)i * (size_t)g.nz + (size_t)k;
}

// HField::calc_field, src/hField.cpp:93-157.  H_r on all nr rows (the last
// row is the loop at :105-117), H_phi and H_z on rows 0..nr-2; k < nz-1.
// H_half = H_old + delta/2, H = H_old + delta (the half-time copy is what the
// particle gather reads, src/hField.cpp:159-267).
__global__ void __launch_bounds__(256) k_h_update(GridDesc g, const double *__restrict__ er,
                                                  const double *__restrict__ ephi,
                                                  const double *__restrict__ ez,
                                                  double *__restrict__ hr, double *__restrict__ hphi,
                                                  double *__restrict__ hz, double *__restrict__ hhr,
                                                  double *__restrict__ hhphi,
                                                  double *__restrict__ hhz) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (k >= g.nz - 1) return;
  const double dt = g.dt, dr = g.dr, dz = g.dz;
  const size_t p = at(g, i, k);
  {
    double alpha_t = dt * (ephi[p + 1] - ephi[p]) / (dz * kMagnConst);
    double h0 = hr[p];
    hhr[p] = h0 + alpha_t / 2;
    hr[p] = h0 + alpha_t;
  }
  if (i < g.nr - 1) {
    const size_t pu = p + g.nz;  // (i+1, k)
    {
      double alpha_t = dt * ((ez[pu] - ez[p]) / dr - (er[p + 1] - er[p]) / dz) / kMagnConst;
      double h0 = hphi[p];
      hhphi[p] = h0 + alpha_t / 2;
      hphi[p] = h0 + alpha_t;
    }
    {
      double alpha_t = dt * ((ephi[pu] + ephi[p]) / (2.0 * dr * (i + 0.5)) + (ephi[pu] - ephi[p]) / dr) /
                       kMagnConst;
      double h0 = hz[p];
      hhz[p] = h0 - alpha_t / 2;
      hz[p] = h0 - alpha_t;
    }
  }
}

// EField::calc_field, src/eField.cpp:85-170, with the per-cell conductivity
// coefficients koef_e = (2 eps eps0 - sigma dt)/(2 eps eps0 + sigma dt) and
// koef_h = 2 dt/(2 eps eps0 + sigma dt) (:105-110) computed once on the host
// in the same arithmetic.  Rows 0..nr-2, columns 0..nz-2:
//   i = 0: E_r for k >= 1 (:99-112), E_z with the 4/dr H_phi axis term (:114-129)
//   i >= 1, k >= 1: E_r, E_phi, E_z (:131-152);  i >= 1, k = 0: E_z (:154-168)
__global__ void __launch_bounds__(256) k_e_update(GridDesc g, const double *__restrict__ hr,
                                                  const double *__restrict__ hphi,
                                                  const double *__restrict__ hz,
                                                  const double *__restrict__ jr,
                                                  const double *__restrict__ jphi,
                                                  const double *__restrict__ jz,
                                                  const double *__restrict__ ke_a,
                                                  const double *__restrict__ kh_a,
                                                  double *__restrict__ er, double *__restrict__ ephi,
                                                  double *__restrict__ ez) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (k >= g.nz - 1) return;
  const double dr = g.dr, dz = g.dz;
  const size_t p = at(g, i, k);
  const double ke = ke_a[p], kh = kh_a[p];
  if (i == 0) {
    if (k >= 1) er[p] = er[p] * ke - (jr[p] + (hphi[p] - hphi[p - 1]) / dz) * kh;
    ez[p] = ez[p] * ke - (jz[p] - 4.0 / dr * hphi[p]) * kh;
    return;
  }
  const size_t pd = p - g.nz;  // (i-1, k)
  if (k >= 1) {
    er[p] = er[p] * ke - (jr[p] + (hphi[p] - hphi[p - 1]) / dz) * kh;
    ephi[p] = ephi[p] * ke - (jphi[p] - (hr[p] - hr[p - 1]) / dz + (hz[p] - hz[pd]) / dr) * kh;
  }
  ez[p] = ez[p] * ke - (jz[p] - (hphi[p] - hphi[pd]) / dr - (hphi[p] + hphi[pd]) / (2.0 * dr * i)) * kh;
}

__global__ void k_fill(double *p, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] = v;
}


// Gauss-law residual R = div E - rho / eps0 on the nodes (r_i, z_k), the
// points where Particles::charge_weighting puts rho (src/particles.cpp:101-226:
// node k collects the cloud part between (k-1/2) dz and (k+1/2) dz, node i the
// ring between (i-1/2) dr and (i+1/2) dr).  With E_r at (r_{i+1/2}, z_k) and
// E_z at (r_i, z_{k+1/2}) (the staggering of src/eField.cpp:85-170), the
// divergence the E update leaves invariant (div curl H = 0 exactly) is
//   i >= 1: (r_{i+1/2} E_r[i] - r_{i-1/2} E_r[i-1]) / (r_i dr) + (E_z[k] - E_z[k-1]) / dz
//   i = 0:  4 E_r[0] / dr + (E_z[k] - E_z[k-1]) / dz   (flux through the r = dr/2 disc)
// so with a charge-conserving deposit R changes only by round-off (outside
// the conductivity layer and the bunch injection nodes).  R = 0 on k = 0,
// k = nz-1 and i = nr-1, where the divergence has no inner stencil.
__global__ void __launch_bounds__(256) k_gauss_residual(GridDesc g, const double *__restrict__ er,
                                                        const double *__restrict__ ez,
                                                        const double *__restrict__ rho,
                                                        double *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (k >= g.nz) return;
  const size_t p = at(g, i, k);
  if (k == 0 || k == g.nz - 1 || i == g.nr - 1) {
    out[p] = 0.0;
    return;
  }
  double div;
  const double dzt = (ez[p] - ez[p - 1]) / g.dz;
  if (i == 0) {
    div = 4.0 * er[p] / g.dr + dzt;
  } else {
    const double ri = (double)i;
    div = ((ri + 0.5) * er[p] - (ri - 0.5) * er[p - g.nz]) / (ri * g.dr) + dzt;
  }
  out[p] = div - rho[p] / kEpsilon0;
}

}  // namespace

void launch_h_update(Ctx &c) {
  dim3 block(256);
  dim3 grid((c.g.nz - 1 + 255) / 256, c.g.nr);
  hipLaunchKernelGGL(k_h_update, grid, block, 0, c.stream, c.g, c.a[kER], c.a[kEPhi], c.a[kEZ], c.a[kHR],
                     c.a[kHPhi], c.a[kHZ], c.a[kHHR], c.a[kHHPhi], c.a[kHHZ]);
}

void launch_e_update(Ctx &c) {
  dim3 block(256);
  dim3 grid((c.g.nz - 1 + 255) / 256, c.g.nr - 1);
  hipLaunchKernelGGL(k_e_update, grid, block, 0, c.stream, c.g, c.a[kHR], c.a[kHPhi], c.a[kHZ], c.a[kJR],
                     c.a[kJPhi], c.a[kJZ], c.koef_e, c.koef_h, c.a[kER], c.a[kEPhi], c.a[kEZ]);
}

void launch_gauss_residual(Ctx &c, const double *rho, double *out) {
  dim3 grid((c.g.nz + 255) / 256, c.g.nr);
  hipLaunchKernelGGL(k_gauss_residual, grid, dim3(256), 0, c.stream, c.g, c.a[kER], c.a[kEZ], rho, out);
}

void launch_fill(double *p, size_t n, double v, hipStream_t s) {
  if (n == 0) return;
  size_t blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
}

}  // namespace pdp3
