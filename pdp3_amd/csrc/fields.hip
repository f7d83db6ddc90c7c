// Yee FDTD updates of H and E on the r-z grid (FP64 in both precision forms).
// Built with -ffp-contract=off so every expression rounds exactly as the
// reference's IEEE build does.
//
// Data layout: each component is a row-major nr x nz FP64 array, z contiguous;
// one thread per node, a wave covers 64 consecutive z nodes of one row, so
// every load and store is a coalesced 512 B line; the i-1 / i+1 neighbours of
// the radial differences are re-read from L2.
#include <algorithm>
#include <vector>

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
struct HPtrs {
  const double *er, *ephi, *ez;
  double *hr, *hphi, *hz, *hhr, *hhphi, *hhz;
};

__device__ __forceinline__ void h_node(const GridDesc &g, const HPtrs &f, int i, int k) {
  const double dt = g.dt, dr = g.dr, dz = g.dz;
  const size_t p = at(g, i, k);
  {
    double alpha_t = dt * (f.ephi[p + 1] - f.ephi[p]) / (dz * kMagnConst);
    double h0 = f.hr[p];
    f.hhr[p] = h0 + alpha_t / 2;
    f.hr[p] = h0 + alpha_t;
  }
  if (i < g.nr - 1) {
    const size_t pu = p + g.nz;  // (i+1, k)
    {
      double alpha_t = dt * ((f.ez[pu] - f.ez[p]) / dr - (f.er[p + 1] - f.er[p]) / dz) / kMagnConst;
      double h0 = f.hphi[p];
      f.hhphi[p] = h0 + alpha_t / 2;
      f.hphi[p] = h0 + alpha_t;
    }
    {
      double alpha_t = dt * ((f.ephi[pu] + f.ephi[p]) / (2.0 * dr * (i + 0.5)) + (f.ephi[pu] - f.ephi[p]) / dr) /
                       kMagnConst;
      double h0 = f.hz[p];
      f.hhz[p] = h0 - alpha_t / 2;
      f.hz[p] = h0 - alpha_t;
    }
  }
}

__global__ void __launch_bounds__(256) k_h_update(GridDesc g, HPtrs f) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (k >= g.nz - 1) return;
  h_node(g, f, i, k);
}

// EField::calc_field, src/eField.cpp:85-170, with the per-cell conductivity
// coefficients koef_e = (2 eps eps0 - sigma dt)/(2 eps eps0 + sigma dt) and
// koef_h = 2 dt/(2 eps eps0 + sigma dt) (:105-110) computed once on the host
// in the same arithmetic.  Rows 0..nr-2, columns 0..nz-2:
//   i = 0: E_r for k >= 1 (:99-112), E_z with the 4/dr H_phi axis term (:114-129)
//   i >= 1, k >= 1: E_r, E_phi, E_z (:131-152);  i >= 1, k = 0: E_z (:154-168)
struct EPtrs {
  const double *hr, *hphi, *hz, *jr, *jphi, *jz, *ke, *kh;
  double *er, *ephi, *ez;
};

__device__ __forceinline__ void e_node(const GridDesc &g, const EPtrs &f, int i, int k) {
  const double dr = g.dr, dz = g.dz;
  const size_t p = at(g, i, k);
  const double ke = f.ke[p], kh = f.kh[p];
  if (i == 0) {
    if (k >= 1) f.er[p] = f.er[p] * ke - (f.jr[p] + (f.hphi[p] - f.hphi[p - 1]) / dz) * kh;
    f.ez[p] = f.ez[p] * ke - (f.jz[p] - 4.0 / dr * f.hphi[p]) * kh;
    return;
  }
  const size_t pd = p - g.nz;  // (i-1, k)
  if (k >= 1) {
    f.er[p] = f.er[p] * ke - (f.jr[p] + (f.hphi[p] - f.hphi[p - 1]) / dz) * kh;
    f.ephi[p] = f.ephi[p] * ke - (f.jphi[p] - (f.hr[p] - f.hr[p - 1]) / dz + (f.hz[p] - f.hz[pd]) / dr) * kh;
  }
  f.ez[p] = f.ez[p] * ke - (f.jz[p] - (f.hphi[p] - f.hphi[pd]) / dr - (f.hphi[p] + f.hphi[pd]) / (2.0 * dr * i)) * kh;
}

__global__ void __launch_bounds__(256) k_e_update(GridDesc g, EPtrs f) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (k >= g.nz - 1) return;
  e_node(g, f, i, k);
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

// ---- ensembles: every instance's update in one launch ----
// a work item is one run of 256 columns of one row of one instance; blocks
// walk the item list with a stride of the grid size
struct EnsRow {
  int inst, i, k0;
};
struct EnsGrid {
  GridDesc g;
  HPtrs h;
  EPtrs e;
  double *jr, *jphi, *jz, *rho;
};

__global__ void __launch_bounds__(256) k_ens_h(const EnsGrid *__restrict__ grids, const EnsRow *__restrict__ rows,
                                               int n_rows) {
  for (int it = blockIdx.x; it < n_rows; it += gridDim.x) {
    const EnsRow r = rows[it];
    const EnsGrid &G = grids[r.inst];
    const int k = r.k0 + threadIdx.x;
    if (k < G.g.nz - 1) h_node(G.g, G.h, r.i, k);
  }
}

__global__ void __launch_bounds__(256) k_ens_e(const EnsGrid *__restrict__ grids, const EnsRow *__restrict__ rows,
                                               int n_rows) {
  for (int it = blockIdx.x; it < n_rows; it += gridDim.x) {
    const EnsRow r = rows[it];
    const EnsGrid &G = grids[r.inst];
    const int k = r.k0 + threadIdx.x;
    if (k < G.g.nz - 1) e_node(G.g, G.e, r.i, k);
  }
}

// Current::reset_j and the bunch density reset of every instance
__global__ void __launch_bounds__(256) k_ens_zero(const EnsGrid *__restrict__ grids) {
  const EnsGrid &G = grids[blockIdx.y];
  const size_t n = (size_t)G.g.nr * (size_t)G.g.nz;
  for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
    G.jr[p] = 0.0;
    G.jphi[p] = 0.0;
    G.jz[p] = 0.0;
    G.rho[p] = 0.0;
  }
}

}  // namespace

HPtrs hptrs(Ctx &c) {
  return HPtrs{c.a[kER], c.a[kEPhi], c.a[kEZ], c.a[kHR], c.a[kHPhi], c.a[kHZ], c.a[kHHR], c.a[kHHPhi], c.a[kHHZ]};
}
EPtrs eptrs(Ctx &c) {
  return EPtrs{c.a[kHR], c.a[kHPhi], c.a[kHZ], c.a[kJR], c.a[kJPhi], c.a[kJZ], c.koef_e, c.koef_h,
               c.a[kER], c.a[kEPhi], c.a[kEZ]};
}

void launch_h_update(Ctx &c) {
  dim3 block(256);
  dim3 grid((c.g.nz - 1 + 255) / 256, c.g.nr);
  hipLaunchKernelGGL(k_h_update, grid, block, 0, c.stream, c.g, hptrs(c));
}

void launch_e_update(Ctx &c) {
  dim3 block(256);
  dim3 grid((c.g.nz - 1 + 255) / 256, c.g.nr - 1);
  hipLaunchKernelGGL(k_e_update, grid, block, 0, c.stream, c.g, eptrs(c));
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

struct EnsFields {
  int n_inst = 0, n_h = 0, n_e = 0;
  EnsGrid *grids = nullptr;
  EnsRow *hrows = nullptr, *erows = nullptr;
};

EnsFields *ens_fields_build(const std::vector<Ctx *> &ctxs) {
  std::vector<EnsGrid> g;
  std::vector<EnsRow> hr, er;
  for (size_t j = 0; j < ctxs.size(); j++) {
    Ctx &c = *ctxs[j];
    g.push_back(EnsGrid{c.g, hptrs(c), eptrs(c), c.a[kJR], c.a[kJPhi], c.a[kJZ], c.a[kRhoBunch]});
    for (int i = 0; i < c.g.nr; i++)
      for (int k0 = 0; k0 < c.g.nz - 1; k0 += 256) {
        hr.push_back(EnsRow{(int)j, i, k0});
        if (i < c.g.nr - 1) er.push_back(EnsRow{(int)j, i, k0});
      }
  }
  auto *e = new EnsFields;
  e->n_inst = (int)g.size();
  e->n_h = (int)hr.size();
  e->n_e = (int)er.size();
  if (hipMalloc((void **)&e->grids, g.size() * sizeof(EnsGrid)) != hipSuccess ||
      hipMalloc((void **)&e->hrows, hr.size() * sizeof(EnsRow)) != hipSuccess ||
      hipMalloc((void **)&e->erows, er.size() * sizeof(EnsRow)) != hipSuccess ||
      hipMemcpy(e->grids, g.data(), g.size() * sizeof(EnsGrid), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->hrows, hr.data(), hr.size() * sizeof(EnsRow), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(e->erows, er.data(), er.size() * sizeof(EnsRow), hipMemcpyHostToDevice) != hipSuccess) {
    ens_fields_free(e);
    return nullptr;
  }
  return e;
}

void ens_fields_free(EnsFields *e) {
  if (!e) return;
  if (e->grids) (void)hipFree(e->grids);
  if (e->hrows) (void)hipFree(e->hrows);
  if (e->erows) (void)hipFree(e->erows);
  delete e;
}

void ens_launch_h(EnsFields *e, hipStream_t st) {
  const int blocks = std::min(e->n_h, 256 * 16);
  hipLaunchKernelGGL(k_ens_h, dim3(blocks), dim3(256), 0, st, e->grids, e->hrows, e->n_h);
}

void ens_launch_e(EnsFields *e, hipStream_t st) {
  const int blocks = std::min(e->n_e, 256 * 16);
  hipLaunchKernelGGL(k_ens_e, dim3(blocks), dim3(256), 0, st, e->grids, e->erows, e->n_e);
}

void ens_launch_zero(EnsFields *e, hipStream_t st) {
  hipLaunchKernelGGL(k_ens_zero, dim3(64, e->n_inst), dim3(256), 0, st, e->grids);
}

}  // namespace pdp3
