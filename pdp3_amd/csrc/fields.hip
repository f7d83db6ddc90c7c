// Yee FDTD updates of H and E on the r-z grid (FP64 in both precision forms).
// Built with -ffp-contract=off so every expression rounds exactly as the
// reference's IEEE build does.
//
// Data layout: each component is a row-major nr x nz FP64 array, z contiguous.
// The updates run on tiles of 2 x 256 (H) and 1 x 256 (E) nodes, one thread
// per node: the block stages the tile of the components whose neighbours the
// stencil reads (E for the H update, H for the E update) with one halo row
// and column in LDS, then reads the stencil from there.  Loads and stores are
// 2 KB row segments.
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

// tiles of R x C nodes, one thread per node; the block stages the E (for H)
// or H (for E) values of its tile plus one halo row and column in LDS, one
// barrier, then every thread reads its stencil from LDS and its own node's
// values straight from HBM.  Measured (rocprofv3) at 513 x 8193 nodes: H on
// 2 x 256 tiles 0.069 ms, E on 1 x 256 tiles 0.085 ms, against 0.070 and
// 0.082 ms for per-node kernels reading the neighbours through the caches;
// 4 x 64 and 2 x 128 tiles were slower for E, and loading the node's own
// values before the barrier was slower for both.
template <int R, int C>
struct TileShape {
  static constexpr int kRows = R, kCols = C, kThreads = R * C, kStride = C + 1;
  static constexpr int kLds = (R + 1) * (C + 1);  // doubles per staged component
};
using HTile = TileShape<2, 256>;
using ETile = TileShape<1, 256>;

// H update of tile (i0, k0): E_phi and E_r need column k+1, E_phi and E_z row
// i+1.  s: 3 x kLds doubles (E_phi, E_z, E_r), s[c][r][k'] = E_c(i0+r, k0+k')
template <class T>
__device__ __forceinline__ void h_tile(const GridDesc &g, const HPtrs &f, int i0, int k0, double *s) {
  double *s_ephi = s, *s_ez = s + T::kLds, *s_er = s + 2 * T::kLds;
  // stage rows i0 .. i0+R and columns k0 .. k0+C (zeros off the grid)
  for (int t = threadIdx.x; t < T::kLds; t += T::kThreads) {
    const int sr = t / T::kStride, sc = t - sr * T::kStride;
    const int si = i0 + sr, sk = k0 + sc;
    const bool in = si < g.nr && sk < g.nz;
    const size_t q = in ? at(g, si, sk) : 0;
    s_ephi[t] = in ? f.ephi[q] : 0.0;
    s_ez[t] = in ? f.ez[q] : 0.0;
    s_er[t] = in ? f.er[q] : 0.0;
  }
  __syncthreads();
  const int r = threadIdx.x / T::kCols, c = threadIdx.x - r * T::kCols;
  const int i = i0 + r, k = k0 + c;
  if (i >= g.nr || k >= g.nz - 1) return;
  const double dt = g.dt, dr = g.dr, dz = g.dz;
  const int q = r * T::kStride + c;
  const size_t p = at(g, i, k);
  const double ephi = s_ephi[q];
  {
    double alpha_t = dt * (s_ephi[q + 1] - ephi) / (dz * kMagnConst);
    double h0 = f.hr[p];
    f.hhr[p] = h0 + alpha_t / 2;
    f.hr[p] = h0 + alpha_t;
  }
  if (i < g.nr - 1) {
    const double ephi_u = s_ephi[q + T::kStride];
    {
      double alpha_t = dt * ((s_ez[q + T::kStride] - s_ez[q]) / dr - (s_er[q + 1] - s_er[q]) / dz) / kMagnConst;
      double h0 = f.hphi[p];
      f.hhphi[p] = h0 + alpha_t / 2;
      f.hphi[p] = h0 + alpha_t;
    }
    {
      double alpha_t = dt * ((ephi_u + ephi) / (2.0 * dr * (i + 0.5)) + (ephi_u - ephi) / dr) / kMagnConst;
      double h0 = f.hz[p];
      f.hhz[p] = h0 - alpha_t / 2;
      f.hz[p] = h0 - alpha_t;
    }
  }
}

__global__ void __launch_bounds__(HTile::kThreads) k_h_update(GridDesc g, HPtrs f) {
  __shared__ double s[3 * HTile::kLds];
  h_tile<HTile>(g, f, blockIdx.y * HTile::kRows, blockIdx.x * HTile::kCols, s);
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

// E update of tile (i0, k0): H_phi and H_r need column k-1, H_phi and H_z row
// i-1.  s: 3 x kLds doubles (H_phi, H_r, H_z), s[c][r][k'] = H_c(i0-1+r, k0-1+k')
template <class T>
__device__ __forceinline__ void e_tile(const GridDesc &g, const EPtrs &f, int i0, int k0, double *s) {
  double *s_hphi = s, *s_hr = s + T::kLds, *s_hz = s + 2 * T::kLds;
  for (int t = threadIdx.x; t < T::kLds; t += T::kThreads) {
    const int sr = t / T::kStride, sc = t - sr * T::kStride;
    const int si = i0 - 1 + sr, sk = k0 - 1 + sc;
    const bool in = si >= 0 && si < g.nr && sk >= 0 && sk < g.nz;
    const size_t q = in ? at(g, si, sk) : 0;
    s_hphi[t] = in ? f.hphi[q] : 0.0;
    s_hr[t] = in ? f.hr[q] : 0.0;
    s_hz[t] = in ? f.hz[q] : 0.0;
  }
  __syncthreads();
  const int r = threadIdx.x / T::kCols, c = threadIdx.x - r * T::kCols;
  const int i = i0 + r, k = k0 + c;
  if (i >= g.nr - 1 || k >= g.nz - 1) return;
  const double dr = g.dr, dz = g.dz;
  const int q = (r + 1) * T::kStride + c + 1;  // node (i, k); q-1 = (i, k-1), q-kStride = (i-1, k)
  const size_t p = at(g, i, k);
  const double ke = f.ke[p], kh = f.kh[p];
  const double hphi = s_hphi[q];
  if (i == 0) {
    if (k >= 1) f.er[p] = f.er[p] * ke - (f.jr[p] + (hphi - s_hphi[q - 1]) / dz) * kh;
    f.ez[p] = f.ez[p] * ke - (f.jz[p] - 4.0 / dr * hphi) * kh;
    return;
  }
  const double hphi_d = s_hphi[q - T::kStride];
  if (k >= 1) {
    f.er[p] = f.er[p] * ke - (f.jr[p] + (hphi - s_hphi[q - 1]) / dz) * kh;
    f.ephi[p] = f.ephi[p] * ke -
                (f.jphi[p] - (s_hr[q] - s_hr[q - 1]) / dz + (s_hz[q] - s_hz[q - T::kStride]) / dr) * kh;
  }
  f.ez[p] = f.ez[p] * ke - (f.jz[p] - (hphi - hphi_d) / dr - (hphi + hphi_d) / (2.0 * dr * i)) * kh;
}

__global__ void __launch_bounds__(ETile::kThreads) k_e_update(GridDesc g, EPtrs f) {
  __shared__ double s[3 * ETile::kLds];
  e_tile<ETile>(g, f, blockIdx.y * ETile::kRows, blockIdx.x * ETile::kCols, s);
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
// a work item is one tile (HTile or ETile) of one instance;
// blocks walk the item list with a stride of the grid size
struct EnsRow {
  int inst, i0, k0;
};
struct EnsGrid {
  GridDesc g;
  HPtrs h;
  EPtrs e;
  double *jr, *jphi, *jz, *rho;
};

// the grid tables are read through mirrors whose pointers are in the global
// address space (pointers read from device memory are otherwise generic, and
// every access would be a flat instruction on a 64-bit per-lane address)
template <class T>
using gptr = __attribute__((address_space(1))) T *;
struct HPtrsG {
  gptr<const double> er, ephi, ez;
  gptr<double> hr, hphi, hz, hhr, hhphi, hhz;
};
struct EPtrsG {
  gptr<const double> hr, hphi, hz, jr, jphi, jz, ke, kh;
  gptr<double> er, ephi, ez;
};
struct EnsGridG {
  GridDesc g;
  HPtrsG h;
  EPtrsG e;
  gptr<double> jr, jphi, jz, rho;
};
static_assert(sizeof(EnsGridG) == sizeof(EnsGrid) && offsetof(EnsGridG, h) == offsetof(EnsGrid, h) &&
                  offsetof(EnsGridG, e) == offsetof(EnsGrid, e) && offsetof(EnsGridG, jr) == offsetof(EnsGrid, jr) &&
                  sizeof(HPtrsG) == sizeof(HPtrs) && sizeof(EPtrsG) == sizeof(EPtrs),
              "EnsGridG mirrors EnsGrid");
__device__ __forceinline__ HPtrs from_global(const HPtrsG &g) {
  return HPtrs{(const double *)g.er,  (const double *)g.ephi, (const double *)g.ez,
               (double *)g.hr,        (double *)g.hphi,       (double *)g.hz,
               (double *)g.hhr,       (double *)g.hhphi,      (double *)g.hhz};
}
__device__ __forceinline__ EPtrs from_global(const EPtrsG &g) {
  return EPtrs{(const double *)g.hr, (const double *)g.hphi, (const double *)g.hz, (const double *)g.jr,
               (const double *)g.jphi, (const double *)g.jz, (const double *)g.ke, (const double *)g.kh,
               (double *)g.er, (double *)g.ephi, (double *)g.ez};
}

__global__ void __launch_bounds__(HTile::kThreads) k_ens_h(const EnsGrid *__restrict__ grids,
                                                           const EnsRow *__restrict__ tiles, int n_tiles) {
  __shared__ double s[3 * HTile::kLds];
  for (int it = blockIdx.x; it < n_tiles; it += gridDim.x) {
    const EnsRow r = tiles[it];
    const EnsGridG &G = reinterpret_cast<const EnsGridG *>(grids)[r.inst];
    h_tile<HTile>(G.g, from_global(G.h), r.i0, r.k0, s);
    __syncthreads();  // the next tile rewrites the staged rows
  }
}

__global__ void __launch_bounds__(ETile::kThreads) k_ens_e(const EnsGrid *__restrict__ grids,
                                                           const EnsRow *__restrict__ tiles, int n_tiles) {
  __shared__ double s[3 * ETile::kLds];
  for (int it = blockIdx.x; it < n_tiles; it += gridDim.x) {
    const EnsRow r = tiles[it];
    const EnsGridG &G = reinterpret_cast<const EnsGridG *>(grids)[r.inst];
    e_tile<ETile>(G.g, from_global(G.e), r.i0, r.k0, s);
    __syncthreads();  // the next tile rewrites the staged rows
  }
}

// Current::reset_j and the bunch density reset of every instance
__global__ void __launch_bounds__(256) k_ens_zero(const EnsGrid *__restrict__ grids) {
  const EnsGridG &G = reinterpret_cast<const EnsGridG *>(grids)[blockIdx.y];
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
  dim3 grid((c.g.nz - 1 + HTile::kCols - 1) / HTile::kCols, (c.g.nr + HTile::kRows - 1) / HTile::kRows);
  hipLaunchKernelGGL(k_h_update, grid, dim3(HTile::kThreads), 0, c.stream, c.g, hptrs(c));
}

void launch_e_update(Ctx &c) {
  dim3 grid((c.g.nz - 1 + ETile::kCols - 1) / ETile::kCols, (c.g.nr - 1 + ETile::kRows - 1) / ETile::kRows);
  hipLaunchKernelGGL(k_e_update, grid, dim3(ETile::kThreads), 0, c.stream, c.g, eptrs(c));
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
    for (int i0 = 0; i0 < c.g.nr; i0 += HTile::kRows)
      for (int k0 = 0; k0 < c.g.nz - 1; k0 += HTile::kCols) hr.push_back(EnsRow{(int)j, i0, k0});
    for (int i0 = 0; i0 < c.g.nr - 1; i0 += ETile::kRows)
      for (int k0 = 0; k0 < c.g.nz - 1; k0 += ETile::kCols) er.push_back(EnsRow{(int)j, i0, k0});
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
  hipLaunchKernelGGL(k_ens_h, dim3(blocks), dim3(HTile::kThreads), 0, st, e->grids, e->hrows, e->n_h);
}

void ens_launch_e(EnsFields *e, hipStream_t st) {
  const int blocks = std::min(e->n_e, 256 * 16);
  hipLaunchKernelGGL(k_ens_e, dim3(blocks), dim3(ETile::kThreads), 0, st, e->grids, e->erows, e->n_e);
}

void ens_launch_zero(EnsFields *e, hipStream_t st) {
  hipLaunchKernelGGL(k_ens_zero, dim3(64, e->n_inst), dim3(256), 0, st, e->grids);
}

}  // namespace pdp3
