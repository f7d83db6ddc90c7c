// Host side of the C ABI (include/pdp3_hip.h): context and species
// management, the step loop of LoadInitParam::run, bunch bookkeeping and live
// kernel timing.  All device work is enqueued on the context's stream.
#include "../../include/pdp3_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <functional>
#include <thread>
#include <vector>

#include "context.hpp"

struct pdp3_ctx {
  pdp3::Ctx c;
};

using namespace pdp3;

// Device memory of pdp3_gauss_residual_noalloc: the charge density of all
// species (first PDP3_NOALLOC_MAX_NODES values), then the residual.  Every GPU
// the library runs on holds its own copy (2 x 36 MiB) until the process ends.
// In host code `scratch` is only a stand-in for the symbol: its device address
// on a GPU comes from hipGetSymbolAddress (noalloc_address below).
static __device__ __align__(256) double scratch[2 * (size_t)PDP3_NOALLOC_MAX_NODES];

namespace {

// per-device address of `scratch`, resolved once per device with that device
// current (at the first context created on it)
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[kMaxDevices] = {};
bool g_noalloc_tried[kMaxDevices] = {};

double *noalloc_address(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_noalloc_tried[device]) {
    g_noalloc_tried[device] = true;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) g_noalloc_ptr[device] = (double *)p;
    else (void)hipGetLastError();  // leave no sticky error for the caller's next call
  }
  return g_noalloc_ptr[device];
}

int fail(pdp3_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->c.last_error = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                         \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      return fail(ctx, e_ == hipErrorOutOfMemory ? PDP3_ERR_OOM : PDP3_ERR_HIP,                    \
                  std::string(#expr) + ": " + hipGetErrorString(e_));                              \
  } while (0)

#define CHECK_LAUNCH(ctx) HIP_TRY(ctx, hipGetLastError())

bool valid_species(pdp3_ctx *ctx, int idx) {
  return ctx && idx >= 0 && idx < (int)ctx->c.species.size();
}

// koef_e / koef_h of EField::calc_field (src/eField.cpp:105-110), per cell
int upload_koef(pdp3_ctx *ctx) {
  Ctx &c = ctx->c;
  const size_t n = c.cells();
  std::vector<double> ke(n), kh(n);
  for (size_t p = 0; p < n; p++) {
    double epsilonx2 = 2 * c.eps_host[p] * kEpsilon0;
    double sigma_t = c.sigma_host[p] * c.g.dt;
    ke[p] = (epsilonx2 - sigma_t) / (epsilonx2 + sigma_t);
    kh[p] = 2 * c.g.dt / (epsilonx2 + sigma_t);
  }
  HIP_TRY(ctx, hipMemcpyAsync(c.koef_e, ke.data(), n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipMemcpyAsync(c.koef_h, kh.data(), n * sizeof(double), hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipMemcpyAsync(c.a[kSigma], c.sigma_host.data(), n * sizeof(double), hipMemcpyHostToDevice,
                              c.stream));
  HIP_TRY(ctx, hipMemcpyAsync(c.a[kEps], c.eps_host.data(), n * sizeof(double), hipMemcpyHostToDevice,
                              c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

template <typename T>
int dalloc(pdp3_ctx *ctx, T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  HIP_TRY(ctx, hipMalloc((void **)p, n * sizeof(T)));
  return PDP3_OK;
}

void dfree(void *p) {
  if (p) (void)hipFree(p);
}

void free_species(Species &sp) {
  for (auto *s : {&sp.f64, &sp.f64_alt}) {
    dfree(s->r); dfree(s->z); dfree(s->vr); dfree(s->vphi); dfree(s->vz); dfree(s->w);
    dfree(s->alive); dfree(s->r_old); dfree(s->z_old); dfree(s->sin_t);
    *s = SpeciesF64{};
  }
  for (auto *s : {&sp.f32, &sp.f32_alt}) {
    dfree(s->cell); dfree(s->fr); dfree(s->fz); dfree(s->vr); dfree(s->vphi); dfree(s->vz); dfree(s->w);
    *s = SpeciesF32{};
  }
  dfree(sp.inject_u);
  sp.inject_u = nullptr;
  if (sp.runs_ev) (void)hipEventDestroy(sp.runs_ev);
  if (sp.runs_host) (void)hipHostFree(sp.runs_host);
  sp.runs_ev = nullptr;
  sp.runs_host = nullptr;
}

bool is_f32(const Ctx &c) { return c.precision == PDP3_PREC_FP32; }

// FP64 arrays: the live state in the species' layout (n_cap slots) in the
// FP64 form, a logical-order staging copy (n entries) in the FP32 form
int alloc_species_state(pdp3_ctx *ctx, Species &sp) {
  Ctx &c = ctx->c;
  sp.n_cap = sp.chunked ? ((long long)sp.n + kChunk - 1) / kChunk * kChunk : sp.n;
  const size_t n = is_f32(c) ? (size_t)sp.n : (size_t)sp.n_cap;
  int st;
  if ((st = dalloc(ctx, &sp.f64.r, n)) || (st = dalloc(ctx, &sp.f64.z, n)) || (st = dalloc(ctx, &sp.f64.vr, n)) ||
      (st = dalloc(ctx, &sp.f64.vphi, n)) || (st = dalloc(ctx, &sp.f64.vz, n)) ||
      (st = dalloc(ctx, &sp.f64.w, n)) || (st = dalloc(ctx, &sp.f64.alive, n)))
    return st;
  HIP_TRY(ctx, hipMemsetAsync(sp.f64.alive, 0, n, c.stream));
  for (double *p : {sp.f64.r, sp.f64.z, sp.f64.vr, sp.f64.vphi, sp.f64.vz, sp.f64.w})
    HIP_TRY(ctx, hipMemsetAsync(p, 0, n * sizeof(double), c.stream));
  if (!is_f32(c) && sp.chunked) {
    launch_mark_padding(c, sp.f64.alive, sp.n, sp.n_cap);
    CHECK_LAUNCH(ctx);
  }
  return PDP3_OK;
}

int ensure_split_scratch(pdp3_ctx *ctx, Species &sp) {
  if (sp.f64.r_old) return PDP3_OK;
  int st;
  if ((st = dalloc(ctx, &sp.f64.r_old, (size_t)sp.n_cap)) || (st = dalloc(ctx, &sp.f64.z_old, (size_t)sp.n_cap)) ||
      (st = dalloc(ctx, &sp.f64.sin_t, (size_t)sp.n_cap)))
    return st;
  return PDP3_OK;
}

// ---- live timing with HIP events on the context's stream ----
struct TimedScope {
  Ctx &c;
  int kernel;
  hipEvent_t a = nullptr, b = nullptr;
  TimedScope(Ctx &cc, int k) : c(cc), kernel(k) {
    if (!c.profiling) return;
    KernelTimer &t = c.timers[kernel];
    if (!t.pool.empty()) {
      a = t.pool.back().first;
      b = t.pool.back().second;
      t.pool.pop_back();
    } else {
      (void)hipEventCreate(&a);
      (void)hipEventCreate(&b);
    }
    (void)hipEventRecord(a, c.stream);
  }
  ~TimedScope() {
    if (!c.profiling || !a) return;
    (void)hipEventRecord(b, c.stream);
    c.timers[kernel].pending.emplace_back(a, b);
  }
};

void harvest_timers(Ctx &c) {
  for (auto &t : c.timers) {
    for (auto &ev : t.pending) {
      float ms = 0;
      if (hipEventElapsedTime(&ms, ev.first, ev.second) == hipSuccess) {
        t.total_ms += ms;
        t.launches++;
      }
      t.pool.push_back(ev);
    }
    t.pending.clear();
  }
}

int check_device_errors(pdp3_ctx *ctx) {
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipMemcpyAsync(c.err_host, c.err, sizeof(int), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  harvest_timers(c);
  int e = *c.err_host;
  if (e) {
    HIP_TRY(ctx, hipMemsetAsync(c.err, 0, sizeof(int), c.stream));
    if (e & kErrNonfinite) return fail(ctx, PDP3_ERR_NONFINITE, "particle position is not a finite number");
    if (e & kErrSuperluminal) return fail(ctx, PDP3_ERR_SUPERLUMINAL, "particle speed reached c (gamma complex)");
    if (e & kErrDomain) return fail(ctx, PDP3_ERR_BAD_ARG, "particle outside the grid (deposit or gather dropped)");
  }
  return PDP3_OK;
}

int inject_bunch(pdp3_ctx *ctx, Species &sp) {
  // Bunch::bunch_inject, src/bunch.cpp:87-124 (the index arithmetic runs here
  // on the host in double exactly as in the reference)
  Ctx &c = ctx->c;
  double bunch_time_begin = sp.duration * sp.bunch_number + sp.hole_duration * sp.bunch_number;
  double bunch_time_end = bunch_time_begin + sp.duration;
  double dl = sp.velocity * c.g.dt;
  int steps_amount = (int)std::ceil(sp.duration / c.g.dt);
  if (steps_amount <= 0) return PDP3_OK;
  unsigned particles_in_step = (unsigned)sp.n / (unsigned)steps_amount;
  int start_number = (int)((c.current_time - bunch_time_begin) / c.g.dt * particles_in_step);
  if (c.current_time >= bunch_time_begin && c.current_time < bunch_time_end) {
    if (is_f32(c))
      launch_bunch_inject_f32(c, sp, (unsigned)start_number, particles_in_step, dl);
    else
      launch_bunch_inject_f64(c, sp, (unsigned)start_number, particles_in_step, dl);
    CHECK_LAUNCH(ctx);
  }
  return PDP3_OK;
}

// LoadInitParam::init_beam, src/loadInitParam.cpp:168-248 (inject false: the
// bookkeeping only, for ensembles, whose injections run on the device)
int init_beam(pdp3_ctx *ctx, bool inject = true) {
  Ctx &c = ctx->c;
  int cand = -1;
  for (int j = 0; j < (int)c.species.size(); j++) {
    Species &sp = c.species[j];
    if (sp.kind == PDP3_SPECIES_BUNCH && !sp.created && sp.bunch_number == c.next_bunch) cand = j;
  }
  if (cand >= 0) {
    Species &sp = c.species[cand];
    double bunch_inject_time = c.next_bunch * (sp.duration + sp.hole_duration);
    if (c.current_time >= bunch_inject_time) {
      sp.created = true;
      c.order.push_back(cand);
      c.next_bunch++;
    }
  }
  if (!inject) return PDP3_OK;
  for (int j : c.order)
    if (c.species[j].kind == PDP3_SPECIES_BUNCH) {
      int st = inject_bunch(ctx, c.species[j]);
      if (st) return st;
    }
  return PDP3_OK;
}

int zero_grid(pdp3_ctx *ctx, int comp, int count = 1) {
  Ctx &c = ctx->c;
  launch_fill(c.a[comp], c.cells() * (size_t)count, 0.0, c.stream);  // the components are contiguous
  CHECK_LAUNCH(ctx);
  return PDP3_OK;
}

int particles_fused(pdp3_ctx *ctx, Species &sp) {
  Ctx &c = ctx->c;
  TimedScope ts(c, sp.kind == PDP3_SPECIES_BUNCH ? PDP3_KERNEL_BUNCH : PDP3_KERNEL_PARTICLES);
  if (is_f32(c))
    launch_fused_f32(c, sp);
  else
    launch_fused_f64(c, sp);
  CHECK_LAUNCH(ctx);
  return PDP3_OK;
}

}  // namespace

extern "C" {

const char *pdp3_version(void) { return "pdp3-mi355x 0.1 (gfx950)"; }

int pdp3_ctx_create(int n_grid_r, int n_grid_z, double r_size, double z_size, double current_time,
                    double start_time, double relaxation_time, double end_time, double delta_t, int pusher,
                    int precision, pdp3_ctx **out) {
  if (!out) return PDP3_ERR_BAD_ARG;
  *out = nullptr;
  if (n_grid_r < 4 || n_grid_z < 4 || !(r_size > 0) || !(z_size > 0) || !(delta_t > 0)) return PDP3_ERR_BAD_ARG;
  if (pusher != PDP3_PUSHER_BORIS && pusher != PDP3_PUSHER_VAY) return PDP3_ERR_BAD_ARG;
  if (precision != PDP3_PREC_FP64 && precision != PDP3_PREC_FP32) return PDP3_ERR_BAD_ARG;
  if ((long long)n_grid_r * n_grid_z >= (1LL << 31)) return PDP3_ERR_BAD_ARG;
  // the FP32 form packs the cell as 15 + 16 bits
  if (precision == PDP3_PREC_FP32 && (n_grid_r > 32768 || n_grid_z > 65536)) return PDP3_ERR_BAD_ARG;
  pdp3_ctx *ctx = new (std::nothrow) pdp3_ctx();
  if (!ctx) return PDP3_ERR_OOM;
  Ctx &c = ctx->c;
  c.g.nr = n_grid_r;
  c.g.nz = n_grid_z;
  c.g.r_size = r_size;
  c.g.z_size = z_size;
  c.g.dr = r_size / (n_grid_r - 1);  // Geometry::get_dr, src/geometry.cpp:45-48
  c.g.dz = z_size / (n_grid_z - 1);
  c.g.dt = delta_t;
  c.current_time = current_time;
  c.start_time = start_time;
  c.relaxation_time = relaxation_time;
  c.end_time = end_time;
  c.pusher = pusher;
  c.precision = precision;
  int st = PDP3_OK;
  auto bail = [&](int code) {
    pdp3_ctx_destroy(ctx);
    return code;
  };
  if (hipGetDevice(&c.device) != hipSuccess) return bail(PDP3_ERR_HIP);
  c.noalloc = noalloc_address(c.device);  // checked by pdp3_gauss_residual_noalloc
  if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) return bail(PDP3_ERR_HIP);
  const size_t n = c.cells();
  if ((st = dalloc(ctx, &c.grid, n * kNComp))) return bail(st);
  for (int k = 0; k < kNComp; k++) c.a[k] = c.grid + (size_t)k * n;
  if ((st = dalloc(ctx, &c.koef_e, n)) || (st = dalloc(ctx, &c.koef_h, n)) || (st = dalloc(ctx, &c.err, 1)))
    return bail(st);
  if (hipHostMalloc((void **)&c.err_host, sizeof(int), hipHostMallocDefault) != hipSuccess)
    return bail(PDP3_ERR_HIP);
  *c.err_host = 0;
  if (hipMemsetAsync(c.grid, 0, n * kNComp * sizeof(double), c.stream) != hipSuccess) return bail(PDP3_ERR_HIP);
  if (hipMemsetAsync(c.err, 0, sizeof(int), c.stream) != hipSuccess) return bail(PDP3_ERR_HIP);
  c.sigma_host.assign(n, 0.0);
  c.eps_host.assign(n, 1.0);  // Geometry::set_epsilon, src/geometry.cpp:33-39
  if ((st = upload_koef(ctx))) return bail(st);
  {
    std::vector<float4> rows(2 * (size_t)n_grid_r);
    for (int i = 0; i < n_grid_r; i++) {
      const double fi = i;
      const float r0 = i ? (float)(1.0 / (2.0 * fi)) : 0.f, i0 = i ? (float)(1.0 / fi) : 0.f;
      rows[2 * i] = make_float4(r0, (float)(1.0 / (2.0 * (fi + 1))), (float)(1.0 / (2.0 * (fi + 2))),
                                (float)(1.0 / (2.0 * (fi + 3))));
      rows[2 * i + 1] = make_float4(i0, (float)(1.0 / (fi + 1)), (float)(1.0 / (fi + 0.5)), (float)(fi * (fi + 1)));
    }
    if ((st = dalloc(ctx, &c.rows, rows.size()))) return bail(st);
    if (hipMemcpy(c.rows, rows.data(), rows.size() * sizeof(float4), hipMemcpyHostToDevice) != hipSuccess)
      return bail(PDP3_ERR_HIP);
  }
  {
    // the FP64 form's divisors: CELL_VOLUME and CYL_VOL in the reference's
    // expression order (include/geometry.h:33-39), with IEEE reciprocals
    const double dr = c.g.dr, dz = c.g.dz;
    std::vector<double2> vol((size_t)n_grid_r + 2);
    for (int i = 0; i < n_grid_r + 2; i++) {
      const double v = kPi * dz * dr * dr * 2.0 * i;
      vol[i] = make_double2(v, i ? 1.0 / v : 0.0);
    }
    c.g.inv_dr = 1.0 / dr;
    c.g.inv_dz = 1.0 / dz;
    c.g.cyl = kPi * dz * dr * dr / 4.;
    c.g.inv_cyl = 1.0 / c.g.cyl;
    if ((st = dalloc(ctx, &c.vol64, vol.size()))) return bail(st);
    if (hipMemcpy(c.vol64, vol.data(), vol.size() * sizeof(double2), hipMemcpyHostToDevice) != hipSuccess)
      return bail(PDP3_ERR_HIP);
    c.g.vol = c.vol64;
  }
  *out = ctx;
  return PDP3_OK;
}

int pdp3_ctx_destroy(pdp3_ctx *ctx) {
  if (!ctx) return PDP3_OK;
  Ctx &c = ctx->c;
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  for (auto &sp : c.species) free_species(sp);
  dfree(c.grid);
  dfree(c.koef_e);
  dfree(c.koef_h);
  dfree(c.err);
  dfree(c.sort_keys);
  dfree(c.sort_perm);
  dfree(c.sort_ws);
  dfree(c.rho_all);
  dfree(c.gauss);
  dfree(c.probe_ws);
  dfree(c.rows);
  dfree(c.vol64);
  if (c.err_host) (void)hipHostFree(c.err_host);
  for (auto &t : c.timers) {
    for (auto &ev : t.pending) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
    for (auto &ev : t.pool) {
      (void)hipEventDestroy(ev.first);
      (void)hipEventDestroy(ev.second);
    }
  }
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete ctx;
  return PDP3_OK;
}

const char *pdp3_last_error(const pdp3_ctx *ctx) { return ctx ? ctx->c.last_error.c_str() : "null context"; }

int pdp3_geometry_get(const pdp3_ctx *ctx, double out[6]) {
  if (!ctx || !out) return PDP3_ERR_BAD_ARG;
  const Ctx &c = ctx->c;
  out[0] = c.g.dr;
  out[1] = c.g.dz;
  out[2] = c.g.r_size;
  out[3] = c.g.z_size;
  out[4] = c.g.nr;
  out[5] = c.g.nz;
  return PDP3_OK;
}

int pdp3_time_get(const pdp3_ctx *ctx, double *current_time, double *delta_t) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  if (current_time) *current_time = ctx->c.current_time;
  if (delta_t) *delta_t = ctx->c.g.dt;
  return PDP3_OK;
}

int pdp3_time_set(pdp3_ctx *ctx, double current_time) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  ctx->c.current_time = current_time;
  return PDP3_OK;
}

// Geometry::set_pml, src/geometry.cpp:56-107 (host, once)
int pdp3_geometry_set_pml(pdp3_ctx *ctx, double l1, double l2, double l3, double sigma1, double sigma2) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  const double dr = c.g.dr, dz = c.g.dz, r_size = c.g.r_size, z_size = c.g.z_size;
  const int nr = c.g.nr, nz = c.g.nz;
  auto S = [&](int i, int k) -> double & { return c.sigma_host[(size_t)i * nz + k]; };
  double len_left = dz * (std::floor(nz * l1));
  double len_right = dz * (std::floor(nz * l2));
  double len_ext = dr * (std::floor(nr * l3));
  if ((l1 == 0) && (l2 == 0) && (l3 != 0)) {
    for (int i = 0; i < nr; i++)
      for (int k = 0; k < nz; k++)
        if (r_size - dr * i <= len_ext)
          S(i, k) = sigma1 + (sigma2 - sigma1) / std::pow(len_ext, 2) * std::pow((dr * (i + 1) - r_size + len_ext), 2);
  } else {
    for (int k = 0; k < nz; k++)
      for (int i = 0; i < nr; i++) {
        if (dz * k < len_left)
          S(i, k) = sigma1 + (sigma2 - sigma1) / std::pow(len_left, 2) * std::pow(len_left - dz * k, 2);
        if (z_size - dz * k < len_right)
          S(i, k) = sigma1 + (sigma2 - sigma1) / std::pow(len_right, 2) *
                                 std::pow(dz * (k + 1) - z_size + len_right, 2);
      }
  }
  if (l3 != 0) {
    for (int i = 0; i < nr; i++)
      for (int k = 0; k < nz; k++)
        if (r_size - dr * i <= len_ext)
          if ((r_size - dr * (i + 1) < len_ext * dz * k / len_left) &&
              (r_size - dr * i <= len_ext / len_right * (z_size - dz * k)))
            S(i, k) = sigma1 + (sigma2 - sigma1) / std::pow(len_ext, 2) * std::pow(dr * (i + 1) - r_size + len_ext, 2);
  }
  return upload_koef(ctx);
}

int pdp3_efield_boundary_conditions(pdp3_ctx *ctx) {
  // EField::boundary_conditions, src/eField.cpp:46-82: the tangential wall
  // values are set once; no update ever writes them afterwards
  if (!ctx) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  const int nr = c.g.nr, nz = c.g.nz;
  std::vector<double> er(c.cells()), ephi(c.cells()), ez(c.cells());
  HIP_TRY(ctx, hipMemcpyAsync(er.data(), c.a[kER], c.cells() * 8, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipMemcpyAsync(ephi.data(), c.a[kEPhi], c.cells() * 8, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipMemcpyAsync(ez.data(), c.a[kEZ], c.cells() * 8, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  for (int i = 0; i < nr - 1; i++) er[(size_t)i * nz] = er[(size_t)i * nz + nz - 1] = 0;
  for (int k = 0; k < nz; k++) ephi[k] = ephi[(size_t)(nr - 1) * nz + k] = 0;
  for (int i = 0; i < nr; i++) ephi[(size_t)i * nz] = ephi[(size_t)i * nz + nz - 1] = 0;
  for (int k = 0; k < nz - 1; k++) ez[k] = ez[(size_t)(nr - 1) * nz + k] = 0;
  HIP_TRY(ctx, hipMemcpyAsync(c.a[kER], er.data(), c.cells() * 8, hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipMemcpyAsync(c.a[kEPhi], ephi.data(), c.cells() * 8, hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipMemcpyAsync(c.a[kEZ], ez.data(), c.cells() * 8, hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

namespace {
// sets a[i][k] = v for i < ni, k < nk (a rectangle anchored at (0,0))
int set_block(pdp3_ctx *ctx, int comp, int ni, int nk, double v) {
  Ctx &c = ctx->c;
  std::vector<double> h(c.cells());
  HIP_TRY(ctx, hipMemcpyAsync(h.data(), c.a[comp], c.cells() * 8, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  for (int i = 0; i < ni; i++)
    for (int k = 0; k < nk; k++) h[(size_t)i * c.g.nz + k] = v;
  HIP_TRY(ctx, hipMemcpyAsync(c.a[comp], h.data(), c.cells() * 8, hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}
}  // namespace

int pdp3_efield_set_homogeneous(pdp3_ctx *ctx, double e_r, double e_phi, double e_z) {
  // EField::set_homogeneous_efield, src/eField.cpp:33-44
  if (!ctx) return PDP3_ERR_BAD_ARG;
  const int nr = ctx->c.g.nr, nz = ctx->c.g.nz;
  int st;
  if ((st = set_block(ctx, kER, nr - 1, nz - 1, e_r)) || (st = set_block(ctx, kEPhi, nr - 1, nz - 1, e_phi)) ||
      (st = set_block(ctx, kEZ, nr - 1, nz - 1, e_z)))
    return st;
  return PDP3_OK;
}

int pdp3_hfield_set_homogeneous(pdp3_ctx *ctx, double h_r, double h_phi, double h_z) {
  // HField::set_homogeneous_h, src/hField.cpp:66-91
  if (!ctx) return PDP3_ERR_BAD_ARG;
  const int nr = ctx->c.g.nr, nz = ctx->c.g.nz;
  int st;
  if ((st = set_block(ctx, kHR, nr, nz - 1, h_r)) || (st = set_block(ctx, kHHR, nr, nz - 1, h_r)) ||
      (st = set_block(ctx, kHPhi, nr - 1, nz - 1, h_phi)) || (st = set_block(ctx, kHHPhi, nr - 1, nz - 1, h_phi)) ||
      (st = set_block(ctx, kHZ, nr - 1, nz, h_z)) || (st = set_block(ctx, kHHZ, nr - 1, nz, h_z)))
    return st;
  return PDP3_OK;
}

int pdp3_boundary_specify_initial_field(pdp3_ctx *ctx, double e_phi_upper, double e_phi_left,
                                        double e_phi_right) {
  // BoundaryMaxwellConditions::specify_initial_field,
  // src/boundaryMaxwellConditions.cpp:17-38
  if (!ctx) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  const int nr = c.g.nr, nz = c.g.nz;
  std::vector<double> h(c.cells());
  HIP_TRY(ctx, hipMemcpyAsync(h.data(), c.a[kEPhi], c.cells() * 8, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  for (int i = 0; i < nr; i++) {
    h[(size_t)i * nz] = e_phi_left;
    h[(size_t)i * nz + nz - 1] = e_phi_right;
  }
  for (int k = 0; k < nz; k++) h[(size_t)(nr - 1) * nz + k] = e_phi_upper;
  HIP_TRY(ctx, hipMemcpyAsync(c.a[kEPhi], h.data(), c.cells() * 8, hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

int pdp3_hfield_calc(pdp3_ctx *ctx) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  TimedScope ts(ctx->c, PDP3_KERNEL_H);
  launch_h_update(ctx->c);
  CHECK_LAUNCH(ctx);
  return PDP3_OK;
}

int pdp3_efield_calc(pdp3_ctx *ctx) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  TimedScope ts(ctx->c, PDP3_KERNEL_E);
  launch_e_update(ctx->c);
  CHECK_LAUNCH(ctx);
  return PDP3_OK;
}

int pdp3_current_reset(pdp3_ctx *ctx) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  int st;
  if ((st = zero_grid(ctx, kJR)) || (st = zero_grid(ctx, kJPhi)) || (st = zero_grid(ctx, kJZ))) return st;
  return PDP3_OK;
}

int pdp3_rho_bunch_reset(pdp3_ctx *ctx) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  return zero_grid(ctx, kRhoBunch);
}

int pdp3_efield_get_field(pdp3_ctx *ctx, const double *r, const double *z, int n, double *out) {
  if (!ctx || n < 0 || (n > 0 && (!r || !z || !out))) return PDP3_ERR_BAD_ARG;
  if (n == 0) return PDP3_OK;
  Ctx &c = ctx->c;
  double *d = nullptr;
  HIP_TRY(ctx, hipMalloc(&d, sizeof(double) * 5 * (size_t)n));
  (void)hipMemcpyAsync(d, r, sizeof(double) * n, hipMemcpyHostToDevice, c.stream);
  (void)hipMemcpyAsync(d + n, z, sizeof(double) * n, hipMemcpyHostToDevice, c.stream);
  launch_gather_f64(c, 0, d, d + n, n, d + 2 * n);
  (void)hipMemcpyAsync(out, d + 2 * n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c.stream);
  hipError_t e = hipStreamSynchronize(c.stream);
  dfree(d);
  if (e != hipSuccess) return fail(ctx, PDP3_ERR_HIP, hipGetErrorString(e));
  return check_device_errors(ctx);
}

int pdp3_hfield_get_field(pdp3_ctx *ctx, const double *r, const double *z, int n, double *out) {
  if (!ctx || n < 0 || (n > 0 && (!r || !z || !out))) return PDP3_ERR_BAD_ARG;
  if (n == 0) return PDP3_OK;
  Ctx &c = ctx->c;
  double *d = nullptr;
  HIP_TRY(ctx, hipMalloc(&d, sizeof(double) * 5 * (size_t)n));
  (void)hipMemcpyAsync(d, r, sizeof(double) * n, hipMemcpyHostToDevice, c.stream);
  (void)hipMemcpyAsync(d + n, z, sizeof(double) * n, hipMemcpyHostToDevice, c.stream);
  launch_gather_f64(c, 1, d, d + n, n, d + 2 * n);
  (void)hipMemcpyAsync(out, d + 2 * n, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, c.stream);
  hipError_t e = hipStreamSynchronize(c.stream);
  dfree(d);
  if (e != hipSuccess) return fail(ctx, PDP3_ERR_HIP, hipGetErrorString(e));
  return check_device_errors(ctx);
}

int pdp3_grid_download(pdp3_ctx *ctx, int component, double *out) {
  if (!ctx || component < 0 || component >= kNComp || !out) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipMemcpyAsync(out, c.a[component], c.cells() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

int pdp3_grid_upload(pdp3_ctx *ctx, int component, const double *in) {
  if (!ctx || component < 0 || component >= kNComp || !in) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  if (component == kSigma || component == kEps) {
    auto &dst = component == kSigma ? c.sigma_host : c.eps_host;
    std::memcpy(dst.data(), in, c.cells() * sizeof(double));
    return upload_koef(ctx);
  }
  HIP_TRY(ctx, hipMemcpyAsync(c.a[component], in, c.cells() * sizeof(double), hipMemcpyHostToDevice, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

int pdp3_species_add(pdp3_ctx *ctx, int charge_e, double mass_e, int number, int *out_index) {
  if (!ctx || number < 0 || !(mass_e > 0)) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  Species sp;
  sp.kind = PDP3_SPECIES_PLASMA;
  sp.charge_e = charge_e;
  sp.mass_e = mass_e;
  sp.charge = charge_e * kElCharge;  // src/particles.cpp:30-33
  sp.mass = mass_e * kElMass;
  sp.n = number;
  sp.created = true;
  sp.chunked = true;
  int st = alloc_species_state(ctx, sp);
  if (st) {
    free_species(sp);
    return st;
  }
  c.species.push_back(sp);
  int idx = (int)c.species.size() - 1;
  c.order.push_back(idx);
  if (out_index) *out_index = idx;
  return PDP3_OK;
}

int pdp3_bunch_add(pdp3_ctx *ctx, int charge_e, double mass_e, int number, const double *r, const double *w,
                   const double *inject_u, double duration, double hole_duration, double velocity,
                   int bunch_number, int *out_index) {
  if (!ctx || number < 0 || !(mass_e > 0) || (number > 0 && (!r || !w || !inject_u)) || !(duration > 0) ||
      bunch_number < 0)
    return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  Species sp;
  sp.kind = PDP3_SPECIES_BUNCH;
  sp.charge_e = charge_e;
  sp.mass_e = mass_e;
  sp.charge = charge_e * kElCharge;
  sp.mass = mass_e * kElMass;
  sp.n = number;
  sp.duration = duration;
  sp.hole_duration = hole_duration;
  sp.velocity = velocity;
  sp.bunch_number = bunch_number;
  sp.created = false;
  sp.chunked = false;
  int st = alloc_species_state(ctx, sp);
  if (!st) st = dalloc(ctx, &sp.inject_u, (size_t)number);
  if (st) {
    free_species(sp);
    return st;
  }
  const size_t b = sizeof(double) * (size_t)number;
  if (number > 0) {
    HIP_TRY(ctx, hipMemcpyAsync(sp.f64.r, r, b, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(ctx, hipMemcpyAsync(sp.f64.w, w, b, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(ctx, hipMemcpyAsync(sp.inject_u, inject_u, b, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  }
  c.species.push_back(sp);
  int idx = (int)c.species.size() - 1;
  if (is_f32(c)) {
    if (!launch_f32_from_f64(c, c.species[idx])) return fail(ctx, PDP3_ERR_OOM, "FP32 species arrays");
    CHECK_LAUNCH(ctx);
  }
  if (out_index) *out_index = idx;
  return PDP3_OK;
}

int pdp3_species_count(const pdp3_ctx *ctx, int index, int *number) {
  if (!ctx || index < 0 || index >= (int)ctx->c.species.size() || !number) return PDP3_ERR_BAD_ARG;
  *number = ctx->c.species[index].n;
  return PDP3_OK;
}

int pdp3_bunch_add_loaded(pdp3_ctx *ctx, int charge_e, double mass_e, int number, const double *r,
                          const double *z, const double *vr, const double *vphi, const double *vz, const double *w,
                          int *out_index) {
  if (!ctx || number < 0 || !(mass_e > 0) || (number > 0 && (!r || !z || !vr || !vphi || !vz || !w)))
    return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  Species sp;
  sp.kind = PDP3_SPECIES_BUNCH;
  sp.charge_e = charge_e;
  sp.mass_e = mass_e;
  sp.charge = charge_e * kElCharge;
  sp.mass = mass_e * kElMass;
  sp.n = number;
  sp.duration = 0;  // nothing left to inject
  sp.created = true;
  sp.chunked = false;
  int st = alloc_species_state(ctx, sp);
  if (st) {
    free_species(sp);
    return st;
  }
  const size_t b = sizeof(double) * (size_t)number;
  if (number > 0) {
    const double *src[6] = {r, z, vr, vphi, vz, w};
    double *dst[6] = {sp.f64.r, sp.f64.z, sp.f64.vr, sp.f64.vphi, sp.f64.vz, sp.f64.w};
    for (int j = 0; j < 6; j++) HIP_TRY(ctx, hipMemcpyAsync(dst[j], src[j], b, hipMemcpyHostToDevice, c.stream));
    HIP_TRY(ctx, hipMemsetAsync(sp.f64.alive, 1, (size_t)number, c.stream));
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  }
  c.species.push_back(sp);
  const int idx = (int)c.species.size() - 1;
  c.order.push_back(idx);
  if (is_f32(c)) {
    if (!launch_f32_from_f64(c, c.species[idx])) return fail(ctx, PDP3_ERR_OOM, "FP32 species arrays");
    CHECK_LAUNCH(ctx);
  }
  if (out_index) *out_index = idx;
  return PDP3_OK;
}

int pdp3_species_upload(pdp3_ctx *ctx, int index, const double *r, const double *z, const double *vr,
                        const double *vphi, const double *vz, const double *w, const uint8_t *alive) {
  if (!valid_species(ctx, index)) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  Species &sp = c.species[index];
  const size_t b = sizeof(double) * (size_t)sp.n;
  if (sp.n == 0) return PDP3_OK;
  const double *src[6] = {r, z, vr, vphi, vz, w};
  double *dst[6] = {sp.f64.r, sp.f64.z, sp.f64.vr, sp.f64.vphi, sp.f64.vz, sp.f64.w};
  if (is_f32(c)) {
    // refresh the logical-order FP64 staging, overwrite what is given, repack
    launch_f64_from_f32(c, sp);
    CHECK_LAUNCH(ctx);
    for (int j = 0; j < 6; j++)
      if (src[j]) HIP_TRY(ctx, hipMemcpyAsync(dst[j], src[j], b, hipMemcpyHostToDevice, c.stream));
    if (alive) HIP_TRY(ctx, hipMemcpyAsync(sp.f64.alive, alive, (size_t)sp.n, hipMemcpyHostToDevice, c.stream));
    if (!launch_f32_from_f64(c, sp)) return fail(ctx, PDP3_ERR_OOM, "FP32 species arrays");
    CHECK_LAUNCH(ctx);
  } else if (sp.chunked) {
    void *tmp = nullptr;
    HIP_TRY(ctx, hipMalloc(&tmp, b));
    for (int j = 0; j < 6; j++)
      if (src[j]) {
        (void)hipMemcpyAsync(tmp, src[j], b, hipMemcpyHostToDevice, c.stream);
        launch_to_phys(c, (const double *)tmp, dst[j], sp.n, sp.n_cap);
      }
    if (alive) {
      (void)hipMemcpyAsync(tmp, alive, (size_t)sp.n, hipMemcpyHostToDevice, c.stream);
      launch_alive_to_phys(c, (const uint8_t *)tmp, sp.f64.alive, sp.n, sp.n_cap);
    }
    hipError_t e = hipStreamSynchronize(c.stream);
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(ctx, PDP3_ERR_HIP, hipGetErrorString(e));
    CHECK_LAUNCH(ctx);
  } else {
    for (int j = 0; j < 6; j++)
      if (src[j]) HIP_TRY(ctx, hipMemcpyAsync(dst[j], src[j], b, hipMemcpyHostToDevice, c.stream));
    if (alive) HIP_TRY(ctx, hipMemcpyAsync(sp.f64.alive, alive, (size_t)sp.n, hipMemcpyHostToDevice, c.stream));
  }
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

int pdp3_species_download(pdp3_ctx *ctx, int index, double *r, double *z, double *vr, double *vphi, double *vz,
                          double *w, uint8_t *alive) {
  if (!valid_species(ctx, index)) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  Species &sp = c.species[index];
  const size_t b = sizeof(double) * (size_t)sp.n;
  if (sp.n == 0) return PDP3_OK;
  double *dst[6] = {r, z, vr, vphi, vz, w};
  const double *src[6] = {sp.f64.r, sp.f64.z, sp.f64.vr, sp.f64.vphi, sp.f64.vz, sp.f64.w};
  if (!is_f32(c) && sp.chunked) {
    void *tmp = nullptr;
    HIP_TRY(ctx, hipMalloc(&tmp, b));
    hipError_t e = hipSuccess;
    for (int j = 0; j < 6 && e == hipSuccess; j++)
      if (dst[j]) {
        launch_to_logical(c, src[j], (double *)tmp, sp.n);
        e = hipMemcpyAsync(dst[j], tmp, b, hipMemcpyDeviceToHost, c.stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
      }
    if (alive && e == hipSuccess) {
      launch_alive_to_logical(c, sp.f64.alive, (uint8_t *)tmp, sp.n);
      e = hipMemcpyAsync(alive, tmp, (size_t)sp.n, hipMemcpyDeviceToHost, c.stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c.stream);
    }
    (void)hipFree(tmp);
    if (e != hipSuccess) return fail(ctx, PDP3_ERR_HIP, hipGetErrorString(e));
    return PDP3_OK;
  }
  if (is_f32(c)) {
    launch_f64_from_f32(c, sp);
    CHECK_LAUNCH(ctx);
  }
  for (int j = 0; j < 6; j++)
    if (dst[j]) HIP_TRY(ctx, hipMemcpyAsync(dst[j], src[j], b, hipMemcpyDeviceToHost, c.stream));
  if (alive) HIP_TRY(ctx, hipMemcpyAsync(alive, sp.f64.alive, (size_t)sp.n, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

int pdp3_species_init_uniform(pdp3_ctx *ctx, int index, uint64_t seed, int ppc, double density, int ramp_cells,
                              double temperature_ev) {
  if (!valid_species(ctx, index) || ppc <= 0 || !(density >= 0) || ramp_cells < 0 || !(temperature_ev >= 0))
    return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  Species &sp = c.species[index];
  long long want = (long long)ppc * (c.g.nr - 1) * (c.g.nz - 1);
  if (sp.kind != PDP3_SPECIES_PLASMA || want != sp.n)
    return fail(ctx, PDP3_ERR_BAD_ARG, "species size must be ppc * (n_grid_r-1) * (n_grid_z-1)");
  launch_init_uniform(c, sp, index, seed, ppc, density, ramp_cells, temperature_ev, !is_f32(c) && sp.chunked);
  CHECK_LAUNCH(ctx);
  if (is_f32(c)) {
    if (!launch_f32_from_f64(c, sp)) return fail(ctx, PDP3_ERR_OOM, "FP32 species arrays");
    CHECK_LAUNCH(ctx);
  }
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

int pdp3_particles_step_v(pdp3_ctx *ctx, int index) {
  if (!valid_species(ctx, index)) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  if (is_f32(c)) return fail(ctx, PDP3_ERR_STATE, "split entry points run in the FP64 form only");
  int st = ensure_split_scratch(ctx, c.species[index]);
  if (st) return st;
  launch_step_v_f64(c, c.species[index]);
  CHECK_LAUNCH(ctx);
  return PDP3_OK;
}

int pdp3_particles_full_current_distribution(pdp3_ctx *ctx, int index) {
  if (!valid_species(ctx, index)) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  if (is_f32(c)) return fail(ctx, PDP3_ERR_STATE, "split entry points run in the FP64 form only");
  if (!c.species[index].f64.r_old) return fail(ctx, PDP3_ERR_STATE, "step_v must run first");
  launch_full_current_f64(c, c.species[index]);
  CHECK_LAUNCH(ctx);
  return PDP3_OK;
}

int pdp3_particles_charge_weighting(pdp3_ctx *ctx, int index) {
  if (!valid_species(ctx, index)) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  if (is_f32(c))
    launch_charge_weighting_f32(c, c.species[index]);
  else
    launch_charge_weighting_f64(c, c.species[index]);
  CHECK_LAUNCH(ctx);
  return PDP3_OK;
}

}  // extern "C"

namespace {
// the Gauss residual into device buffers rho and res (cells() values each):
// rho cleared by a fill kernel, then the charge weighting of every species of
// the loop (bunches once they are created), then the residual, which writes
// every node
int gauss_into(pdp3_ctx *ctx, double *rho, double *res, double *residual_out, double *rho_out) {
  Ctx &c = ctx->c;
  int st;
  launch_fill(rho, c.cells(), 0.0, c.stream);
  for (int j : c.order) {
    if (is_f32(c))
      launch_charge_weighting_f32(c, c.species[j], rho);
    else
      launch_charge_weighting_f64(c, c.species[j], rho);
  }
  launch_gauss_residual(c, rho, res);
  CHECK_LAUNCH(ctx);
  if ((st = check_device_errors(ctx))) return st;
  const size_t b = c.cells() * sizeof(double);
  if (residual_out) HIP_TRY(ctx, hipMemcpyAsync(residual_out, res, b, hipMemcpyDeviceToHost, c.stream));
  if (rho_out) HIP_TRY(ctx, hipMemcpyAsync(rho_out, rho, b, hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}
}  // namespace

extern "C" {

int pdp3_gauss_residual(pdp3_ctx *ctx, double *residual_out, double *rho_out) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  int st;
  if (!c.rho_all && (st = dalloc(ctx, &c.rho_all, c.cells()))) return st;
  if (!c.gauss && (st = dalloc(ctx, &c.gauss, c.cells()))) return st;
  return gauss_into(ctx, c.rho_all, c.gauss, residual_out, rho_out);
}

int pdp3_gauss_residual_noalloc(pdp3_ctx *ctx, double *residual_out, double *rho_out) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  if (c.cells() > (size_t)PDP3_NOALLOC_MAX_NODES)
    return fail(ctx, PDP3_ERR_BAD_ARG,
                "grid has " + std::to_string(c.cells()) + " nodes; pdp3_gauss_residual_noalloc takes at most " +
                    std::to_string((long long)PDP3_NOALLOC_MAX_NODES) + " (PDP3_NOALLOC_MAX_NODES)");
  if (!c.noalloc)
    return fail(ctx, PDP3_ERR_HIP, "the device address of the no-allocation array could not be resolved");
  return gauss_into(ctx, c.noalloc, c.noalloc + (size_t)PDP3_NOALLOC_MAX_NODES, residual_out, rho_out);
}

int pdp3_temperature(pdp3_ctx *ctx, int index, double *out) {
  if (!valid_species(ctx, index) || !out) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  int st;
  if (!c.probe_ws && (st = dalloc(ctx, &c.probe_ws, 6 * c.cells()))) return st;
  double *res = c.probe_ws + 5 * c.cells();
  if (!launch_temperature(c, c.species[index], c.probe_ws, res))
    return fail(ctx, PDP3_ERR_STATE, "the species has no particle state yet");
  CHECK_LAUNCH(ctx);
  if ((st = check_device_errors(ctx))) return st;
  HIP_TRY(ctx, hipMemcpyAsync(out, res, c.cells() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

int pdp3_density(pdp3_ctx *ctx, int index, double *out) {
  if (!valid_species(ctx, index) || !out) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  int st;
  if (!c.probe_ws && (st = dalloc(ctx, &c.probe_ws, 6 * c.cells()))) return st;
  double *res = c.probe_ws + 5 * c.cells();
  if (!launch_density(c, c.species[index], res))
    return fail(ctx, PDP3_ERR_STATE, "the species has no particle state yet");
  CHECK_LAUNCH(ctx);
  if ((st = check_device_errors(ctx))) return st;
  HIP_TRY(ctx, hipMemcpyAsync(out, res, c.cells() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

int pdp3_bunch_inject(pdp3_ctx *ctx, int index) {
  if (!valid_species(ctx, index) || ctx->c.species[index].kind != PDP3_SPECIES_BUNCH) return PDP3_ERR_BAD_ARG;
  return inject_bunch(ctx, ctx->c.species[index]);
}

}  // extern "C"

namespace {
int sort_plasma(pdp3_ctx *ctx, bool exact);

// the periodic sorts' buffers, allocated on the first step that sorts
// periodically (not inside the sort itself, so that a warmed-up loop
// allocates nothing)
int prepare_sorts(pdp3_ctx *ctx) {
  Ctx &c = ctx->c;
  for (int j : c.order) {
    Species &sp = c.species[j];
    if (sp.kind != PDP3_SPECIES_PLASMA || sp.sort_prepared) continue;
    if (sort_prepare(c, sp)) return fail(ctx, PDP3_ERR_OOM, "sort buffers");
  }
  return PDP3_OK;
}

// one loop iteration (src/loadInitParam.cpp:471-587); after_h (may be null)
// runs right after the H update, when E^n and H at the half time are final
int step_one(pdp3_ctx *ctx, const std::function<int()> *after_h) {
  Ctx &c = ctx->c;
  int st;
  if (c.sort_interval > 0 && (st = prepare_sorts(ctx))) return st;
  if (c.sort_interval > 0 && c.steps_done > 0 && c.steps_done % c.sort_interval == 0) {
    if ((st = sort_plasma(ctx, false))) return st;
  }
  if ((st = init_beam(ctx))) return st;
  if ((st = pdp3_hfield_calc(ctx))) return st;
  if (after_h && (st = (*after_h)())) return st;
  // Current::reset_j and the bunch rho's reset in one launch (J_r, J_phi,
  // J_z and rho_bunch are consecutive components of the grid)
  static_assert(kJPhi == kJR + 1 && kJZ == kJR + 2 && kRhoBunch == kJR + 3, "grid component order");
  if ((st = zero_grid(ctx, kJR, 4))) return st;
  for (int j : c.order)
    if ((st = particles_fused(ctx, c.species[j]))) return st;
  if ((st = pdp3_efield_calc(ctx))) return st;
  c.current_time = c.current_time + c.g.dt;  // src/loadInitParam.cpp:586
  c.steps_done++;
  return PDP3_OK;
}
}  // namespace

extern "C" {

int pdp3_step(pdp3_ctx *ctx, int nsteps) {
  if (!ctx || nsteps < 0) return PDP3_ERR_BAD_ARG;
  for (int n = 0; n < nsteps; n++) {
    int st = step_one(ctx, nullptr);
    if (st) return st;
  }
  return check_device_errors(ctx);
}

int pdp3_sync(pdp3_ctx *ctx) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  return check_device_errors(ctx);
}

namespace {
// exact: see sort_species; the periodic sorts of the loop are not exact
int sort_plasma(pdp3_ctx *ctx, bool exact) {
  Ctx &c = ctx->c;
  TimedScope ts(c, PDP3_KERNEL_SORT);
  // bunches are injected by particle index (src/bunch.cpp:106-117) and stay unsorted
  for (int j : c.order) {
    if (c.species[j].kind != PDP3_SPECIES_PLASMA) continue;
    int st = sort_species(c, c.species[j], exact);
    if (st) return fail(ctx, st == -3 ? PDP3_ERR_OOM : PDP3_ERR_HIP, "sort failed");
  }
  CHECK_LAUNCH(ctx);
  // an exact sort sizes its record space for one run per particle (~24 bytes
  // per particle): give that back down to what the periodic sorts use
  if (exact && sort_shrink_ws(c, sort_periodic_bytes(c)))
    return fail(ctx, PDP3_ERR_OOM, "sort workspace");
  return PDP3_OK;
}
}  // namespace

int pdp3_sort_particles(pdp3_ctx *ctx) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  return sort_plasma(ctx, true);
}

int pdp3_set_sort_interval(pdp3_ctx *ctx, int sort_interval) {
  if (!ctx || sort_interval < 0) return PDP3_ERR_BAD_ARG;
  ctx->c.sort_interval = sort_interval;
  return PDP3_OK;
}

int pdp3_set_strict_radial_fix(pdp3_ctx *ctx, int on) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  ctx->c.g.strict_fix = on ? 1 : 0;
  return PDP3_OK;
}

int pdp3_set_tile(pdp3_ctx *ctx, int ppt, int waves) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  for (int t = 0; t < kNumTiles; t++)
    if (kTileShapes[t][0] == ppt && kTileShapes[t][1] == waves) {
      ctx->c.tile = t;
      return PDP3_OK;
    }
  return fail(ctx, PDP3_ERR_BAD_ARG, "unsupported tile shape");
}

int pdp3_get_tile(const pdp3_ctx *ctx, int *ppt, int *waves) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  if (ppt) *ppt = kTileShapes[ctx->c.tile][0];
  if (waves) *waves = kTileShapes[ctx->c.tile][1];
  return PDP3_OK;
}

int pdp3_autotune(pdp3_ctx *ctx, int steps_per_candidate, int *ppt, int *waves) {
  if (!ctx || steps_per_candidate < 1) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  if (!is_f32(c)) return pdp3_get_tile(ctx, ppt, waves);  // one tile shape in the FP64 form
  // each candidate times real steps of the simulation on the plasma kernel
  const bool was_profiling = c.profiling;
  KernelTimer saved = c.timers[PDP3_KERNEL_PARTICLES];
  int st, best = 0;
  double best_ms = 1e300;
  for (int t = 0; t < kNumTiles; t++) {
    c.tile = t;
    c.profiling = true;
    if ((st = check_device_errors(ctx))) break;
    c.timers[PDP3_KERNEL_PARTICLES].total_ms = 0;
    c.timers[PDP3_KERNEL_PARTICLES].launches = 0;
    if ((st = pdp3_step(ctx, steps_per_candidate))) break;
    if ((st = check_device_errors(ctx))) break;
    const double ms = c.timers[PDP3_KERNEL_PARTICLES].total_ms;
    if (ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  }
  c.tile = best;
  c.profiling = was_profiling;
  c.timers[PDP3_KERNEL_PARTICLES].total_ms = saved.total_ms;
  c.timers[PDP3_KERNEL_PARTICLES].launches = saved.launches;
  if (st) return st;
  return pdp3_get_tile(ctx, ppt, waves);
}

// ----------------------------------------------------------------- adjoint

}  // extern "C"

namespace {

// a device buffer owned by the adjoint driver
struct DBuf {
  void *p = nullptr;
  ~DBuf() {
    if (p) (void)hipFree(p);
  }
  template <class T>
  T *as() const {
    return (T *)p;
  }
};

struct SpeciesCkpt {
  std::vector<DBuf> bufs;  // native state arrays
};

// the state a loop iteration reads: fields E, H (and the rest of the grid for
// the final restore) and every species' native particle arrays, plus the host
// bookkeeping of time and bunches
struct Checkpoint {
  DBuf grid;  // components kER .. kRhoBunch
  std::vector<SpeciesCkpt> sp;
  double time = 0;
  long long steps = 0;
  int next_bunch = 0;
  std::vector<int> order;
  std::vector<char> created;
};

constexpr int kCkComps = kRhoBunch + 1;

std::vector<std::pair<void *, size_t>> native_arrays(Ctx &c, Species &sp) {
  std::vector<std::pair<void *, size_t>> out;
  const size_t n = (size_t)sp.n_cap;
  if (is_f32(c)) {
    out.push_back({sp.f32.cell, n * sizeof(int32_t)});
    for (float *a : {sp.f32.fr, sp.f32.fz, sp.f32.vr, sp.f32.vphi, sp.f32.vz}) out.push_back({a, n * sizeof(float)});
  } else {
    for (double *a : {sp.f64.r, sp.f64.z, sp.f64.vr, sp.f64.vphi, sp.f64.vz}) out.push_back({a, n * sizeof(double)});
    out.push_back({sp.f64.alive, n});
  }
  return out;
}

int ckpt_save(pdp3_ctx *ctx, Checkpoint &ck) {
  Ctx &c = ctx->c;
  const size_t gb = kCkComps * c.cells() * sizeof(double);
  if (!ck.grid.p) HIP_TRY(ctx, hipMalloc(&ck.grid.p, gb));
  HIP_TRY(ctx, hipMemcpyAsync(ck.grid.p, c.a[0], gb, hipMemcpyDeviceToDevice, c.stream));
  if (ck.sp.empty()) ck.sp.resize(c.species.size());
  for (size_t j = 0; j < c.species.size(); j++) {
    auto arrs = native_arrays(c, c.species[j]);
    auto &b = ck.sp[j].bufs;
    if (b.empty()) {
      b = std::vector<DBuf>(arrs.size());
      for (size_t a = 0; a < arrs.size(); a++) HIP_TRY(ctx, hipMalloc(&b[a].p, std::max<size_t>(arrs[a].second, 1)));
    }
    for (size_t a = 0; a < arrs.size(); a++)
      if (arrs[a].second)
        HIP_TRY(ctx, hipMemcpyAsync(b[a].p, arrs[a].first, arrs[a].second, hipMemcpyDeviceToDevice, c.stream));
  }
  ck.time = c.current_time;
  ck.steps = c.steps_done;
  ck.next_bunch = c.next_bunch;
  ck.order = c.order;
  ck.created.clear();
  for (auto &sp : c.species) ck.created.push_back(sp.created ? 1 : 0);
  return PDP3_OK;
}

int ckpt_restore(pdp3_ctx *ctx, const Checkpoint &ck) {
  Ctx &c = ctx->c;
  const size_t gb = kCkComps * c.cells() * sizeof(double);
  HIP_TRY(ctx, hipMemcpyAsync(c.a[0], ck.grid.p, gb, hipMemcpyDeviceToDevice, c.stream));
  for (size_t j = 0; j < c.species.size(); j++) {
    auto arrs = native_arrays(c, c.species[j]);
    for (size_t a = 0; a < arrs.size(); a++)
      if (arrs[a].second)
        HIP_TRY(ctx, hipMemcpyAsync(arrs[a].first, ck.sp[j].bufs[a].p, arrs[a].second, hipMemcpyDeviceToDevice,
                                    c.stream));
  }
  c.current_time = ck.time;
  c.steps_done = ck.steps;
  c.next_bunch = ck.next_bunch;
  c.order = ck.order;
  for (size_t j = 0; j < c.species.size(); j++) c.species[j].created = ck.created[j] != 0;
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  return PDP3_OK;
}

// FP64 particle state of one step, in the index space of the species (slots
// of the FP64 form, logical order of the FP32 form's staging)
struct SpeciesSnap {
  DBuf r, z, vr, vphi, vz, alive;
};

long long adj_count(const Ctx &c, const Species &sp) { return is_f32(c) ? sp.n : sp.n_cap; }

// the tangent-linear model carried along a forward sweep of the loop:
// kTangentDirs directions of the particle weights; its grids are advanced by
// the context's own (linear) field updates through shadow contexts that
// share the geometry, PML coefficients and stream; the particle tangents by
// the dual-number step (launch_tlm_particles) right after each H update
struct TangentSweep {
  static constexpr int D = kTangentDirs;
  pdp3_ctx *ctx = nullptr;
  DBuf tgrid_mem, phi_dev;
  std::vector<Ctx> shadow;
  double *tg[D][kNComp] = {};
  double *const *tgp[D] = {};
  std::vector<DBuf> tstate;  // per species: r, z, vr, vphi, vz tangents and the weight tangents (D x n each)

  int init(pdp3_ctx *c_, const double *const *dw) {
    ctx = c_;
    Ctx &c = ctx->c;
    const size_t cells = c.cells();
    const int nsp = (int)c.species.size();
    HIP_TRY(ctx, hipMalloc(&tgrid_mem.p, (size_t)D * kNComp * cells * sizeof(double)));
    HIP_TRY(ctx, hipMemsetAsync(tgrid_mem.p, 0, (size_t)D * kNComp * cells * sizeof(double), c.stream));
    HIP_TRY(ctx, hipMalloc(&phi_dev.p, (1 + D) * sizeof(double)));
    HIP_TRY(ctx, hipMemsetAsync(phi_dev.p, 0, (1 + D) * sizeof(double), c.stream));
    shadow.assign(D, Ctx());
    for (int d = 0; d < D; d++) {
      Ctx &sh = shadow[d];
      sh.g = c.g;
      sh.stream = c.stream;
      sh.koef_e = c.koef_e;
      sh.koef_h = c.koef_h;
      for (int k = 0; k < kNComp; k++)
        sh.a[k] = tg[d][k] = tgrid_mem.as<double>() + ((size_t)d * kNComp + k) * cells;
      tgp[d] = tg[d];
    }
    tstate = std::vector<DBuf>((size_t)nsp * 6);
    for (int j = 0; j < nsp; j++) {
      Species &sp = c.species[j];
      const long long n = adj_count(c, sp);
      const size_t b = (size_t)std::max<long long>(n, 1) * D * sizeof(double);
      for (int a = 0; a < 6; a++) {
        HIP_TRY(ctx, hipMalloc(&tstate[j * 6 + a].p, b));
        HIP_TRY(ctx, hipMemsetAsync(tstate[j * 6 + a].p, 0, b, c.stream));
      }
      if (dw && dw[j] && n > 0) {
        // logical order -> the index space of the adjoint (FP64 form: slots)
        std::vector<double> h((size_t)n * D, 0.0);
        for (int d = 0; d < D; d++)
          for (long long p = 0; p < n; p++) {
            const long long l = (!is_f32(c) && sp.chunked) ? logical_of(p) : p;
            if (l < sp.n) h[(size_t)d * n + p] = dw[j][(size_t)d * sp.n + l];
          }
        HIP_TRY(ctx, hipMemcpyAsync(tstate[j * 6 + 5].p, h.data(), h.size() * sizeof(double),
                                    hipMemcpyHostToDevice, c.stream));
        HIP_TRY(ctx, hipStreamSynchronize(c.stream));
      }
    }
    return PDP3_OK;
  }
  // after the H update of an iteration: E^n and Hh^n are final and the
  // particles are still at step n
  int after_h() {
    Ctx &c = ctx->c;
    const size_t cells = c.cells();
    for (int d = 0; d < D; d++) {
      launch_h_update(shadow[d]);
      launch_fill(tg[d][kJR], 3 * cells, 0.0, c.stream);  // J_r, J_phi, J_z are contiguous
    }
    const double *f6[6] = {c.a[kER], c.a[kEPhi], c.a[kEZ], c.a[kHHR], c.a[kHHPhi], c.a[kHHZ]};
    for (int j = 0; j < (int)c.species.size(); j++) {
      Species &sp = c.species[j];
      const long long cnt = adj_count(c, sp);
      if (cnt <= 0) continue;
      if (is_f32(c)) launch_f64_from_f32(c, sp);  // logical-order FP64 staging
      double *ts[5];
      for (int a = 0; a < 5; a++) ts[a] = tstate[j * 6 + a].as<double>();
      launch_tlm_particles(c.g, f6, tgp, sp.f64.r, sp.f64.z, sp.f64.vr, sp.f64.vphi, sp.f64.vz, sp.f64.w,
                           sp.f64.alive, ts, tstate[j * 6 + 5].as<double>(), sp.charge, sp.mass, cnt,
                           sp.kind == PDP3_SPECIES_BUNCH, c.pusher, c.stream);
    }
    CHECK_LAUNCH(ctx);
    return PDP3_OK;
  }
  // after the iteration (E^{n+1} final): the tangent E update and, inside the
  // objective's window, its tangent
  int after_step(bool in_window, int k0, int k1) {
    Ctx &c = ctx->c;
    for (int d = 0; d < D; d++) launch_e_update(shadow[d]);
    if (in_window && k1 >= k0)
      launch_tlm_objective(c.g, c.a[kEZ], tgp, k0, k1, c.g.dt * c.g.dz, phi_dev.as<double>(), c.stream);
    CHECK_LAUNCH(ctx);
    return PDP3_OK;
  }
  int result(double *phi_out, double *dphi_out) {
    double out[1 + D];
    HIP_TRY(ctx, hipMemcpy(out, phi_dev.p, sizeof(out), hipMemcpyDeviceToHost));
    if (phi_out) *phi_out = out[0];
    if (dphi_out)
      for (int d = 0; d < D; d++) dphi_out[d] = out[1 + d];
    return PDP3_OK;
  }
};

// the axis window of the objective: E_z(0, k) sits at z = (k + 1/2) dz
void objective_window(const Ctx &c, double z_lo, double z_hi, int &k0, int &k1) {
  k0 = std::max((int)std::ceil(z_lo / c.g.dz - 0.5), 0);
  k1 = std::min((int)std::floor(z_hi / c.g.dz - 0.5), c.g.nz - 2);
}

// the adjoint run, optionally carrying the tangent-linear model along its
// first forward sweep (dw non-null): the tangents then follow exactly the
// trajectory that the checkpoints record
int adjoint_impl(pdp3_ctx *ctx, int nsteps, int checkpoint_every, int window_steps, double z_lo, double z_hi,
                 double *phi_out, double *const *dphi_dw, const double *const *dw, double *dphi_tangent);

}  // namespace

extern "C" {

int pdp3_adjoint_run(pdp3_ctx *ctx, int nsteps, int checkpoint_every, int window_steps, double z_lo, double z_hi,
                     double *phi_out, double *const *dphi_dw) {
  return adjoint_impl(ctx, nsteps, checkpoint_every, window_steps, z_lo, z_hi, phi_out, dphi_dw, nullptr, nullptr);
}

int pdp3_adjoint_tangent_run(pdp3_ctx *ctx, int nsteps, int checkpoint_every, int window_steps, double z_lo,
                             double z_hi, const double *const *dw, double *phi_out, double *const *dphi_dw,
                             double *dphi_tangent) {
  if (!dw) return PDP3_ERR_BAD_ARG;
  return adjoint_impl(ctx, nsteps, checkpoint_every, window_steps, z_lo, z_hi, phi_out, dphi_dw, dw, dphi_tangent);
}

int pdp3_adjoint_run_batch(pdp3_ctx *const *ctxs, int n, int nsteps, int checkpoint_every, int window_steps,
                           double z_lo, double z_hi, double *phi_out, double *const *const *dphi_dw,
                           int *status_out) {
  if (!ctxs || n < 1) return PDP3_ERR_BAD_ARG;
  for (int m = 0; m < n; m++) {
    if (!ctxs[m]) return PDP3_ERR_BAD_ARG;
    for (int q = 0; q < m; q++)
      if (ctxs[q] == ctxs[m]) return fail(ctxs[m], PDP3_ERR_BAD_ARG, "a context appears twice in the batch");
  }
  // one host thread per member, each driving its own context (and stream);
  // the members' launches interleave on the device
  std::vector<int> st((size_t)n, PDP3_OK);
  std::vector<std::thread> workers;
  workers.reserve((size_t)n);
  for (int m = 0; m < n; m++)
    workers.emplace_back([&, m]() {
      if (hipSetDevice(ctxs[m]->c.device) != hipSuccess) {
        st[m] = fail(ctxs[m], PDP3_ERR_HIP, "hipSetDevice failed in the batch worker");
        return;
      }
      st[m] = adjoint_impl(ctxs[m], nsteps, checkpoint_every, window_steps, z_lo, z_hi,
                           phi_out ? phi_out + m : nullptr, dphi_dw ? dphi_dw[m] : nullptr, nullptr, nullptr);
    });
  for (auto &w : workers) w.join();
  int first = PDP3_OK;
  for (int m = 0; m < n; m++) {
    if (status_out) status_out[m] = st[m];
    if (first == PDP3_OK) first = st[m];
  }
  return first;
}

int pdp3_tangent_run(pdp3_ctx *ctx, int nsteps, int window_steps, double z_lo, double z_hi,
                     const double *const *dw, double *phi_out, double *dphi_out) {
  if (!ctx || nsteps < 1 || window_steps < 0 || window_steps > nsteps || !(z_hi >= z_lo)) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  if (c.sort_interval != 0)
    return fail(ctx, PDP3_ERR_STATE, "the tangent-linear run needs a fixed particle order (sort_interval 0)");
  int st;
  if ((st = check_device_errors(ctx))) return st;
  int k0, k1;
  objective_window(c, z_lo, z_hi, k0, k1);
  TangentSweep tl;
  if ((st = tl.init(ctx, dw))) return st;
  std::function<int()> hook = [&]() { return tl.after_h(); };
  for (int n = 0; n < nsteps; n++) {
    if ((st = step_one(ctx, &hook))) return st;
    if ((st = tl.after_step(n >= nsteps - window_steps, k0, k1))) return st;
  }
  if ((st = check_device_errors(ctx))) return st;
  return tl.result(phi_out, dphi_out);
}

}  // extern "C"

namespace {

int adjoint_impl(pdp3_ctx *ctx, int nsteps, int checkpoint_every, int window_steps, double z_lo, double z_hi,
                 double *phi_out, double *const *dphi_dw, const double *const *dw, double *dphi_tangent) {
  if (!ctx || nsteps < 1 || checkpoint_every < 1 || window_steps < 0 || window_steps > nsteps || !(z_hi >= z_lo))
    return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  if (c.sort_interval != 0) return fail(ctx, PDP3_ERR_STATE, "the adjoint needs a fixed particle order (sort_interval 0)");
  int st;
  if ((st = check_device_errors(ctx))) return st;
  const size_t cells = c.cells();
  const int nsp = (int)c.species.size();
  int k0, k1;
  objective_window(c, z_lo, z_hi, k0, k1);
  const double scale = c.g.dt * c.g.dz;

  // forward with checkpoints (and the tangents, when requested)
  const int nck = (nsteps + checkpoint_every - 1) / checkpoint_every;
  std::vector<Checkpoint> ck(nck);
  Checkpoint final_state;
  TangentSweep tl;
  if (dw && (st = tl.init(ctx, dw))) return st;
  std::function<int()> tl_hook = [&]() { return tl.after_h(); };
  for (int n = 0; n < nsteps; n++) {
    if (n % checkpoint_every == 0 && (st = ckpt_save(ctx, ck[n / checkpoint_every]))) return st;
    if ((st = step_one(ctx, dw ? &tl_hook : nullptr))) return st;
    if (dw && (st = tl.after_step(n >= nsteps - window_steps, k0, k1))) return st;
  }
  if ((st = check_device_errors(ctx)) || (st = ckpt_save(ctx, final_state))) return st;
  if (dw && (st = tl.result(nullptr, dphi_tangent))) return st;

  // adjoint state
  DBuf lam_grid, phi_dev, ez_rows;
  HIP_TRY(ctx, hipMalloc(&lam_grid.p, 12 * cells * sizeof(double)));
  HIP_TRY(ctx, hipMemsetAsync(lam_grid.p, 0, 12 * cells * sizeof(double), c.stream));
  double *lam[12];
  for (int j = 0; j < 12; j++) lam[j] = lam_grid.as<double>() + (size_t)j * cells;
  HIP_TRY(ctx, hipMalloc(&phi_dev.p, sizeof(double)));
  HIP_TRY(ctx, hipMemsetAsync(phi_dev.p, 0, sizeof(double), c.stream));
  const int K = checkpoint_every;
  HIP_TRY(ctx, hipMalloc(&ez_rows.p, (size_t)K * c.g.nz * sizeof(double)));
  std::vector<DBuf> lstate((size_t)nsp * 6);  // r, z, vr, vphi, vz, w per species
  for (int j = 0; j < nsp; j++) {
    const size_t n = (size_t)std::max<long long>(adj_count(c, c.species[j]), 1);
    for (int a = 0; a < 6; a++) {
      HIP_TRY(ctx, hipMalloc(&lstate[j * 6 + a].p, n * sizeof(double)));
      HIP_TRY(ctx, hipMemsetAsync(lstate[j * 6 + a].p, 0, n * sizeof(double), c.stream));
    }
  }
  // per-step snapshots of one segment: E^n, Hh^n and the FP64 particle states
  DBuf fsnap;
  HIP_TRY(ctx, hipMalloc(&fsnap.p, (size_t)K * 6 * cells * sizeof(double)));
  std::vector<SpeciesSnap> psnap((size_t)K * nsp);
  for (int t = 0; t < K; t++)
    for (int j = 0; j < nsp; j++) {
      const size_t n = (size_t)std::max<long long>(adj_count(c, c.species[j]), 1);
      SpeciesSnap &s = psnap[(size_t)t * nsp + j];
      for (DBuf *b : {&s.r, &s.z, &s.vr, &s.vphi, &s.vz}) HIP_TRY(ctx, hipMalloc(&b->p, n * sizeof(double)));
      HIP_TRY(ctx, hipMalloc(&s.alive.p, n));
    }

  for (int seg = nck - 1; seg >= 0; seg--) {
    const int n0 = seg * K, n1 = std::min(nsteps, n0 + K);
    if ((st = ckpt_restore(ctx, ck[seg]))) return st;
    // recompute the segment, keeping what each step's adjoint reads
    for (int n = n0; n < n1; n++) {
      const int t = n - n0;
      std::function<int()> snap = [&]() -> int {
        double *fs = fsnap.as<double>() + (size_t)t * 6 * cells;
        const int comps[6] = {kER, kEPhi, kEZ, kHHR, kHHPhi, kHHZ};
        for (int a = 0; a < 6; a++)
          HIP_TRY(ctx, hipMemcpyAsync(fs + (size_t)a * cells, c.a[comps[a]], cells * sizeof(double),
                                      hipMemcpyDeviceToDevice, c.stream));
        for (int j = 0; j < nsp; j++) {
          Species &sp = c.species[j];
          SpeciesSnap &s = psnap[(size_t)t * nsp + j];
          const long long n = adj_count(c, sp);
          if (n <= 0) continue;
          if (is_f32(c)) {
            launch_f64_from_f32(c, sp);  // logical-order FP64 staging
            CHECK_LAUNCH(ctx);
          }
          const double *src[5] = {sp.f64.r, sp.f64.z, sp.f64.vr, sp.f64.vphi, sp.f64.vz};
          DBuf *dst[5] = {&s.r, &s.z, &s.vr, &s.vphi, &s.vz};
          for (int a = 0; a < 5; a++)
            HIP_TRY(ctx, hipMemcpyAsync(dst[a]->p, src[a], (size_t)n * sizeof(double), hipMemcpyDeviceToDevice,
                                        c.stream));
          HIP_TRY(ctx, hipMemcpyAsync(s.alive.p, sp.f64.alive, (size_t)n, hipMemcpyDeviceToDevice, c.stream));
        }
        return PDP3_OK;
      };
      if ((st = step_one(ctx, &snap))) return st;
      HIP_TRY(ctx, hipMemcpyAsync(ez_rows.as<double>() + (size_t)t * c.g.nz, c.a[kEZ], c.g.nz * sizeof(double),
                                  hipMemcpyDeviceToDevice, c.stream));
    }
    if ((st = check_device_errors(ctx))) return st;
    // backward through the segment
    for (int n = n1 - 1; n >= n0; n--) {
      const int t = n - n0;
      if (n >= nsteps - window_steps && k1 >= k0)
        launch_adj_seed(c.g, ez_rows.as<double>() + (size_t)t * c.g.nz, lam[2], k0, k1, scale, phi_dev.as<double>(),
                        c.stream);
      launch_adj_fields_pre(c.g, lam, c.koef_e, c.koef_h, c.stream);
      const double *fs = fsnap.as<double>() + (size_t)t * 6 * cells;
      const double *f6[6] = {fs, fs + cells, fs + 2 * cells, fs + 3 * cells, fs + 4 * cells, fs + 5 * cells};
      for (int j = 0; j < nsp; j++) {
        Species &sp = c.species[j];
        SpeciesSnap &s = psnap[(size_t)t * nsp + j];
        double *ls[5];
        for (int a = 0; a < 5; a++) ls[a] = lstate[j * 6 + a].as<double>();
        launch_adj_particles(c.g, f6, lam, s.r.as<double>(), s.z.as<double>(), s.vr.as<double>(),
                             s.vphi.as<double>(), s.vz.as<double>(), sp.f64.w, s.alive.as<uint8_t>(), ls,
                             lstate[j * 6 + 5].as<double>(), sp.charge, sp.mass, adj_count(c, sp),
                             sp.kind == PDP3_SPECIES_BUNCH, c.pusher, c.stream);
      }
      launch_adj_fields_post(c.g, lam, c.koef_e, c.koef_h, c.stream);
      CHECK_LAUNCH(ctx);
    }
    HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  }
  // results
  double phi = 0;
  HIP_TRY(ctx, hipMemcpy(&phi, phi_dev.p, sizeof(double), hipMemcpyDeviceToHost));
  if (phi_out) *phi_out = phi;
  if (dphi_dw) {
    for (int j = 0; j < nsp; j++) {
      if (!dphi_dw[j]) continue;
      Species &sp = c.species[j];
      const long long n = adj_count(c, sp);
      std::vector<double> h((size_t)std::max<long long>(n, 1));
      HIP_TRY(ctx, hipMemcpy(h.data(), lstate[j * 6 + 5].p, (size_t)n * sizeof(double), hipMemcpyDeviceToHost));
      if (!is_f32(c) && sp.chunked) {
        for (long long p = 0; p < n; p++) {
          const long long l = logical_of(p);
          if (l < sp.n) dphi_dw[j][l] = h[(size_t)p];
        }
      } else {
        std::memcpy(dphi_dw[j], h.data(), (size_t)sp.n * sizeof(double));
      }
    }
  }
  return ckpt_restore(ctx, final_state);
}

}  // namespace

extern "C" {

int pdp3_objective_run(pdp3_ctx *ctx, int nsteps, int window_steps, double z_lo, double z_hi, double *phi_out) {
  if (!ctx || nsteps < 0 || window_steps < 0 || window_steps > nsteps || !(z_hi >= z_lo) || !phi_out)
    return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  int k0 = (int)std::ceil(z_lo / c.g.dz - 0.5), k1 = (int)std::floor(z_hi / c.g.dz - 0.5);
  k0 = std::max(k0, 0);
  k1 = std::min(k1, c.g.nz - 2);
  DBuf phi_dev, scratch;
  HIP_TRY(ctx, hipMalloc(&phi_dev.p, sizeof(double)));
  HIP_TRY(ctx, hipMalloc(&scratch.p, c.g.nz * sizeof(double)));
  HIP_TRY(ctx, hipMemsetAsync(phi_dev.p, 0, sizeof(double), c.stream));
  int st;
  for (int n = 0; n < nsteps; n++) {
    if ((st = step_one(ctx, nullptr))) return st;
    if (n >= nsteps - window_steps && k1 >= k0)
      launch_adj_seed(c.g, c.a[kEZ], scratch.as<double>(), k0, k1, c.g.dt * c.g.dz, phi_dev.as<double>(), c.stream);
  }
  if ((st = check_device_errors(ctx))) return st;
  HIP_TRY(ctx, hipMemcpy(phi_out, phi_dev.p, sizeof(double), hipMemcpyDeviceToHost));
  return PDP3_OK;
}

// ---------------------------------------------------------------- ensembles

}  // extern "C"

struct pdp3_ensemble {
  std::vector<pdp3_ctx *> ctxs;
  std::vector<Ctx *> cs;
  hipStream_t stream = nullptr;
  EnsFields *fields = nullptr;
  EnsParticles *parts = nullptr;
  int tile = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_tile = -1;
  int sort_interval = 0;
  long long steps_done = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double timed_ms = 0;
  long long timed_steps = 0;
  std::string last_error;
  // injected bunches: one item per (instance, bunch), their start tables and
  // the device step counter the graph's injection launch reads
  EnsInjectItem *inj = nullptr;
  int n_inj = 0, inj_max = 0;
  int *inj_tables = nullptr;
  long long *step_dev = nullptr;
  // the periodic sorts: every instance's plasma in one launch sequence
  MultiSort ms;
};

namespace {

int efail(pdp3_ensemble *e, int code, const std::string &msg) {
  e->last_error = msg;
  return code;
}

#define ENS_HIP(e, expr)                                                         \
  do {                                                                           \
    hipError_t _err = (expr);                                                    \
    if (_err != hipSuccess) return efail((e), PDP3_ERR_HIP, hipGetErrorString(_err)); \
  } while (0)

void ens_drop_graph(pdp3_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  e->exec = nullptr;
  e->graph = nullptr;
  e->graph_tile = -1;
}

// one loop iteration of every instance (src/loadInitParam.cpp:471-587; the
// bunch injections of init_beam from the device tables), captured once per
// tile shape and replayed
int ens_graph(pdp3_ensemble *e) {
  if (e->exec && e->graph_tile == e->tile) return PDP3_OK;
  ens_drop_graph(e);
  ENS_HIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
  ens_launch_inject(e->inj, e->n_inj, e->inj_max, e->step_dev, e->stream);
  ens_launch_h(e->fields, e->stream);
  ens_launch_zero(e->fields, e->stream);
  ens_launch_particles(e->parts, e->tile, e->stream);
  ens_launch_e(e->fields, e->stream);
  ens_launch_count(e->step_dev, e->stream);
  hipGraph_t g = nullptr;
  ENS_HIP(e, hipStreamEndCapture(e->stream, &g));
  e->graph = g;
  ENS_HIP(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
  e->graph_tile = e->tile;
  return PDP3_OK;
}

int ens_check(pdp3_ensemble *e) {
  ENS_HIP(e, hipStreamSynchronize(e->stream));
  for (pdp3_ctx *ctx : e->ctxs) {
    int st = check_device_errors(ctx);
    if (st) return efail(e, st, std::string("instance: ") + ctx->c.last_error);
  }
  return PDP3_OK;
}

// the injection tables: per instance, LoadInitParam::init_beam's creation of
// bunches and Bunch::bunch_inject's start index (src/loadInitParam.cpp:168-248,
// src/bunch.cpp:87-124) replayed on the host from the instance's clock,
// accumulated as the loop accumulates it, one entry per ensemble step until
// every injected bunch has finished
int ens_inject_build(pdp3_ensemble *e) {
  std::vector<EnsInjectItem> items;
  std::vector<std::vector<int>> tables;
  for (Ctx *cp : e->cs) {
    Ctx &c = *cp;
    double t = c.current_time, t_last = -1e300;
    std::vector<int> injected;  // species index of each injected bunch
    for (int j = 0; j < (int)c.species.size(); j++) {
      const Species &sp = c.species[j];
      if (sp.kind != PDP3_SPECIES_BUNCH || !(sp.duration > 0) || !sp.inject_u) continue;
      injected.push_back(j);
      t_last = std::max(t_last, sp.duration * (sp.bunch_number + 1) + sp.hole_duration * sp.bunch_number);
    }
    if (injected.empty()) continue;
    std::vector<char> created;
    for (const Species &sp : c.species) created.push_back(sp.created ? 1 : 0);
    int next = c.next_bunch;
    std::vector<std::vector<int>> tab(injected.size());
    for (long long step = 0; t < t_last + 2 * c.g.dt; step++) {
      if (step > 100000000LL) return efail(e, PDP3_ERR_BAD_ARG, "bunch injection spans more than 1e8 steps");
      int cand = -1;
      for (int j = 0; j < (int)c.species.size(); j++) {
        const Species &sp = c.species[j];
        if (sp.kind == PDP3_SPECIES_BUNCH && !created[j] && sp.bunch_number == next) cand = j;
      }
      if (cand >= 0 && t >= next * (c.species[cand].duration + c.species[cand].hole_duration)) {
        created[cand] = 1;
        next++;
      }
      for (size_t b = 0; b < injected.size(); b++) {
        const Species &sp = c.species[injected[b]];
        int start = -1;
        const double begin = sp.duration * sp.bunch_number + sp.hole_duration * sp.bunch_number;
        const double end = begin + sp.duration;
        const int steps_amount = (int)std::ceil(sp.duration / c.g.dt);
        if (created[injected[b]] && steps_amount > 0 && t >= begin && t < end) {
          const unsigned pis = (unsigned)sp.n / (unsigned)steps_amount;
          start = (int)((t - begin) / c.g.dt * pis);
        }
        tab[b].push_back(start);
      }
      t = t + c.g.dt;
    }
    for (size_t b = 0; b < injected.size(); b++) {
      Species &sp = c.species[injected[b]];
      EnsInjectItem it{};
      it.g = c.g;
      it.cell = (uint32_t *)sp.f32.cell;
      it.fb = sp.f32.fz;
      it.ur = sp.f32.vr;
      it.uphi = sp.f32.vphi;
      it.uz = sp.f32.vz;
      it.u = sp.inject_u;
      it.n = sp.n;
      it.count = (int)((unsigned)sp.n / (unsigned)std::max(1, (int)std::ceil(sp.duration / c.g.dt)));
      it.table_len = (int)tab[b].size();
      it.dl = sp.velocity * c.g.dt;
      it.velocity = sp.velocity;
      items.push_back(it);
      tables.push_back(std::move(tab[b]));
    }
  }
  ENS_HIP(e, hipMalloc((void **)&e->step_dev, sizeof(long long)));
  ENS_HIP(e, hipMemset(e->step_dev, 0, sizeof(long long)));
  e->n_inj = (int)items.size();
  if (items.empty()) return PDP3_OK;
  size_t total = 0;
  for (auto &t : tables) total += t.size();
  ENS_HIP(e, hipMalloc((void **)&e->inj_tables, std::max<size_t>(total, 1) * sizeof(int)));
  size_t off = 0;
  for (size_t b = 0; b < items.size(); b++) {
    ENS_HIP(e, hipMemcpy(e->inj_tables + off, tables[b].data(), tables[b].size() * sizeof(int),
                         hipMemcpyHostToDevice));
    items[b].start = e->inj_tables + off;
    off += tables[b].size();
    e->inj_max = std::max(e->inj_max, items[b].count);
  }
  ENS_HIP(e, hipMalloc((void **)&e->inj, items.size() * sizeof(EnsInjectItem)));
  ENS_HIP(e, hipMemcpy(e->inj, items.data(), items.size() * sizeof(EnsInjectItem), hipMemcpyHostToDevice));
  return PDP3_OK;
}

}  // namespace

extern "C" {

int pdp3_ensemble_create(pdp3_ctx *const *ctxs, int n, pdp3_ensemble **out) {
  if (!ctxs || n < 1 || !out) return PDP3_ERR_BAD_ARG;
  *out = nullptr;
  auto *e = new pdp3_ensemble;
  for (int j = 0; j < n; j++) {
    pdp3_ctx *ctx = ctxs[j];
    if (!ctx) {
      delete e;
      return PDP3_ERR_BAD_ARG;
    }
    Ctx &c = ctx->c;
    const bool ok = is_f32(c) && c.g.dt == ctxs[0]->c.g.dt;
    if (!ok) {
      delete e;
      return PDP3_ERR_BAD_ARG;
    }
    if (hipStreamSynchronize(c.stream) != hipSuccess) {
      delete e;
      return PDP3_ERR_HIP;
    }
    e->ctxs.push_back(ctx);
    e->cs.push_back(&c);
  }
  if (hipStreamCreate(&e->stream) != hipSuccess || hipEventCreate(&e->ev0) != hipSuccess ||
      hipEventCreate(&e->ev1) != hipSuccess) {
    pdp3_ensemble_destroy(e);
    return PDP3_ERR_HIP;
  }
  e->fields = ens_fields_build(e->cs);
  e->parts = e->fields ? ens_particles_build(e->cs) : nullptr;
  if (!e->fields || !e->parts) {
    pdp3_ensemble_destroy(e);
    return PDP3_ERR_OOM;
  }
  if (int st = ens_inject_build(e)) {
    pdp3_ensemble_destroy(e);
    return st;
  }
  *out = e;
  return PDP3_OK;
}

int pdp3_ensemble_destroy(pdp3_ensemble *e) {
  if (!e) return PDP3_OK;
  if (e->stream) (void)hipStreamSynchronize(e->stream);
  ens_drop_graph(e);
  ens_fields_free(e->fields);
  ens_particles_free(e->parts);
  if (e->inj) (void)hipFree(e->inj);
  if (e->inj_tables) (void)hipFree(e->inj_tables);
  if (e->step_dev) (void)hipFree(e->step_dev);
  sort_multi_free(e->ms);
  if (e->ev0) (void)hipEventDestroy(e->ev0);
  if (e->ev1) (void)hipEventDestroy(e->ev1);
  if (e->stream) (void)hipStreamDestroy(e->stream);
  delete e;
  return PDP3_OK;
}

const char *pdp3_ensemble_last_error(const pdp3_ensemble *e) { return e ? e->last_error.c_str() : ""; }

int pdp3_ensemble_set_sort_interval(pdp3_ensemble *e, int sort_interval) {
  if (!e || sort_interval < 0) return PDP3_ERR_BAD_ARG;
  e->sort_interval = sort_interval;
  return PDP3_OK;
}

int pdp3_ensemble_step(pdp3_ensemble *e, int nsteps) {
  if (!e || nsteps < 0) return PDP3_ERR_BAD_ARG;
  int st;
  bool timing = false;
  for (int s = 0; s < nsteps; s++) {
    if (e->sort_interval > 0 && e->steps_done > 0 && e->steps_done % e->sort_interval == 0) {
      // every instance's plasma in one sort on the ensemble's stream; the
      // previous step has finished before its particle tables are rewritten
      ENS_HIP(e, hipStreamSynchronize(e->stream));
      st = sort_multi(e->ms, e->cs, e->stream);
      if (st) return efail(e, st == -3 ? PDP3_ERR_OOM : PDP3_ERR_HIP, "ensemble sort failed");
      if (ens_particles_update(e->parts, e->cs) != 0) return efail(e, PDP3_ERR_HIP, "ensemble tables");
    }
    if ((st = ens_graph(e))) return st;
    // init_beam's bookkeeping (the bunches join the species list on the
    // host; their injection runs in the graph)
    for (pdp3_ctx *ctx : e->ctxs)
      if ((st = init_beam(ctx, false))) return efail(e, st, ctx->c.last_error);
    if (!timing) {
      ENS_HIP(e, hipEventRecord(e->ev0, e->stream));
      timing = true;
    }
    ENS_HIP(e, hipGraphLaunch(e->exec, e->stream));
    for (Ctx *c : e->cs) {
      c->current_time = c->current_time + c->g.dt;  // src/loadInitParam.cpp:586
      c->steps_done++;
    }
    e->steps_done++;
    // close the timed interval before a sort (which runs outside the graph)
    const bool sort_next = e->sort_interval > 0 && e->steps_done % e->sort_interval == 0;
    if (timing && (sort_next || s == nsteps - 1)) {
      ENS_HIP(e, hipEventRecord(e->ev1, e->stream));
      ENS_HIP(e, hipEventSynchronize(e->ev1));
      float ms = 0;
      ENS_HIP(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
      e->timed_ms += ms;
      timing = false;
    }
  }
  e->timed_steps += nsteps;
  return ens_check(e);
}

int pdp3_ensemble_timing(pdp3_ensemble *e, double *total_ms, long long *steps) {
  if (!e) return PDP3_ERR_BAD_ARG;
  if (total_ms) *total_ms = e->timed_ms;
  if (steps) *steps = e->timed_steps;
  e->timed_ms = 0;
  e->timed_steps = 0;
  return PDP3_OK;
}

int pdp3_ensemble_set_tile(pdp3_ensemble *e, int ppt, int waves) {
  if (!e) return PDP3_ERR_BAD_ARG;
  for (int t = 0; t < kNumTiles; t++)
    if (kTileShapes[t][0] == ppt && kTileShapes[t][1] == waves) {
      e->tile = t;
      return PDP3_OK;
    }
  return efail(e, PDP3_ERR_BAD_ARG, "unsupported tile shape");
}

int pdp3_ensemble_get_tile(const pdp3_ensemble *e, int *ppt, int *waves) {
  if (!e) return PDP3_ERR_BAD_ARG;
  if (ppt) *ppt = kTileShapes[e->tile][0];
  if (waves) *waves = kTileShapes[e->tile][1];
  return PDP3_OK;
}

int pdp3_ensemble_autotune(pdp3_ensemble *e, int steps_per_candidate, int *ppt, int *waves) {
  if (!e || steps_per_candidate < 1) return PDP3_ERR_BAD_ARG;
  double saved_ms = e->timed_ms;
  long long saved_steps = e->timed_steps;
  int best = 0, st = PDP3_OK;
  double best_ms = 1e300;
  for (int t = 0; t < kNumTiles; t++) {
    e->tile = t;
    e->timed_ms = 0;
    e->timed_steps = 0;
    if ((st = pdp3_ensemble_step(e, steps_per_candidate))) break;
    if (e->timed_ms < best_ms) {
      best_ms = e->timed_ms;
      best = t;
    }
  }
  e->tile = best;
  e->timed_ms = saved_ms;
  e->timed_steps = saved_steps;
  if (st) return st;
  return pdp3_ensemble_get_tile(e, ppt, waves);
}

int pdp3_profile_enable(pdp3_ctx *ctx, int enable) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  ctx->c.profiling = enable != 0;
  return PDP3_OK;
}

int pdp3_profile_get(pdp3_ctx *ctx, int kernel, double *total_ms, long long *launches) {
  if (!ctx || kernel < 0 || kernel >= PDP3_NKERNELS) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  harvest_timers(c);
  if (total_ms) *total_ms = c.timers[kernel].total_ms;
  if (launches) *launches = c.timers[kernel].launches;
  return PDP3_OK;
}

int pdp3_profile_reset(pdp3_ctx *ctx) {
  if (!ctx) return PDP3_ERR_BAD_ARG;
  Ctx &c = ctx->c;
  HIP_TRY(ctx, hipStreamSynchronize(c.stream));
  harvest_timers(c);
  for (auto &t : c.timers) {
    t.total_ms = 0;
    t.launches = 0;
  }
  return PDP3_OK;
}

}  // extern "C"
