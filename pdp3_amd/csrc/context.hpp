// Internal state of a pdp3 context and the launchers shared between the HIP
// translation units.  Not part of the public ABI (include/pdp3_hip.h is).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace pdp3 {

// constants, include/constant.h:10-27 (the reference's rounded SI values)
constexpr double kPi = 3.1415926535897932;
constexpr double kEpsilon0 = 8.85E-12;
constexpr double kElMass = 9.1E-31;
constexpr double kElCharge = 1.6E-19;
constexpr double kLightSpeed = 3.0E8;
constexpr double kMagnConst = 1.26E-6;
constexpr double kMnzl = 1e-15;

// device error bits, reported by pdp3_sync / pdp3_step
constexpr int kErrNonfinite = 1;
constexpr int kErrSuperluminal = 2;
constexpr int kErrDomain = 4;

enum GridComp {
  kER = 0, kEPhi, kEZ, kHR, kHPhi, kHZ, kHHR, kHHPhi, kHHZ,
  kJR, kJPhi, kJZ, kRhoBunch, kSigma, kEps, kNComp
};

// geometry and step constants, passed by value to every kernel
struct GridDesc {
  int nr, nz;           // node counts n_grid_r, n_grid_z
  double dr, dz;        // src/geometry.cpp:45-53
  double r_size, z_size;
  double dt;
  // the FP64 form's shared divisors, correctly rounded on the host: 1/dr,
  // 1/dz, CYL_VOL(dz, dr) and its reciprocal, and per radial index
  // i = 0 .. nr+1 the pair {CELL_VOLUME(i, dr, dz), 1/CELL_VOLUME(i, dr, dz)}
  // (include/geometry.h:33-39), built at context creation
  double inv_dr, inv_dz;
  double cyl, inv_cyl;
  const double2 *vol;
  // 1: the FP64 form evaluates the second piece of a strictly radial move
  // across a radial face with r0^2 / (i dr), as its other pieces do, instead
  // of the reference's r0 * r0 / i_n * dr (= (r0^2 / i) dr,
  // src/particles.cpp:941, :976); 0 (default): as the reference
  int strict_fix = 0;
};

// Physical layout of plasma species.  Particles are processed by wavefronts
// in chunks of 64 * kPPT: lane L of a chunk handles the kPPT consecutive
// logical (cell-sorted) particles L*kPPT .. L*kPPT+kPPT-1, and iteration j of
// all lanes reads the contiguous slots j*64 .. j*64+63, so loads stay
// coalesced while every lane walks its own run of same-cell particles.
// Species arrays hold n_cap = n rounded up to a whole number of chunks; the
// extra slots are dead padding.  Bunches use the plain layout (slot =
// logical index; their injection addresses particles by index).
constexpr int kPPT = 16;
constexpr int kChunk = 64 * kPPT;
constexpr int kWavesPerBlock = 4;
constexpr int kBlockParticles = kWavesPerBlock * kChunk;

__host__ __device__ __forceinline__ long long phys_of(long long s) {
  long long c = s / kChunk, w = s - c * kChunk;
  return c * kChunk + (w % kPPT) * 64 + w / kPPT;
}
__host__ __device__ __forceinline__ long long logical_of(long long p) {
  long long c = p / kChunk, w = p - c * kChunk;
  return c * kChunk + (w % 64) * kPPT + w / 64;
}

// tile shapes of the FP32 particle kernel: particles per lane (PPT) and waves
// per block; the autotuner picks one per context
constexpr int kNumTiles = 6;
constexpr int kTileShapes[kNumTiles][2] = {{16, 4}, {16, 2}, {16, 1}, {8, 4}, {8, 2}, {4, 4}};

// alive byte of the FP64 state: 0 dead, 1 alive, 2 padding slot
constexpr uint8_t kPadSlot = 2;
// FP32 cell word of a padding slot (a dead particle's word never equals it)
constexpr uint32_t kPadCell = 0xFFFFFFFFu;

// FP64 particle state of one species (structure of arrays)
struct SpeciesF64 {
  double *r = nullptr, *z = nullptr, *vr = nullptr, *vphi = nullptr, *vz = nullptr, *w = nullptr;
  uint8_t *alive = nullptr;
  // split-entry-point scratch: position before the push and sin(theta)
  double *r_old = nullptr, *z_old = nullptr, *sin_t = nullptr;
};

// FP32 cell-relative state of one species: cell = i * nz + k (or -1 dead),
// offsets in units of the cell size in (0, 1]
struct SpeciesF32 {
  int32_t *cell = nullptr;
  float *fr = nullptr, *fz = nullptr, *vr = nullptr, *vphi = nullptr, *vz = nullptr, *w = nullptr;
};

struct Species {
  int kind = 0;  // PDP3_SPECIES_PLASMA / PDP3_SPECIES_BUNCH
  int charge_e = 0;
  double mass_e = 0;
  double charge = 0, mass = 0;  // per real particle, src/particles.cpp:30-33
  int n = 0;          // particles (logical count)
  long long n_cap = 0;  // slots in the device arrays
  bool chunked = false; // chunk-transposed layout (plasma) or plain (bunch)
  SpeciesF64 f64;
  SpeciesF32 f32;
  // sort scratch (second buffer of the same layout) and keys
  SpeciesF64 f64_alt;
  SpeciesF32 f32_alt;
  // sort: capacity for run records, and the previous sort's run count (pinned,
  // copied back asynchronously; complete once runs_ev has fired)
  long long rec_cap = 0;
  uint32_t *runs_host = nullptr;
  hipEvent_t runs_ev = nullptr;
  bool sort_prepared = false;  // the buffers of periodic sorts are allocated
  // bunch
  double *inject_u = nullptr;
  double duration = 0, hole_duration = 0, velocity = 0;
  int bunch_number = 0;
  bool created = true;
};

struct KernelTimer {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
  double total_ms = 0;
  long long launches = 0;
};

struct Ctx {
  GridDesc g{};
  double current_time = 0, start_time = 0, relaxation_time = 0, end_time = 0;
  int pusher = 0;
  int precision = 0;
  hipStream_t stream = nullptr;
  int device = 0;
  double *grid = nullptr;  // kNComp contiguous nr*nz arrays
  double *a[kNComp] = {};
  double *koef_e = nullptr, *koef_h = nullptr;
  std::vector<double> sigma_host, eps_host;
  std::vector<Species> species;
  std::vector<int> order;  // ParticlesList order of created species
  int next_bunch = 0;
  int *err = nullptr;      // device error bits
  int *err_host = nullptr; // pinned mirror
  std::string last_error;
  bool profiling = false;
  KernelTimer timers[5];
  int sort_interval = 0;
  // tile shape of the FP32 particle kernel (index into kTileShapes)
  int tile = 0;
  long long steps_done = 0;
  // sort workspace
  int32_t *sort_keys = nullptr;
  int32_t *sort_perm = nullptr;
  size_t sort_ws_bytes = 0;
  void *sort_ws = nullptr;
  size_t sort_cap = 0;
  // Gauss-law diagnostic: charge density of all species and the residual
  double *rho_all = nullptr;
  double *gauss = nullptr;
  // probe workspace: 5 per-cell sums of the temperature probe, then the
  // probe output (6 x cells doubles, allocated by the first probe)
  double *probe_ws = nullptr;
  // device address of the library's fixed no-allocation array on this
  // context's GPU (nullptr when it could not be resolved)
  double *noalloc = nullptr;
  // FP32 form: per radial cell index i (0 .. nr-1) two float4 entries,
  // {1/(2i), 1/(2(i+1)), 1/(2(i+2)), 1/(2(i+3))} and {1/i, 1/(i+1), 1/(i+1/2),
  // i(i+1)}, with 1/0 stored as 0 (built at context creation)
  float4 *rows = nullptr;
  // FP64 form: the cell-volume table GridDesc::vol points to
  double2 *vol64 = nullptr;
  size_t cells() const { return (size_t)g.nr * (size_t)g.nz; }
};

// ---- launchers (each enqueues on ctx.stream) ----
void launch_h_update(Ctx &c);
void launch_e_update(Ctx &c);
void launch_fill(double *p, size_t n, double v, hipStream_t s);
void launch_gauss_residual(Ctx &c, const double *rho, double *out);

// ensembles of independent instances: device tables of every instance's
// grid (built once; the field pointers of a context never move)
struct EnsFields;
EnsFields *ens_fields_build(const std::vector<Ctx *> &ctxs);  // nullptr: out of memory
void ens_fields_free(EnsFields *e);
void ens_launch_h(EnsFields *e, hipStream_t st);
void ens_launch_e(EnsFields *e, hipStream_t st);
void ens_launch_zero(EnsFields *e, hipStream_t st);
// the plasma tiles and bunch blocks of every instance, one launch;
// rebuild after a sort (particle arrays move)
struct EnsParticles;
EnsParticles *ens_particles_build(const std::vector<Ctx *> &ctxs);
int ens_particles_update(EnsParticles *e, const std::vector<Ctx *> &ctxs);  // 0 on success
void ens_particles_free(EnsParticles *e);
void ens_launch_particles(EnsParticles *e, int tile, hipStream_t st);
// bunch injection of an ensemble (Bunch::bunch_inject, src/bunch.cpp:87-124)
// driven by a device step counter: per (instance, bunch) pair a table of the
// first particle injected at each ensemble step (-1: none), computed on the
// host from the instance's accumulated clock; the counter is incremented by
// the last launch of each step
struct EnsInjectItem {
  GridDesc g;
  uint32_t *cell;
  float *fb, *ur, *uphi, *uz;
  const double *u;
  const int *start;
  int n, count, table_len;
  double dl, velocity;
};
void ens_launch_inject(const EnsInjectItem *items, int n_items, int max_count, const long long *step,
                       hipStream_t st);
void ens_launch_count(long long *step, hipStream_t st);

// adjoint of one loop iteration (FP64); lam = adjoints of E_r, E_phi, E_z,
// H_r, H_phi, H_z, Hh_r, Hh_phi, Hh_z, J_r, J_phi, J_z (12 grids)
void launch_adj_fields_pre(const GridDesc &g, double *const *lam, const double *ke, const double *kh, hipStream_t st);
void launch_adj_fields_post(const GridDesc &g, double *const *lam, const double *ke, const double *kh,
                            hipStream_t st);
void launch_adj_particles(const GridDesc &g, const double *const *fields6, double *const *lam, const double *r,
                          const double *z, const double *vr, const double *vphi, const double *vz, const double *w,
                          const uint8_t *alive, double *const *lstate, double *lw, double charge, double mass,
                          long long n, int is_bunch, int pusher, hipStream_t st);
void launch_adj_seed(const GridDesc &g, const double *ez, double *lez, int k0, int k1, double scale, double *phi,
                     hipStream_t st);
// tangent-linear model of one iteration's particle step (kTangentDirs
// directions at once): tgrid[d][comp] are direction d's tangent grids (E, Hh
// read, J accumulated); tstate = 5 arrays (r, z, vr, vphi, vz) of
// kTangentDirs x n tangents, updated in place; tw = kTangentDirs x n weight
// tangents
constexpr int kTangentDirs = 4;
void launch_tlm_particles(const GridDesc &g, const double *const *fields6, double *const *const *tgrid,
                          const double *r, const double *z, const double *vr, const double *vphi, const double *vz,
                          const double *w, const uint8_t *alive, double *const *tstate, const double *tw,
                          double charge, double mass, long long n, int is_bunch, int pusher, hipStream_t st);
// phi5[0] += scale sum E_z(0,k)^2, phi5[1 + d] += scale sum 2 E_z E_z-dot_d over k0..k1
void launch_tlm_objective(const GridDesc &g, const double *ez, double *const *const *tgrid, int k0, int k1,
                          double scale, double *phi5, hipStream_t st);

// FP64 plain form
void launch_fused_f64(Ctx &c, Species &sp);
void launch_step_v_f64(Ctx &c, Species &sp);
void launch_full_current_f64(Ctx &c, Species &sp);
// rho: target grid (nullptr: the bunch density)
void launch_charge_weighting_f64(Ctx &c, Species &sp, double *rho = nullptr);
void launch_gather_f64(Ctx &c, int which, const double *r, const double *z, int n, double *out);
void launch_bunch_inject_f64(Ctx &c, Species &sp, unsigned start, unsigned count, double dl);

// FP32 form
void launch_fused_f32(Ctx &c, Species &sp);
void launch_charge_weighting_f32(Ctx &c, Species &sp, double *rho = nullptr);
void launch_bunch_inject_f32(Ctx &c, Species &sp, unsigned start, unsigned count, double dl);
bool launch_f32_from_f64(Ctx &c, Species &sp);  // pack absolute FP64 state into cell form
void launch_f64_from_f32(Ctx &c, Species &sp);  // unpack for download

// probes of one species from the resident state (either form):
// TemperatureClassic into out (sums: 5 x cells workspace) and
// ParticlesDensity into out; false when the species has no state
bool launch_temperature(Ctx &c, Species &sp, double *sums, double *out);
bool launch_density(Ctx &c, Species &sp, double *out);

// synthetic loader (writes the FP64 arrays; to_phys: in the chunked layout)
void launch_init_uniform(Ctx &c, Species &sp, int species_index, uint64_t seed, int ppc,
                         double density, int ramp_cells, double temperature_ev, bool to_phys);
// logical <-> chunked-layout copies of FP64 / byte arrays; padding slots of
// `alive` are marked kPadSlot
void launch_mark_padding(Ctx &c, uint8_t *alive, long long n, long long n_cap);
void launch_to_phys(Ctx &c, const double *logical, double *phys, long long n, long long n_cap);
void launch_to_logical(Ctx &c, const double *phys, double *logical, long long n);
void launch_alive_to_phys(Ctx &c, const uint8_t *logical, uint8_t *phys, long long n, long long n_cap);
void launch_alive_to_logical(Ctx &c, const uint8_t *phys, uint8_t *logical, long long n);

// cell sort of a plasma species.  exact: always sorts (record space for one
// run per particle); otherwise the record space follows the previous sort's
// run count, and a sort that finds more runs than that is skipped (the state
// is left in its order) and sizes the next one
int sort_species(Ctx &c, Species &sp, bool exact);
// allocates what the periodic sorts of a plasma species use (the second
// particle buffers, the workspace for the initial record space, the pinned
// run count), so that a loop that sorts allocates nothing after its first
// step; returns 0, or -3 when out of memory
int sort_prepare(Ctx &c, Species &sp);
// workspace bytes the periodic sorts of the context's plasma species need
// (their current record space), and a shrink of the shared workspace back to
// that size after an exact sort grew it (waits for the stream; 0 or -3)
// several plasma species (an ensemble's) sorted as one: the segment table
// the kernels read (global slots [slot_base, slot_base + n_cap), keys from
// key_base), and the state the multi-segment sort keeps between sorts
struct MultiSeg {
  const uint32_t *cell;
  long long slot_base;
  // keys [key_base, key_base + ncell) for the living particles (ncell: the
  // segment's number of live keys), then one for dead ones and one for
  // padding; band: cells ordered in bands of two rows, column by column
  uint32_t key_base, ncell, nzm1, band;
  void *in[7], *out[7];
};
struct MultiSort {
  MultiSeg *seg = nullptr;             // device table
  uint32_t *seg_of_chunk = nullptr;    // device, one entry per global chunk
  long long n = 0, nseg = 0, rec_cap = 0;
  void *ws = nullptr;
  size_t ws_bytes = 0;
  uint32_t *runs_host = nullptr;       // pinned: the previous sort's run count
  hipEvent_t runs_ev = nullptr;
  bool counted = false;
  std::vector<MultiSeg> seg_host;      // the table's source (kept for the asynchronous copy)
};
// sort the plasma species of the given contexts in one launch sequence on
// `stream` (same result as sorting each with sort_species, exact = false);
// swaps each species' arrays with its second buffer.  0, -1 (a species not
// in whole chunks), -3 (allocation or HIP error)
int sort_multi(MultiSort &ms, const std::vector<Ctx *> &cs, hipStream_t stream);
void sort_multi_free(MultiSort &ms);
size_t sort_periodic_bytes(const Ctx &c);
int sort_shrink_ws(Ctx &c, size_t need);

}  // namespace pdp3
