// Sort of a species' particles by cell (no reference counterpart: it keeps
// gathers and deposits of neighbouring particles on the same cache lines).
// Key = i * (nz-1) + k of the reference's cell index (include/geometry.h:44,
// clamped at 0); dead particles get the next key and end up behind the living
// ones, chunk padding slots after those.  The result is written back in the
// chunk-transposed layout.
//
// Grids of up to 2^26 cells: one counting sort over the whole key, with one
// atomic per run of equal keys in a wavefront (the order among equal keys
// from different wavefronts is not kept).  Larger grids: a stable LSD radix
// sort with 8-bit digits (one counting sort per digit):
//   hist:    per block of 16384 elements, a 256-bin digit histogram
//   scan:    exclusive scan of the bin-major histogram matrix
//   scatter: per block, stable ranks from wavefront ballots (lanes holding
//            the same digit are matched with 8 ballots), running per-wave
//            digit counters in LDS, then a coalesced-per-digit write
// then every particle array is gathered through the resulting permutation.
#include "context.hpp"

namespace pdp3 {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kPerWave = 4096;
constexpr int kPerBlock = kWaves * kPerWave;  // 16384
constexpr int kBins = 256;

__global__ void __launch_bounds__(kThreads) k_hist(const uint32_t *__restrict__ keys, long long n, int shift,
                                                   uint32_t *__restrict__ hist, int nblocks) {
  __shared__ uint32_t cnt[kBins];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  long long base = (long long)blockIdx.x * kPerBlock;
  long long end = min(base + (long long)kPerBlock, n);
  for (long long j = base + threadIdx.x; j < end; j += kThreads) atomicAdd(&cnt[(keys[j] >> shift) & 0xFF], 1u);
  __syncthreads();
  hist[(size_t)threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

// exclusive scan, three phases: per-block sums, scan of the sums, add back
constexpr int kScanBlock = 1024;
__global__ void __launch_bounds__(kScanBlock) k_scan_local(uint32_t *data, long long n, uint32_t *sums) {
  __shared__ uint32_t s[kScanBlock];
  long long j = (long long)blockIdx.x * kScanBlock + threadIdx.x;
  uint32_t v = j < n ? data[j] : 0u;
  s[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    uint32_t t = threadIdx.x >= (unsigned)off ? s[threadIdx.x - off] : 0u;
    __syncthreads();
    s[threadIdx.x] += t;
    __syncthreads();
  }
  if (j < n) data[j] = s[threadIdx.x] - v;
  if (threadIdx.x == kScanBlock - 1) sums[blockIdx.x] = s[threadIdx.x];
}
__global__ void k_scan_add(uint32_t *data, long long n, const uint32_t *sums) {
  long long j = (long long)blockIdx.x * kScanBlock + threadIdx.x;
  if (j < n) data[j] += sums[blockIdx.x];
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  unsigned lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// lanes (among `active`) whose digit equals this lane's digit
__device__ __forceinline__ uint64_t match_digit(uint32_t d, uint64_t active) {
  uint64_t m = active;
  for (int b = 0; b < 8; b++) {
    uint64_t bal = __ballot((d >> b) & 1u);
    m &= ((d >> b) & 1u) ? bal : ~bal;
  }
  return m;
}

__global__ void __launch_bounds__(kThreads) k_scatter(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                                      uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                      long long n, int shift, const uint32_t *__restrict__ offs,
                                                      int nblocks) {
  __shared__ uint32_t run[kWaves][kBins];
  const int wave = threadIdx.x / 64;
  for (int t = threadIdx.x; t < kWaves * kBins; t += kThreads) run[t / kBins][t % kBins] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kPerBlock + (long long)wave * kPerWave;
  const unsigned lane = __lane_id();
  // phase 1: per-wave digit totals
  for (int it = 0; it < kPerWave / 64; it++) {
    long long j = base + it * 64 + lane;
    bool valid = j < n;
    uint64_t active = __ballot(valid);
    if (active == 0) break;
    uint32_t d = valid ? (kin[j] >> shift) & 0xFF : 0u;
    uint64_t m = match_digit(d, active);
    if (valid && (m & lanemask_lt()) == 0) run[wave][d] += (uint32_t)__popcll(m);
  }
  __syncthreads();
  // bases: global offset of this block for the digit plus earlier waves
  {
    int d = threadIdx.x;
    uint32_t acc = offs[(size_t)d * nblocks + blockIdx.x];
    for (int w = 0; w < kWaves; w++) {
      uint32_t c = run[w][d];
      run[w][d] = acc;
      acc += c;
    }
  }
  __syncthreads();
  // phase 2: stable positions and scatter
  for (int it = 0; it < kPerWave / 64; it++) {
    long long j = base + it * 64 + lane;
    bool valid = j < n;
    uint64_t active = __ballot(valid);
    if (active == 0) break;
    uint32_t key = valid ? kin[j] : 0u;
    uint32_t d = (key >> shift) & 0xFF;
    uint64_t m = match_digit(d, active);
    uint32_t pos = run[wave][d] + (uint32_t)__popcll(m & lanemask_lt());
    if (valid) {
      kout[pos] = key;
      vout[pos] = vin[j];
    }
    // make every lane's read of run[] happen before the leaders update it
    __builtin_amdgcn_wave_barrier();
    if (valid && (m & lanemask_lt()) == 0) run[wave][d] += (uint32_t)__popcll(m);
    __builtin_amdgcn_wave_barrier();
  }
}

// ---- counting sort by the whole key (grids of up to kCountBins cells) ----
//
// The particles of a sorted species stay nearly sorted between sorts, so the
// 64 consecutive keys of a wavefront form a few runs of equal keys.  Each run
// is counted (pass 1) and placed (pass 2) with one global atomic by its first
// lane; the lanes of a run take consecutive positions in their input order.
// Runs of equal keys from different wavefronts are placed in the order their
// atomics arrive, so the result is sorted by key but not stable across
// wavefronts.
constexpr uint32_t kCountBins = 1u << 26;

__global__ void k_fill_u32(uint32_t *p, long long n, uint32_t v) {
  long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) p[j] = v;
}

// this lane's run: first lane, length, and whether this lane leads it
__device__ __forceinline__ void key_run(uint32_t key, bool valid, int &head, int &len) {
  const unsigned lane = __lane_id();
  const uint32_t prev = __shfl_up(key, 1);
  const bool starts = valid && (lane == 0 || prev != key);
  const uint64_t heads = __ballot(starts);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const uint64_t hm = heads & upto;                     // heads at or below this lane
  head = hm ? 63 - __clzll(hm) : 0;
  const uint64_t after = heads & ~upto;                 // heads above this lane
  const int nvalid = __popcll(__ballot(valid));         // the valid lanes are a prefix
  const int end = after ? __ffsll((long long)after) - 1 : nvalid;
  len = end - head;
}

// sort keys straight from the particle state, by logical index s (the state
// is read through the chunk-transposed layout)
struct KeysF32 {
  GridDesc g;
  const uint32_t *cell;
  uint32_t dead_key;
  __device__ __forceinline__ uint32_t operator()(long long s) const {
    const uint32_t c = cell[phys_of(s)];
    if (c == kPadCell) return dead_key + 1;
    if (c & 0x80000000u) return dead_key;
    return ((c >> 16) & 0x7FFF) * (uint32_t)(g.nz - 1) + (c & 0xFFFF);
  }
};
struct KeysF64 {
  GridDesc g;
  const double *r, *z;
  const uint8_t *alive;
  uint32_t dead_key;
  __device__ __forceinline__ uint32_t operator()(long long s) const {
    const long long j = phys_of(s);
    if (alive[j] == kPadSlot) return dead_key + 1;
    if (alive[j] != 1) return dead_key;
    int i = (int)ceil(r[j] / g.dr) - 1, k = (int)ceil(z[j] / g.dz) - 1;
    i = max(0, min(i, g.nr - 2));
    k = max(0, min(k, g.nz - 2));
    return (uint32_t)i * (uint32_t)(g.nz - 1) + (uint32_t)k;
  }
};

template <class K>
__global__ void __launch_bounds__(256) k_count_runs(K keys, long long n, uint32_t *__restrict__ cnt) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = s < n;
  const uint32_t key = valid ? keys(s) : 0xFFFFFFFFu;
  int head, len;
  key_run(key, valid, head, len);
  if (valid && head == (int)__lane_id()) atomicAdd(&cnt[key], (uint32_t)len);
}

// cnt holds each key's first position (exclusive scan of the counts); it is
// advanced as runs are placed.  src[slot of sorted rank] = slot of the particle
template <class K>
__global__ void __launch_bounds__(256) k_place_runs(K keys, long long n, uint32_t *__restrict__ cnt,
                                                    uint32_t *__restrict__ src) {
  const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = s < n;
  const uint32_t key = valid ? keys(s) : 0xFFFFFFFFu;
  int head, len;
  key_run(key, valid, head, len);
  const unsigned lane = __lane_id();
  uint32_t base = 0;
  if (valid && head == (int)lane) base = atomicAdd(&cnt[key], (uint32_t)len);
  base = __shfl(base, head);
  if (valid) src[phys_of((long long)base + (lane - (unsigned)head))] = (uint32_t)phys_of(s);
}

// every particle array in one pass: out[p] = in[src[p]]
struct Arrays7 {
  void *a[7];
};
__global__ void k_gather_f32(Arrays7 in, Arrays7 out, const uint32_t *__restrict__ src, long long n) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t q = src[p];
#pragma unroll
  for (int k = 0; k < 7; k++) static_cast<uint32_t *>(out.a[k])[p] = static_cast<const uint32_t *>(in.a[k])[q];
}
__global__ void k_gather_f64(Arrays7 in, Arrays7 out, const uint32_t *__restrict__ src, long long n) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const uint32_t q = src[p];
#pragma unroll
  for (int k = 0; k < 6; k++) static_cast<double *>(out.a[k])[p] = static_cast<const double *>(in.a[k])[q];
  static_cast<uint8_t *>(out.a[6])[p] = static_cast<const uint8_t *>(in.a[6])[q];
}

// keys in logical order (so that ties keep their logical order); vals = slot
__global__ void k_keys_f64(GridDesc g, const double *r, const double *z, const uint8_t *alive, long long n,
                           uint32_t *keys, uint32_t *vals, uint32_t dead_key) {
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  long long j = phys_of(s);
  uint32_t key = alive[j] == kPadSlot ? dead_key + 1 : dead_key;
  if (alive[j] == 1) {
    int i = (int)ceil(r[j] / g.dr) - 1, k = (int)ceil(z[j] / g.dz) - 1;
    i = max(0, min(i, g.nr - 2));
    k = max(0, min(k, g.nz - 2));
    key = (uint32_t)i * (uint32_t)(g.nz - 1) + (uint32_t)k;
  }
  keys[s] = key;
  vals[s] = (uint32_t)j;
}

__global__ void k_keys_f32(GridDesc g, const uint32_t *cell, long long n, uint32_t *keys, uint32_t *vals,
                           uint32_t dead_key) {
  long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  long long j = phys_of(s);
  uint32_t c = cell[j];
  uint32_t key = c == kPadCell ? dead_key + 1 : dead_key;
  if (!(c & 0x80000000u)) key = ((c >> 16) & 0x7FFF) * (uint32_t)(g.nz - 1) + (c & 0xFFFF);
  keys[s] = key;
  vals[s] = (uint32_t)j;
}

// slot p of the chunked layout holds sorted rank logical_of(p)
template <typename T>
__global__ void k_gather(const T *__restrict__ in, T *__restrict__ out, const uint32_t *__restrict__ perm, long long n) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) out[p] = in[perm[logical_of(p)]];
}

void exclusive_scan(uint32_t *data, long long n, uint32_t *tmp, hipStream_t s) {
  long long nb = (n + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(kScanBlock), 0, s, data, n, tmp);
  if (nb > 1) {
    exclusive_scan(tmp, nb, tmp + nb, s);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kScanBlock), 0, s, data, n, tmp);
  }
}

template <typename T>
bool ensure_alt(T *&alt, long long n) {
  if (alt) return true;
  if (hipMalloc((void **)&alt, (size_t)n * sizeof(T)) != hipSuccess) {
    alt = nullptr;
    return false;
  }
  return true;
}

template <typename T>
bool gather_swap(T *&a, T *&alt, const uint32_t *perm, long long n, hipStream_t s) {
  if (!a) return true;
  if (!alt && hipMalloc((void **)&alt, (size_t)n * sizeof(T)) != hipSuccess) {
    alt = nullptr;
    return false;
  }
  unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_gather<T>, dim3(blocks), dim3(256), 0, s, a, alt, perm, n);
  T *t = a;
  a = alt;
  alt = t;
  return true;
}

}  // namespace

int sort_species(Ctx &c, Species &sp) {
  const long long n = sp.n_cap;
  if (n <= 1 || !sp.chunked) return 0;
  const uint32_t ncell = (uint32_t)(c.g.nr - 1) * (uint32_t)(c.g.nz - 1);
  const uint32_t dead_key = ncell;
  int bits = 1;
  while ((1ull << bits) <= (unsigned long long)dead_key + 1) bits++;
  const int passes = (bits + 7) / 8;
  const int nblocks = (int)((n + kPerBlock - 1) / kPerBlock);
  const bool counting = (unsigned long long)dead_key + 2 <= kCountBins;
  // counting sort: one bin per key; radix sort: the per-block digit histograms
  const long long hist_n = counting ? (long long)dead_key + 2 : (long long)kBins * nblocks;
  // workspace: key/val buffers (counting sort: the source slots only) +
  // histogram + scan scratch
  const size_t bufs = counting ? 1 : 4;
  size_t need = (size_t)n * bufs * sizeof(uint32_t) + (size_t)(hist_n + hist_n / kScanBlock + 4096) * 2 * sizeof(uint32_t);
  if (c.sort_ws_bytes < need) {
    if (c.sort_ws) (void)hipFree(c.sort_ws);
    c.sort_ws = nullptr;
    c.sort_ws_bytes = 0;
    if (hipMalloc(&c.sort_ws, need) != hipSuccess) return -3;
    c.sort_ws_bytes = need;
  }
  uint32_t *ka = (uint32_t *)c.sort_ws, *va = ka + n, *kb = va + n, *vb = kb + n;
  uint32_t *hist = (uint32_t *)c.sort_ws + bufs * n, *tmp = hist + hist_n;
  unsigned eb = (unsigned)((n + 255) / 256);
  if (counting) {
    // count runs of keys, scan, place runs: src[slot] = source slot; then all
    // arrays in one gather
    const unsigned hb = (unsigned)((hist_n + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3(hb), dim3(256), 0, c.stream, hist, hist_n, 0u);
    Arrays7 in{}, out{};
    unsigned gb = eb;
    if (c.precision == 1) {
      SpeciesF32 &a = sp.f32, &b = sp.f32_alt;
      if (!a.cell) return 0;
      if (!(ensure_alt(b.cell, n) && ensure_alt(b.fr, n) && ensure_alt(b.fz, n) && ensure_alt(b.vr, n) &&
            ensure_alt(b.vphi, n) && ensure_alt(b.vz, n) && ensure_alt(b.w, n)))
        return -3;
      const KeysF32 keys{c.g, (const uint32_t *)a.cell, dead_key};
      hipLaunchKernelGGL(k_count_runs<KeysF32>, dim3(eb), dim3(256), 0, c.stream, keys, n, hist);
      exclusive_scan(hist, hist_n, tmp, c.stream);
      hipLaunchKernelGGL(k_place_runs<KeysF32>, dim3(eb), dim3(256), 0, c.stream, keys, n, hist, ka);
      in = Arrays7{{a.cell, a.fr, a.fz, a.vr, a.vphi, a.vz, a.w}};
      out = Arrays7{{b.cell, b.fr, b.fz, b.vr, b.vphi, b.vz, b.w}};
      hipLaunchKernelGGL(k_gather_f32, dim3(gb), dim3(256), 0, c.stream, in, out, ka, n);
      std::swap(a.cell, b.cell);
      std::swap(a.fr, b.fr);
      std::swap(a.fz, b.fz);
      std::swap(a.vr, b.vr);
      std::swap(a.vphi, b.vphi);
      std::swap(a.vz, b.vz);
      std::swap(a.w, b.w);
    } else {
      SpeciesF64 &a = sp.f64, &b = sp.f64_alt;
      if (!(ensure_alt(b.r, n) && ensure_alt(b.z, n) && ensure_alt(b.vr, n) && ensure_alt(b.vphi, n) &&
            ensure_alt(b.vz, n) && ensure_alt(b.w, n) && ensure_alt(b.alive, n)))
        return -3;
      const KeysF64 keys{c.g, a.r, a.z, a.alive, dead_key};
      hipLaunchKernelGGL(k_count_runs<KeysF64>, dim3(eb), dim3(256), 0, c.stream, keys, n, hist);
      exclusive_scan(hist, hist_n, tmp, c.stream);
      hipLaunchKernelGGL(k_place_runs<KeysF64>, dim3(eb), dim3(256), 0, c.stream, keys, n, hist, ka);
      in = Arrays7{{a.r, a.z, a.vr, a.vphi, a.vz, a.w, a.alive}};
      out = Arrays7{{b.r, b.z, b.vr, b.vphi, b.vz, b.w, b.alive}};
      hipLaunchKernelGGL(k_gather_f64, dim3(gb), dim3(256), 0, c.stream, in, out, ka, n);
      std::swap(a.r, b.r);
      std::swap(a.z, b.z);
      std::swap(a.vr, b.vr);
      std::swap(a.vphi, b.vphi);
      std::swap(a.vz, b.vz);
      std::swap(a.w, b.w);
      std::swap(a.alive, b.alive);
    }
    return 0;
  }
  if (c.precision == 1) {
    if (!sp.f32.cell) return 0;
    hipLaunchKernelGGL(k_keys_f32, dim3(eb), dim3(256), 0, c.stream, c.g, (const uint32_t *)sp.f32.cell, n, ka, va,
                       dead_key);
  } else {
    hipLaunchKernelGGL(k_keys_f64, dim3(eb), dim3(256), 0, c.stream, c.g, sp.f64.r, sp.f64.z, sp.f64.alive, n, ka,
                       va, dead_key);
  }
  for (int p = 0; p < passes; p++) {
    int shift = 8 * p;
    hipLaunchKernelGGL(k_hist, dim3(nblocks), dim3(kThreads), 0, c.stream, ka, n, shift, hist, nblocks);
    exclusive_scan(hist, hist_n, tmp, c.stream);
    hipLaunchKernelGGL(k_scatter, dim3(nblocks), dim3(kThreads), 0, c.stream, ka, va, kb, vb, n, shift, hist,
                       nblocks);
    uint32_t *t = ka;
    ka = kb;
    kb = t;
    t = va;
    va = vb;
    vb = t;
  }
  const uint32_t *perm = va;
  bool ok = true;
  if (c.precision == 1) {
    ok = ok && gather_swap(sp.f32.cell, sp.f32_alt.cell, perm, n, c.stream);
    ok = ok && gather_swap(sp.f32.fr, sp.f32_alt.fr, perm, n, c.stream);
    ok = ok && gather_swap(sp.f32.fz, sp.f32_alt.fz, perm, n, c.stream);
    ok = ok && gather_swap(sp.f32.vr, sp.f32_alt.vr, perm, n, c.stream);
    ok = ok && gather_swap(sp.f32.vphi, sp.f32_alt.vphi, perm, n, c.stream);
    ok = ok && gather_swap(sp.f32.vz, sp.f32_alt.vz, perm, n, c.stream);
    ok = ok && gather_swap(sp.f32.w, sp.f32_alt.w, perm, n, c.stream);
  } else {
    ok = ok && gather_swap(sp.f64.r, sp.f64_alt.r, perm, n, c.stream);
    ok = ok && gather_swap(sp.f64.z, sp.f64_alt.z, perm, n, c.stream);
    ok = ok && gather_swap(sp.f64.vr, sp.f64_alt.vr, perm, n, c.stream);
    ok = ok && gather_swap(sp.f64.vphi, sp.f64_alt.vphi, perm, n, c.stream);
    ok = ok && gather_swap(sp.f64.vz, sp.f64_alt.vz, perm, n, c.stream);
    ok = ok && gather_swap(sp.f64.w, sp.f64_alt.w, perm, n, c.stream);
    ok = ok && gather_swap(sp.f64.alive, sp.f64_alt.alive, perm, n, c.stream);
  }
  if (!ok) return -3;
  return 0;
}

}  // namespace pdp3
