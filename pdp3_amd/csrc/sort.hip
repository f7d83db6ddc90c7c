// Stable counting sort of a species' particles by cell (no reference
// counterpart: it keeps gathers and deposits of neighbouring particles on the
// same cache lines).  Key = i * (nz-1) + k of the reference's cell index
// (include/geometry.h:44, clamped at 0); dead particles get the next key and
// end up behind the living ones, chunk padding slots after those.  The order
// among equal keys is the input order, so the result equals a stable sort by
// key; it is written back in the chunk-transposed layout.
//
// LSD radix sort with 8-bit digits (one counting sort per digit):
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
  const long long hist_n = (long long)kBins * nblocks;
  // workspace: 4 key/val buffers + histogram + scan scratch
  size_t need = (size_t)n * 4 * sizeof(uint32_t) + (size_t)(hist_n + hist_n / kScanBlock + 4096) * 2 * sizeof(uint32_t);
  if (c.sort_ws_bytes < need) {
    if (c.sort_ws) (void)hipFree(c.sort_ws);
    c.sort_ws = nullptr;
    c.sort_ws_bytes = 0;
    if (hipMalloc(&c.sort_ws, need) != hipSuccess) return -3;
    c.sort_ws_bytes = need;
  }
  uint32_t *ka = (uint32_t *)c.sort_ws, *va = ka + n, *kb = va + n, *vb = kb + n;
  uint32_t *hist = vb + n, *tmp = hist + hist_n;
  unsigned eb = (unsigned)((n + 255) / 256);
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
