// Stable sort of a species' particles by cell (no reference counterpart: it
// keeps gathers and deposits of neighbouring particles on the same cache
// lines).  Key = i * (nz-1) + k of the reference's cell index
// (include/geometry.h:44, clamped at 0); dead particles get the next key and
// end up behind the living ones, chunk padding slots after those.  The order
// among equal keys is the input order; the result is written back in the
// chunk-transposed layout.
//
// A sorted species stays nearly sorted between sorts, so the sort works on
// runs rather than particles:
//   runs:    each wavefront's 64 consecutive keys split into runs of equal
//            keys (a run never spans two wavefronts); one record (key,
//            length) per run, in logical order
//   records: stable LSD radix sort of the records by key, 8-bit digits:
//              hist:    per block of 16384 records, a 256-bin digit histogram
//              scan:    exclusive scan of the bin-major histogram matrix
//              scatter: per block, stable ranks from wavefront ballots (lanes
//                       holding the same digit are matched with 8 ballots),
//                       running per-wave digit counters in LDS
//   ranks:   exclusive scan of the run lengths in sorted order: each run's
//            first rank; a particle's rank = its run's + its place in the run
// then every particle array is gathered through the resulting permutation in
// one pass (for a particle that stayed in place the gather reads its own
// slot, so a nearly sorted species is moved with coalesced reads).
#include "context.hpp"

namespace pdp3 {

namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr int kPerWave = 4096;
constexpr int kPerBlock = kWaves * kPerWave;  // 16384
constexpr int kBins = 256;

// the number of records present: the count on the device, or none when it
// exceeds the capacity (that sort is then skipped)
__device__ __forceinline__ long long present(const uint32_t *count, long long cap) {
  const long long m = *count;
  return m <= cap ? m : 0;
}

__global__ void __launch_bounds__(kThreads) k_hist(const uint32_t *__restrict__ keys, const uint32_t *count,
                                                   long long cap, int shift, uint32_t *__restrict__ hist,
                                                   int nblocks) {
  __shared__ uint32_t cnt[kBins];
  const long long n = present(count, cap);
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
                                                      const uint32_t *count, long long cap, int shift,
                                                      const uint32_t *__restrict__ offs, int nblocks) {
  __shared__ uint32_t run[kWaves][kBins];
  const long long n = present(count, cap);
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

// ---- runs of equal keys ----

__global__ void k_fill_u32(uint32_t *p, long long n, uint32_t v) {
  long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) p[j] = v;
}

// the runs of this wavefront's keys (lanes >= the valid prefix are not
// keys): heads = mask of first lanes; head, len = this lane's run
__device__ __forceinline__ uint64_t wave_runs(uint32_t key, bool valid, int &head, int &len) {
  const unsigned lane = __lane_id();
  const uint32_t prev = __shfl_up(key, 1);
  const bool starts = valid && (lane == 0 || prev != key);
  const uint64_t heads = __ballot(starts);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1);
  const uint64_t hm = heads & upto;  // heads at or below this lane
  head = hm ? 63 - __clzll(hm) : 0;
  const uint64_t after = heads & ~upto;
  const int nvalid = __popcll(__ballot(valid));  // the valid lanes are a prefix
  const int end = after ? __ffsll((long long)after) - 1 : nvalid;
  len = end - head;
  return heads;
}

// The run kernels take kCPB chunks per block (their loads all in flight
// before the block's one barrier).  A chunk's keys are evaluated in slot order
// (coalesced) and transposed to logical order in LDS (rows padded to 17
// words, so that neither direction has bank conflicts); the block then walks
// them in groups of 256 logical keys, one record count (block_runs /
// block_base entry) per group.  nch: the species' number of chunks.
constexpr int kCPB = 2;
constexpr int kGroups = kChunk / 256;         // groups per chunk
constexpr int kBlockGroups = kCPB * kGroups;  // groups per block
constexpr int kRow = kPPT + 1;
constexpr int kChunkLds = 64 * kRow;

// the number of this block's chunks (the last block may hold fewer)
__device__ __forceinline__ int block_chunks(long long nch) {
  return (int)min((long long)kCPB, nch - (long long)blockIdx.x * kCPB);
}
template <class K>
__device__ __forceinline__ void chunk_keys(const K &keys, int nc, uint32_t *t) {
  const long long c0 = (long long)blockIdx.x * kCPB * kChunk;
  uint32_t v[kCPB * kGroups];
#pragma unroll
  for (int c = 0; c < kCPB; c++) {
    // the key function of this chunk (uniform over the block; a chunk past
    // the species' last takes the block's first, whose keys are not used)
    const auto kc = keys.at_chunk(c < nc ? c0 + (long long)c * kChunk : c0);
#pragma unroll
    for (int m = 0; m < kGroups; m++)
      v[c * kGroups + m] = c < nc ? kc(c0 + (long long)c * kChunk + m * 256 + (int)threadIdx.x) : 0u;
  }
#pragma unroll
  for (int c = 0; c < kCPB; c++)
#pragma unroll
    for (int m = 0; m < kGroups; m++) {
      const int w = m * 256 + (int)threadIdx.x;  // slot offset: iteration w / 64 of lane w % 64
      t[c * kChunkLds + (w % 64) * kRow + w / 64] = v[c * kGroups + m];
    }
  __syncthreads();
}
// logical offset s in the block's chunks: chunk s / kChunk, and in it lane
// (s % kChunk) / kPPT, iteration s % kPPT
__device__ __forceinline__ uint32_t chunk_key(const uint32_t *t, int s) {
  const int c = s / kChunk, l = s % kChunk;
  return t[c * kChunkLds + (l / kPPT) * kRow + l % kPPT];
}

// the runs of the block's groups of 256 logical keys: per lane the key and
// its run (head lane, length) in each group, and per group and wave the
// number of runs in wc[group * 4 + wave] (LDS, read after a barrier)
struct ChunkRuns {
  uint32_t key[kBlockGroups];
  uint64_t heads[kBlockGroups];
  int head[kBlockGroups], len[kBlockGroups];
};
__device__ __forceinline__ void chunk_runs(const uint32_t *t, ChunkRuns &cr, uint32_t *wc) {
#pragma unroll
  for (int m = 0; m < kBlockGroups; m++) {
    cr.key[m] = chunk_key(t, m * 256 + (int)threadIdx.x);
    cr.heads[m] = wave_runs(cr.key[m], true, cr.head[m], cr.len[m]);
    if (__lane_id() == 0) wc[m * 4 + (threadIdx.x >> 6)] = (uint32_t)__popcll(cr.heads[m]);
  }
  __syncthreads();
}
// the record index of this lane's run in group m (base: exclusive scan of the
// group counts at this group)
__device__ __forceinline__ uint32_t run_index(const ChunkRuns &cr, int m, uint32_t base, const uint32_t *wc) {
  const int wave = threadIdx.x >> 6;
  uint32_t r = base;
  for (int w = 0; w < wave; w++) r += wc[m * 4 + w];
  const int head = cr.head[m];
  const uint64_t below = head == 0 ? 0ull : (cr.heads[m] & (~0ull >> (64 - head)));
  return r + (uint32_t)__popcll(below);
}

// runs per group of 256 logical keys
template <class K>
__global__ void __launch_bounds__(256) k_run_count(K keys, long long nch, uint32_t *__restrict__ block_runs) {
  __shared__ uint32_t t[kCPB * kChunkLds];
  __shared__ uint32_t wc[kBlockGroups * 4];
  const int nc = block_chunks(nch);
  chunk_keys(keys, nc, t);
  ChunkRuns cr;
  chunk_runs(t, cr, wc);
  if (threadIdx.x < nc * kGroups) {
    const uint32_t *g = wc + threadIdx.x * 4;
    block_runs[(long long)blockIdx.x * kBlockGroups + threadIdx.x] = g[0] + g[1] + g[2] + g[3];
  }
}

// one record (key, index, length) per run
template <class K>
__global__ void __launch_bounds__(256) k_run_emit(K keys, long long nch, const uint32_t *__restrict__ block_base,
                                                  const uint32_t *count, long long cap, uint32_t *__restrict__ rkey,
                                                  uint32_t *__restrict__ ridx, uint32_t *__restrict__ rlen) {
  if (*count > cap) return;  // more runs than record slots: this sort is skipped
  __shared__ uint32_t t[kCPB * kChunkLds];
  __shared__ uint32_t wc[kBlockGroups * 4];
  const int nc = block_chunks(nch);
  chunk_keys(keys, nc, t);
  ChunkRuns cr;
  chunk_runs(t, cr, wc);
#pragma unroll
  for (int m = 0; m < kBlockGroups; m++) {
    if (m >= nc * kGroups) break;
    const uint32_t r = run_index(cr, m, block_base[(long long)blockIdx.x * kBlockGroups + m], wc);
    if (cr.head[m] == (int)__lane_id()) {
      rkey[r] = cr.key[m];
      ridx[r] = r;
      rlen[r] = (uint32_t)cr.len[m];
    }
  }
}

// lens[t] = length of the t-th run in key order; after the scan of lens,
// rfirst[run] = its first rank
__global__ void k_run_lens(const uint32_t *__restrict__ order, const uint32_t *__restrict__ rlen,
                           const uint32_t *count, long long cap, uint32_t *__restrict__ lens) {
  const long long m = present(count, cap);
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < cap) lens[t] = t < m ? rlen[order[t]] : 0u;
}
__global__ void k_run_first(const uint32_t *__restrict__ order, const uint32_t *__restrict__ lens,
                            const uint32_t *count, long long cap, uint32_t *__restrict__ rfirst) {
  const long long m = present(count, cap);
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < m) rfirst[order[t]] = lens[t];
}

// src[rank] = the slot of the particle of that rank, indexed by logical
// position (consecutive ranks are written by consecutive lanes); a skipped
// sort leaves every particle in its own slot
template <class K>
__global__ void __launch_bounds__(256) k_run_place(K keys, long long nch, const uint32_t *__restrict__ block_base,
                                                   const uint32_t *count, long long cap,
                                                   const uint32_t *__restrict__ rfirst, uint32_t *__restrict__ src) {
  const long long c0 = (long long)blockIdx.x * kCPB * kChunk;
  const int nc = block_chunks(nch);
  if (*count > cap) {
    for (int s = threadIdx.x; s < nc * kChunk; s += 256) src[c0 + s] = (uint32_t)phys_of(c0 + s);
    return;
  }
  __shared__ uint32_t t[kCPB * kChunkLds];
  __shared__ uint32_t wc[kBlockGroups * 4];
  chunk_keys(keys, nc, t);
  ChunkRuns cr;
  chunk_runs(t, cr, wc);
  const unsigned lane = __lane_id();
#pragma unroll
  for (int m = 0; m < kBlockGroups; m++) {
    if (m >= nc * kGroups) break;
    const uint32_t r = run_index(cr, m, block_base[(long long)blockIdx.x * kBlockGroups + m], wc);
    const int head = cr.head[m];
    const uint32_t first = head == (int)lane ? rfirst[r] : 0u;
    const uint32_t rank = __shfl(first, head) + (lane - (unsigned)head);
    src[rank] = (uint32_t)phys_of(c0 + m * 256 + (int)threadIdx.x);
  }
}

// every particle array in one pass, one chunk per block: slot p of the
// output takes the input slot src[logical position of p].  The block reads
// its chunk of src coalesced and transposes it to slot order in LDS (rows
// padded to 17 words); each thread then moves kGatherPer particles (256
// slots apart) so that all their loads are in flight together (non-temporal
// loads were slower: the slots of neighbouring particles share cache lines;
// the stores, each written once, are non-temporal: 5% off the sort).
struct Arrays7 {
  void *a[7];
};
constexpr int kGatherPer = kChunk / 256;
template <class T>
__device__ __forceinline__ void gather_array(const void *in, void *out, const uint32_t (&q)[kGatherPer],
                                             long long p0) {
  T x[kGatherPer];
#pragma unroll
  for (int u = 0; u < kGatherPer; u++) x[u] = static_cast<const T *>(in)[q[u]];
#pragma unroll
  for (int u = 0; u < kGatherPer; u++) __builtin_nontemporal_store(x[u], static_cast<T *>(out) + p0 + u * 256);
}
// the source slots of this thread's kGatherPer output slots c0 + u*256 + tid
__device__ __forceinline__ void gather_sources(const uint32_t *__restrict__ src, long long c0,
                                               uint32_t (&q)[kGatherPer]) {
  __shared__ uint32_t t[kChunkLds];
#pragma unroll
  for (int u = 0; u < kGatherPer; u++) {
    const int s = u * 256 + (int)threadIdx.x;  // logical offset
    t[(s / kPPT) * kRow + s % kPPT] = src[c0 + s];
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < kGatherPer; u++) {
    const int w = u * 256 + (int)threadIdx.x;  // slot offset: lane w % 64, iteration w / 64
    q[u] = t[(w % 64) * kRow + w / 64];
  }
}
__global__ void __launch_bounds__(256) k_gather_f32(Arrays7 in, Arrays7 out, const uint32_t *__restrict__ src) {
  const long long c0 = (long long)blockIdx.x * kChunk;
  uint32_t q[kGatherPer];
  gather_sources(src, c0, q);
#pragma unroll
  for (int k = 0; k < 7; k++) gather_array<uint32_t>(in.a[k], out.a[k], q, c0 + threadIdx.x);
}
__global__ void __launch_bounds__(256) k_gather_f64(Arrays7 in, Arrays7 out, const uint32_t *__restrict__ src) {
  const long long c0 = (long long)blockIdx.x * kChunk;
  uint32_t q[kGatherPer];
  gather_sources(src, c0, q);
#pragma unroll
  for (int k = 0; k < 6; k++) gather_array<double>(in.a[k], out.a[k], q, c0 + threadIdx.x);
  gather_array<uint8_t>(in.a[6], out.a[6], q, c0 + threadIdx.x);
}

// the multi-segment gather: one chunk per block, the chunk's segment gives
// the arrays, the permutation holds global slots of the same segment
__global__ void __launch_bounds__(256) k_gather_f32_multi(const MultiSeg *__restrict__ seg,
                                                          const uint32_t *__restrict__ seg_of_chunk,
                                                          const uint32_t *__restrict__ src) {
  const long long c0 = (long long)blockIdx.x * kChunk;
  const MultiSeg &d = seg[seg_of_chunk[blockIdx.x]];
  uint32_t q[kGatherPer];
  gather_sources(src, c0, q);
#pragma unroll
  for (int u = 0; u < kGatherPer; u++) q[u] -= (uint32_t)d.slot_base;
  const long long p0 = c0 - d.slot_base + threadIdx.x;
#pragma unroll
  for (int k = 0; k < 7; k++) gather_array<uint32_t>(d.in[k], d.out[k], q, p0);
}

// the sort key of the particle in slot j, from its state
struct KeysF32 {
  GridDesc g;
  const uint32_t *cell;
  uint32_t dead_key;
  __device__ __forceinline__ uint32_t operator()(long long j) const {
    const uint32_t c = cell[j];
    if (c == kPadCell) return dead_key + 1;
    if (c & 0x80000000u) return dead_key;
    return ((c >> 16) & 0x7FFF) * (uint32_t)(g.nz - 1) + (c & 0xFFFF);
  }
  __device__ __forceinline__ const KeysF32 &at_chunk(long long) const { return *this; }
};
struct KeysF64 {
  GridDesc g;
  const double *r, *z;
  const uint8_t *alive;
  uint32_t dead_key;
  __device__ __forceinline__ uint32_t operator()(long long j) const {
    if (alive[j] == kPadSlot) return dead_key + 1;
    if (alive[j] != 1) return dead_key;
    int i = (int)ceil(r[j] / g.dr) - 1, k = (int)ceil(z[j] / g.dz) - 1;
    i = max(0, min(i, g.nr - 2));
    k = max(0, min(k, g.nz - 2));
    return (uint32_t)i * (uint32_t)(g.nz - 1) + (uint32_t)k;
  }
  __device__ __forceinline__ const KeysF64 &at_chunk(long long) const { return *this; }
};

// several species sorted as one: segment s holds global slots
// [slot_base, slot_base + n_cap) (whole chunks) and keys
// [key_base, key_base + ncell + 2), so the sorted order keeps every
// particle inside its own segment
__device__ __forceinline__ int seg_of(const MultiSeg *seg, const uint32_t *seg_of_chunk, long long j) {
  return (int)seg_of_chunk[j / kChunk];
}
// the keys of one segment (global slot j): row-major cells, or bands of two
// rows ordered column by column (cell (i, k) -> (i / 2) 2 (nz-1) + 2 k + i % 2)
struct KeysSeg {
  const uint32_t *cell;  // shifted so that cell[j] is global slot j
  uint32_t key_base, ncell, nzm1, band;
  __device__ __forceinline__ uint32_t operator()(long long j) const {
    const uint32_t c = cell[j];
    if (c == kPadCell) return key_base + ncell + 1;
    if (c & 0x80000000u) return key_base + ncell;
    const uint32_t i = (c >> 16) & 0x7FFF, k = c & 0xFFFF;
    return key_base + (band ? (i >> 1) * (2 * nzm1) + 2 * k + (i & 1) : i * nzm1 + k);
  }
};
struct KeysMulti {
  const MultiSeg *seg;
  const uint32_t *seg_of_chunk;
  // the segment of the chunk starting at global slot j0
  __device__ __forceinline__ KeysSeg at_chunk(long long j0) const {
    const int s = __builtin_amdgcn_readfirstlane(seg_of(seg, seg_of_chunk, j0));
    const MultiSeg &d = seg[s];
    return KeysSeg{d.cell - d.slot_base, d.key_base, d.ncell, d.nzm1, d.band};
  }
  __device__ __forceinline__ uint32_t operator()(long long j) const { return at_chunk(j - j % kChunk)(j); }
};

void exclusive_scan(uint32_t *data, long long n, uint32_t *tmp, hipStream_t s) {
  long long nb = (n + kScanBlock - 1) / kScanBlock;
  hipLaunchKernelGGL(k_scan_local, dim3((unsigned)nb), dim3(kScanBlock), 0, s, data, n, tmp);
  if (nb > 1) {
    exclusive_scan(tmp, nb, tmp + nb, s);
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)nb), dim3(kScanBlock), 0, s, data, n, tmp);
  }
}
// scratch words exclusive_scan needs for n elements
long long scan_scratch(long long n) {
  long long t = 0;
  while (n > 1) {
    n = (n + kScanBlock - 1) / kScanBlock;
    t += n;
  }
  return t + 1;
}

// stable LSD radix sort of m (key, val) pairs by the low `bits` bits of the
// key; returns the buffers holding the result (a or b)
void radix_pairs(uint32_t *&ka, uint32_t *&va, uint32_t *&kb, uint32_t *&vb, const uint32_t *count, long long cap,
                 int bits, uint32_t *hist, uint32_t *tmp, hipStream_t st) {
  const int passes = (bits + 7) / 8;
  const int nblocks = (int)((cap + kPerBlock - 1) / kPerBlock);
  const long long hist_n = (long long)kBins * nblocks;
  for (int p = 0; p < passes; p++) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_hist, dim3(nblocks), dim3(kThreads), 0, st, ka, count, cap, shift, hist, nblocks);
    exclusive_scan(hist, hist_n, tmp, st);
    hipLaunchKernelGGL(k_scatter, dim3(nblocks), dim3(kThreads), 0, st, ka, va, kb, vb, count, cap, shift, hist,
                       nblocks);
    std::swap(ka, kb);
    std::swap(va, vb);
  }
}

bool ensure_alt(void *&alt, size_t bytes) {
  if (alt) return true;
  if (hipMalloc(&alt, bytes) != hipSuccess) {
    alt = nullptr;
    return false;
  }
  return true;
}

// workspace bytes for n slots and cap run records (u32 words): src (n); group
// counts (nb + 1); scratch for the scans of the group counts and run lengths
// and for the records' digit histograms with their scan; the records (5 per
// slot)
struct SortWs {
  long long nb, rec_hist, scratch;
  size_t bytes;
};
SortWs sort_ws(long long n, long long cap) {
  SortWs w;
  w.nb = (n + 255) / 256;
  w.rec_hist = (long long)kBins * ((cap + kPerBlock - 1) / kPerBlock);
  w.scratch = std::max(std::max(scan_scratch(w.nb + 1), scan_scratch(cap)), w.rec_hist + scan_scratch(w.rec_hist)) +
              1024;
  w.bytes = (size_t)(n + (w.nb + 1) + w.scratch + 5 * cap) * sizeof(uint32_t);
  return w;
}
bool ensure_ws(Ctx &c, size_t need) {
  if (c.sort_ws_bytes >= need) return true;
  if (c.sort_ws) (void)hipFree(c.sort_ws);
  c.sort_ws = nullptr;
  c.sort_ws_bytes = 0;
  if (hipMalloc(&c.sort_ws, need) != hipSuccess) return false;
  c.sort_ws_bytes = need;
  return true;
}
// the pinned run count and its event
bool ensure_runs(Species &sp) {
  if (sp.runs_host) return true;
  if (hipHostMalloc((void **)&sp.runs_host, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
    sp.runs_host = nullptr;
    return false;
  }
  *sp.runs_host = 0;
  if (hipEventCreateWithFlags(&sp.runs_ev, hipEventDisableTiming) != hipSuccess) {
    sp.runs_ev = nullptr;
    return false;
  }
  return true;
}
// the second buffer of every particle array
bool ensure_alts(Ctx &c, Species &sp) {
  const long long n = sp.n_cap;
  if (c.precision == 1) {
    SpeciesF32 &b = sp.f32_alt;
    void **bp[7] = {(void **)&b.cell, (void **)&b.fr, (void **)&b.fz, (void **)&b.vr, (void **)&b.vphi,
                    (void **)&b.vz, (void **)&b.w};
    for (int k = 0; k < 7; k++)
      if (!ensure_alt(*bp[k], (size_t)n * 4)) return false;
  } else {
    SpeciesF64 &b = sp.f64_alt;
    void **bp[7] = {(void **)&b.r, (void **)&b.z, (void **)&b.vr, (void **)&b.vphi, (void **)&b.vz, (void **)&b.w,
                    (void **)&b.alive};
    for (int k = 0; k < 7; k++)
      if (!ensure_alt(*bp[k], (size_t)n * (k < 6 ? 8 : 1))) return false;
  }
  return true;
}
bool sortable(const Ctx &c, const Species &sp) {
  return sp.n_cap > 1 && sp.chunked && !(c.precision == 1 && !sp.f32.cell);
}

}  // namespace

size_t sort_periodic_bytes(const Ctx &c) {
  size_t need = 0;
  for (const Species &sp : c.species) {
    if (sp.kind != 0 || !sortable(c, sp)) continue;
    const long long cap = sp.rec_cap > 0 ? sp.rec_cap : std::min(sp.n_cap, sp.n_cap / 4 + 65536);
    need = std::max(need, sort_ws(sp.n_cap, cap).bytes);
  }
  return need;
}

int sort_shrink_ws(Ctx &c, size_t need) {
  if (c.sort_ws_bytes <= need + need / 2) return 0;
  if (hipStreamSynchronize(c.stream) != hipSuccess) return -3;
  (void)hipFree(c.sort_ws);
  c.sort_ws = nullptr;
  c.sort_ws_bytes = 0;
  return need == 0 || ensure_ws(c, need) ? 0 : -3;
}

int sort_prepare(Ctx &c, Species &sp) {
  if (sp.sort_prepared || !sortable(c, sp)) return 0;
  const long long n = sp.n_cap;
  if (sp.rec_cap <= 0) sp.rec_cap = std::min(n, n / 4 + 65536);
  if (!ensure_runs(sp) || !ensure_ws(c, sort_ws(n, sp.rec_cap).bytes) || !ensure_alts(c, sp)) return -3;
  sp.sort_prepared = true;
  return 0;
}

int sort_species(Ctx &c, Species &sp, bool exact) {
  const long long n = sp.n_cap;
  if (!sortable(c, sp)) return 0;
  if (n % kChunk != 0) return -1;  // a chunked species fills whole chunks
  const uint32_t ncell = (uint32_t)(c.g.nr - 1) * (uint32_t)(c.g.nz - 1);
  const uint32_t dead_key = ncell;
  int bits = 1;
  while ((1ull << bits) <= (unsigned long long)dead_key + 1) bits++;
  // record capacity: grown from the previous sort's run count (read back
  // asynchronously), so that no sort waits for the device
  const bool first = !sp.runs_host;
  if (!ensure_runs(sp)) return -3;
  if (!first && hipEventSynchronize(sp.runs_ev) != hipSuccess) return -3;
  const long long last = *sp.runs_host;
  long long cap = sp.rec_cap > 0 ? sp.rec_cap : std::min(n, n / 4 + 65536);
  if (4 * last > 3 * cap) cap = std::min(n, 2 * last + 65536);
  sp.rec_cap = cap;
  if (exact) cap = n;  // room for one run per particle: this sort cannot be skipped
  const SortWs ws = sort_ws(n, cap);
  const long long nb = ws.nb, rec_hist = ws.rec_hist, scratch = ws.scratch;
  if (!ensure_ws(c, ws.bytes) || !ensure_alts(c, sp)) return -3;
  uint32_t *src = (uint32_t *)c.sort_ws, *block = src + n, *tmp = block + nb + 1;
  uint32_t *rk = tmp + scratch, *ri = rk + cap, *rk2 = ri + cap, *ri2 = rk2 + cap, *rlen = ri2 + cap;
  const uint32_t *count = block + nb;  // the number of runs, after the scan
  const KeysF32 k32{c.g, (const uint32_t *)sp.f32.cell, dead_key};
  const KeysF64 k64{c.g, sp.f64.r, sp.f64.z, sp.f64.alive, dead_key};
  const bool f32 = c.precision == 1;
  // runs per group of 256 logical keys and their scan (block[nb] = the
  // number of runs); n is a whole number of chunks, one per block
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, c.stream, block + nb, 1LL, 0u);
  const long long nch = n / kChunk;
  const unsigned chunks = (unsigned)nch, rblocks = (unsigned)((nch + kCPB - 1) / kCPB);
  if (f32)
    hipLaunchKernelGGL(k_run_count<KeysF32>, dim3(rblocks), dim3(256), 0, c.stream, k32, nch, block);
  else
    hipLaunchKernelGGL(k_run_count<KeysF64>, dim3(rblocks), dim3(256), 0, c.stream, k64, nch, block);
  exclusive_scan(block, nb + 1, tmp, c.stream);
  if (f32)
    hipLaunchKernelGGL(k_run_emit<KeysF32>, dim3(rblocks), dim3(256), 0, c.stream, k32, nch, block, count, cap, rk,
                       ri, rlen);
  else
    hipLaunchKernelGGL(k_run_emit<KeysF64>, dim3(rblocks), dim3(256), 0, c.stream, k64, nch, block, count, cap, rk,
                       ri, rlen);
  // records by key (stable), then each run's first rank
  radix_pairs(rk, ri, rk2, ri2, count, cap, bits, tmp, tmp + rec_hist, c.stream);
  const unsigned mb = (unsigned)((cap + 255) / 256);
  uint32_t *lens = rk2;  // free after the sort
  hipLaunchKernelGGL(k_run_lens, dim3(mb), dim3(256), 0, c.stream, ri, rlen, count, cap, lens);
  exclusive_scan(lens, cap, tmp, c.stream);
  hipLaunchKernelGGL(k_run_first, dim3(mb), dim3(256), 0, c.stream, ri, lens, count, cap, rlen);
  if (f32)
    hipLaunchKernelGGL(k_run_place<KeysF32>, dim3(rblocks), dim3(256), 0, c.stream, k32, nch, block, count, cap, rlen,
                       src);
  else
    hipLaunchKernelGGL(k_run_place<KeysF64>, dim3(rblocks), dim3(256), 0, c.stream, k64, nch, block, count, cap, rlen,
                       src);
  // the run count for the next sort's capacity
  if (hipMemcpyAsync(sp.runs_host, count, sizeof(uint32_t), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
      hipEventRecord(sp.runs_ev, c.stream) != hipSuccess)
    return -3;
  // all arrays in one gather
  Arrays7 in{}, out{};
  if (c.precision == 1) {
    SpeciesF32 &a = sp.f32, &b = sp.f32_alt;
    void **ap[7] = {(void **)&a.cell, (void **)&a.fr, (void **)&a.fz, (void **)&a.vr, (void **)&a.vphi,
                    (void **)&a.vz, (void **)&a.w};
    void **bp[7] = {(void **)&b.cell, (void **)&b.fr, (void **)&b.fz, (void **)&b.vr, (void **)&b.vphi,
                    (void **)&b.vz, (void **)&b.w};
    for (int k = 0; k < 7; k++) {
      in.a[k] = *ap[k];
      out.a[k] = *bp[k];
    }
    hipLaunchKernelGGL(k_gather_f32, dim3(chunks), dim3(256), 0, c.stream, in, out, src);
    for (int k = 0; k < 7; k++) std::swap(*ap[k], *bp[k]);
  } else {
    SpeciesF64 &a = sp.f64, &b = sp.f64_alt;
    void **ap[7] = {(void **)&a.r, (void **)&a.z, (void **)&a.vr, (void **)&a.vphi, (void **)&a.vz, (void **)&a.w,
                    (void **)&a.alive};
    void **bp[7] = {(void **)&b.r, (void **)&b.z, (void **)&b.vr, (void **)&b.vphi, (void **)&b.vz, (void **)&b.w,
                    (void **)&b.alive};
    for (int k = 0; k < 7; k++) {
      in.a[k] = *ap[k];
      out.a[k] = *bp[k];
    }
    hipLaunchKernelGGL(k_gather_f64, dim3(chunks), dim3(256), 0, c.stream, in, out, src);
    for (int k = 0; k < 7; k++) std::swap(*ap[k], *bp[k]);
  }
  return 0;
}


void sort_multi_free(MultiSort &ms) {
  if (ms.seg) (void)hipFree(ms.seg);
  if (ms.seg_of_chunk) (void)hipFree(ms.seg_of_chunk);
  if (ms.ws) (void)hipFree(ms.ws);
  if (ms.runs_host) (void)hipHostFree(ms.runs_host);
  if (ms.runs_ev) (void)hipEventDestroy(ms.runs_ev);
  ms = MultiSort();
}

int sort_multi(MultiSort &ms, const std::vector<Ctx *> &cs, hipStream_t stream) {
  // the segments: every sortable plasma species of every context, in order
  std::vector<std::pair<Ctx *, Species *>> list;
  for (Ctx *c : cs)
    for (Species &sp : c->species)
      if (sp.kind == 0 && sortable(*c, sp) && c->precision == 1) list.emplace_back(c, &sp);
  if (list.empty()) return 0;
  std::vector<MultiSeg> segs(list.size());
  long long n = 0, keys = 0;
  for (size_t s = 0; s < list.size(); s++) {
    Ctx &c = *list[s].first;
    Species &sp = *list[s].second;
    if (sp.n_cap % kChunk != 0) return -1;
    if (!ensure_alts(c, sp)) return -3;
    MultiSeg &d = segs[s];
    d.cell = (const uint32_t *)sp.f32.cell;
    d.slot_base = n;
    d.key_base = (uint32_t)keys;
    d.nzm1 = (uint32_t)(c.g.nz - 1);
    // A particle tile (a few chunks of consecutive particles) spans
    // tile / ppc cells; in row order that overruns the tile's 64-column field
    // box below about 32 particles per cell, in bands of two rows only below
    // about 16.  Instances with fewer particles per cell sort in bands.
    const long long rows = c.g.nr - 1;
    d.band = sp.n < 32LL * rows * (long long)d.nzm1 ? 1u : 0u;
    d.ncell = d.band ? (uint32_t)(2 * ((rows + 1) / 2)) * d.nzm1 : (uint32_t)rows * d.nzm1;
    SpeciesF32 &a = sp.f32, &b = sp.f32_alt;
    void *ap[7] = {a.cell, a.fr, a.fz, a.vr, a.vphi, a.vz, a.w};
    void *bp[7] = {b.cell, b.fr, b.fz, b.vr, b.vphi, b.vz, b.w};
    for (int k = 0; k < 7; k++) {
      d.in[k] = ap[k];
      d.out[k] = bp[k];
    }
    n += sp.n_cap;
    keys += (long long)d.ncell + 2;
  }
  if (n >= (1ll << 32) || keys >= (1ll << 32)) return -1;  // slots and keys are 32-bit
  const long long nch = n / kChunk;
  // the tables (rebuilt only when the segments change size)
  if (ms.n != n || ms.nseg != (long long)list.size()) {
    if (ms.seg) (void)hipFree(ms.seg);
    if (ms.seg_of_chunk) (void)hipFree(ms.seg_of_chunk);
    ms.seg = nullptr;
    ms.seg_of_chunk = nullptr;
    if (hipMalloc((void **)&ms.seg, list.size() * sizeof(MultiSeg)) != hipSuccess ||
        hipMalloc((void **)&ms.seg_of_chunk, (size_t)nch * sizeof(uint32_t)) != hipSuccess)
      return -3;
    std::vector<uint32_t> soc((size_t)nch);
    for (size_t s = 0; s < list.size(); s++) {
      const long long c0 = segs[s].slot_base / kChunk, c1 = c0 + list[s].second->n_cap / kChunk;
      for (long long ch = c0; ch < c1; ch++) soc[(size_t)ch] = (uint32_t)s;
    }
    if (hipMemcpy(ms.seg_of_chunk, soc.data(), soc.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess)
      return -3;
    ms.n = n;
    ms.nseg = (long long)list.size();
    ms.rec_cap = 0;
    ms.counted = false;
  }
  // the segment table changes with every sort (the arrays swap); the
  // previous copy from seg_host has completed (the last sort's run-count
  // event, waited for below or above, comes after it on the stream)
  if (ms.counted && hipEventSynchronize(ms.runs_ev) != hipSuccess) return -3;
  ms.seg_host = segs;
  if (hipMemcpyAsync(ms.seg, ms.seg_host.data(), segs.size() * sizeof(MultiSeg), hipMemcpyHostToDevice, stream) !=
      hipSuccess)
    return -3;
  int bits = 1;
  while ((1ull << bits) <= (unsigned long long)keys) bits++;
  // record capacity from the previous sort's run count (read back
  // asynchronously), as sort_species does
  if (!ms.runs_host) {
    if (hipHostMalloc((void **)&ms.runs_host, sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) return -3;
    *ms.runs_host = 0;
    if (hipEventCreateWithFlags(&ms.runs_ev, hipEventDisableTiming) != hipSuccess) return -3;
  }
  const long long last = ms.counted ? (long long)*ms.runs_host : 0;
  long long cap = ms.rec_cap > 0 ? ms.rec_cap : std::min(n, n / 4 + 65536);
  if (4 * last > 3 * cap) cap = std::min(n, 2 * last + 65536);
  ms.rec_cap = cap;
  const SortWs ws = sort_ws(n, cap);
  if (ms.ws_bytes < ws.bytes) {
    if (ms.ws) (void)hipFree(ms.ws);
    ms.ws = nullptr;
    ms.ws_bytes = 0;
    if (hipMalloc(&ms.ws, ws.bytes) != hipSuccess) return -3;
    ms.ws_bytes = ws.bytes;
  }
  const long long nb = ws.nb, rec_hist = ws.rec_hist, scratch = ws.scratch;
  uint32_t *src = (uint32_t *)ms.ws, *block = src + n, *tmp = block + nb + 1;
  uint32_t *rk = tmp + scratch, *ri = rk + cap, *rk2 = ri + cap, *ri2 = rk2 + cap, *rlen = ri2 + cap;
  const uint32_t *count = block + nb;
  const KeysMulti km{ms.seg, ms.seg_of_chunk};
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(1), 0, stream, block + nb, 1LL, 0u);
  const unsigned chunks = (unsigned)nch, rblocks = (unsigned)((nch + kCPB - 1) / kCPB);
  hipLaunchKernelGGL(k_run_count<KeysMulti>, dim3(rblocks), dim3(256), 0, stream, km, nch, block);
  exclusive_scan(block, nb + 1, tmp, stream);
  hipLaunchKernelGGL(k_run_emit<KeysMulti>, dim3(rblocks), dim3(256), 0, stream, km, nch, block, count, cap, rk, ri,
                     rlen);
  radix_pairs(rk, ri, rk2, ri2, count, cap, bits, tmp, tmp + rec_hist, stream);
  const unsigned mb = (unsigned)((cap + 255) / 256);
  uint32_t *lens = rk2;
  hipLaunchKernelGGL(k_run_lens, dim3(mb), dim3(256), 0, stream, ri, rlen, count, cap, lens);
  exclusive_scan(lens, cap, tmp, stream);
  hipLaunchKernelGGL(k_run_first, dim3(mb), dim3(256), 0, stream, ri, lens, count, cap, rlen);
  hipLaunchKernelGGL(k_run_place<KeysMulti>, dim3(rblocks), dim3(256), 0, stream, km, nch, block, count, cap, rlen,
                     src);
  if (hipMemcpyAsync(ms.runs_host, count, sizeof(uint32_t), hipMemcpyDeviceToHost, stream) != hipSuccess ||
      hipEventRecord(ms.runs_ev, stream) != hipSuccess)
    return -3;
  ms.counted = true;
  hipLaunchKernelGGL(k_gather_f32_multi, dim3(chunks), dim3(256), 0, stream, ms.seg, ms.seg_of_chunk, src);
  if (hipGetLastError() != hipSuccess) return -3;
  for (auto &e : list) {
    SpeciesF32 &a = e.second->f32, &b = e.second->f32_alt;
    void **ap[7] = {(void **)&a.cell, (void **)&a.fr, (void **)&a.fz, (void **)&a.vr, (void **)&a.vphi,
                    (void **)&a.vz, (void **)&a.w};
    void **bp[7] = {(void **)&b.cell, (void **)&b.fr, (void **)&b.fz, (void **)&b.vr, (void **)&b.vphi,
                    (void **)&b.vz, (void **)&b.w};
    for (int k = 0; k < 7; k++) std::swap(*ap[k], *bp[k]);
  }
  return 0;
}

}  // namespace pdp3
