// Where the particle kernels put their grid contributions (J_r, J_phi, J_z
// and the bunch rho, all FP64).
//
// GlobalSink: bounds-checked hardware FP64 atomics straight to HBM (used by
// the split entry points and by bunches, whose particles are not sorted).
//
// TileSink: three levels, for cell-sorted particles walked by one lane in
// runs of the same cell:
//   1. registers: the contributions of the current cell (ic, kc) to its
//      J_phi / rho corners (ic..ic+1, kc..kc+1), J_z nodes (ic..ic+1, kc) and
//      J_r nodes (ic, kc..kc+1) are summed in FP64 registers and flushed only
//      when the lane moves to a particle in another cell;
//   2. an LDS box of kBoxR x kBoxK nodes per component anchored at the
//      block's first particle, filled with LDS FP64 atomics;
//   3. nodes outside the box go to GlobalSink.
// After the block the box is added to HBM once (one atomic per touched node).
#pragma once

#include "context.hpp"

namespace pdp3 {

enum SinkComp { kCR = 0, kCPhi = 1, kCZ = 2, kCRho = 3 };

// j: J_r of the context's grid, with J_phi and J_z the next nr*nz arrays
// (one base pointer instead of three keeps the particle kernels' scalar
// registers from spilling); rho: the bunch charge density, or a caller's array
struct GlobalSink {
  int nr, nz;
  double *j, *rho;
  int *err;
  __device__ __forceinline__ double *base(int comp) const {
    return comp == kCRho ? rho : j + (size_t)comp * ((size_t)nr * (size_t)nz);
  }
  __device__ __forceinline__ void put(int comp, int i, int k, double v) const {
    if ((unsigned)i < (unsigned)nr && (unsigned)k < (unsigned)nz)
      unsafeAtomicAdd(base(comp) + (size_t)i * (size_t)nz + (size_t)k, v);
    else
      atomicOr(err, kErrDomain);
  }
  template <int C>
  __device__ __forceinline__ void add(int i, int k, double v) {
    put(C, i, k, v);
  }
  // a pair of nodes: (i, k), (i+1, k) for J_z; (i, k), (i, k+1) for J_r
  template <int C>
  __device__ __forceinline__ void add2(int i, int k, double v0, double v1) {
    put(C, i, k, v0);
    if (C == kCZ)
      put(C, i + 1, k, v1);
    else
      put(C, i, k + 1, v1);
  }
  // the four corners (i..i+1, k..k+1) of J_phi / rho
  template <int C>
  __device__ __forceinline__ void add4(int i, int k, double v00, double v10, double v01, double v11) {
    put(C, i, k, v00);
    put(C, i + 1, k, v10);
    put(C, i, k + 1, v01);
    put(C, i + 1, k + 1, v11);
  }
  __device__ __forceinline__ void begin_cell(int, int) {}
  __device__ __forceinline__ void flush() {}
};

constexpr int kBoxComps = 4;

// BR x BK nodes per component; the plasma box spans 4 rows (i0-1 .. i0+2) by
// 128 columns (a sorted block covers a run of cells along one row), the bunch
// box 16 rows by 32 columns (a bunch block is a thin disc of particles
// injected in the same step)
template <bool RHO, int BR = 4, int BK = 128>
struct TileSinkT {
  static constexpr int kR = BR, kK = BK, kSize = kBoxComps * BR * BK;
  GlobalSink g;
  double *box;  // LDS, kBoxComps x kBoxR x kBoxK
  int bi, bk;   // node coordinates of box[.][0][0]
  int ic, kc;   // cell of the register accumulator
  // scalar accumulators (no arrays: a runtime index would push them to scratch)
  double p00, p10, p01, p11;  // J_phi corners (ic+di, kc+dk)
  double r00, r10, r01, r11;  // rho corners (bunches only)
  double z0, z1;              // J_z at (ic, kc), (ic+1, kc)
  double x0, x1;              // J_r at (ic, kc), (ic, kc+1)

  __device__ __forceinline__ void box_put(int comp, int i, int k, double v) {
    int li = i - bi, lk = k - bk;
    if ((unsigned)li < (unsigned)BR && (unsigned)lk < (unsigned)BK)
      atomicAdd(box + (comp * BR + li) * BK + lk, v);
    else
      g.put(comp, i, k, v);
  }
  template <int C>
  __device__ __forceinline__ void add(int i, int k, double v) {
    const int di = i - ic, dk = k - kc;
    if (C == kCPhi || (C == kCRho && RHO)) {
      if ((unsigned)di < 2u && (unsigned)dk < 2u) {
        const double a00 = (di == 0 && dk == 0) ? v : 0.0, a10 = (di == 1 && dk == 0) ? v : 0.0;
        const double a01 = (di == 0 && dk == 1) ? v : 0.0, a11 = (di == 1 && dk == 1) ? v : 0.0;
        if (C == kCPhi) {
          p00 += a00; p10 += a10; p01 += a01; p11 += a11;
        } else {
          r00 += a00; r10 += a10; r01 += a01; r11 += a11;
        }
        return;
      }
    } else if (C == kCZ) {
      if (dk == 0 && (unsigned)di < 2u) {
        z0 += di == 0 ? v : 0.0;
        z1 += di == 1 ? v : 0.0;
        return;
      }
    } else if (C == kCR) {
      if (di == 0 && (unsigned)dk < 2u) {
        x0 += dk == 0 ? v : 0.0;
        x1 += dk == 1 ? v : 0.0;
        return;
      }
    }
    box_put(C, i, k, v);
  }
  // whole stencils: straight into the accumulators when they are centred on
  // the accumulator's cell (the common case: no select chains), else the box
  template <int C>
  __device__ __forceinline__ void add2(int i, int k, double v0, double v1) {
    if (i == ic && k == kc) {
      if (C == kCZ) {
        z0 += v0;
        z1 += v1;
      } else {
        x0 += v0;
        x1 += v1;
      }
      return;
    }
    box_put(C, i, k, v0);
    if (C == kCZ)
      box_put(C, i + 1, k, v1);
    else
      box_put(C, i, k + 1, v1);
  }
  template <int C>
  __device__ __forceinline__ void add4(int i, int k, double v00, double v10, double v01, double v11) {
    if (i == ic && k == kc && (C == kCPhi || RHO)) {
      if (C == kCPhi) {
        p00 += v00; p10 += v10; p01 += v01; p11 += v11;
      } else {
        r00 += v00; r10 += v10; r01 += v01; r11 += v11;
      }
      return;
    }
    box_put(C, i, k, v00);
    box_put(C, i + 1, k, v10);
    box_put(C, i, k + 1, v01);
    box_put(C, i + 1, k + 1, v11);
  }
  __device__ __forceinline__ void put_nz(int comp, int i, int k, double v) {
    if (v != 0.0) g.put(comp, i, k, v);
  }
  // the accumulators' stencil (nodes ic..ic+1, kc..kc+1) in one piece: when
  // it lies inside the box, one bounds check and LDS atomics at fixed offsets
  // from one address (zeros included: cheaper than testing each value);
  // otherwise node by node to HBM
  __device__ __forceinline__ void flush() {
    if (ic < 0) return;
    const int li = ic - bi, lk = kc - bk;
    if ((unsigned)li < (unsigned)(BR - 1) && (unsigned)lk < (unsigned)(BK - 1)) {
      double *q = box + li * BK + lk;
      constexpr int C = BR * BK;
      atomicAdd(q + kCPhi * C, p00);
      atomicAdd(q + kCPhi * C + BK, p10);
      atomicAdd(q + kCPhi * C + 1, p01);
      atomicAdd(q + kCPhi * C + BK + 1, p11);
      if (RHO) {
        atomicAdd(q + kCRho * C, r00);
        atomicAdd(q + kCRho * C + BK, r10);
        atomicAdd(q + kCRho * C + 1, r01);
        atomicAdd(q + kCRho * C + BK + 1, r11);
      }
      atomicAdd(q + kCZ * C, z0);
      atomicAdd(q + kCZ * C + BK, z1);
      atomicAdd(q + kCR * C, x0);
      atomicAdd(q + kCR * C + 1, x1);
    } else {
      put_nz(kCPhi, ic, kc, p00);
      put_nz(kCPhi, ic + 1, kc, p10);
      put_nz(kCPhi, ic, kc + 1, p01);
      put_nz(kCPhi, ic + 1, kc + 1, p11);
      if (RHO) {
        put_nz(kCRho, ic, kc, r00);
        put_nz(kCRho, ic + 1, kc, r10);
        put_nz(kCRho, ic, kc + 1, r01);
        put_nz(kCRho, ic + 1, kc + 1, r11);
      }
      put_nz(kCZ, ic, kc, z0);
      put_nz(kCZ, ic + 1, kc, z1);
      put_nz(kCR, ic, kc, x0);
      put_nz(kCR, ic, kc + 1, x1);
    }
    p00 = p10 = p01 = p11 = 0.0;
    r00 = r10 = r01 = r11 = 0.0;
    z0 = z1 = x0 = x1 = 0.0;
  }
  // the next particle's contributions are centred on cell (i, k)
  __device__ __forceinline__ void begin_cell(int i, int k) {
    if (i != ic || k != kc) {
      flush();
      ic = i;
      kc = k;
    }
  }
};
using TileSink = TileSinkT<false>;
using BunchSink = TileSinkT<true, 16, 32>;

// The anchor cell of a tile of cell-sorted particles, from the cells (i, k)
// of its logical particles 0, 16 and 32 (live: a living particle; uniform
// values, so that the choice runs on the scalar unit): the median row of the
// living samples (the smaller when two live) and the smallest column among
// the samples in that row; (0, 0) when none lives.  Between sorts a few
// particles have moved to a neighbouring cell; a box anchored at one of them
// that moved to another row would serve the tile's own row only at its edge,
// so one such stray among the three samples is outvoted.  A sample one
// column off is harmless: the boxes keep two columns below the anchor.
__device__ __forceinline__ void tile_anchor(const bool (&live)[3], const int (&i)[3], const int (&k)[3], int &ai,
                                            int &ak) {
  constexpr int kNone = 0x7fffffff;
  int r[3];
#pragma unroll
  for (int j = 0; j < 3; j++) r[j] = live[j] ? i[j] : kNone;
  const int n = (int)live[0] + (int)live[1] + (int)live[2];
  if (n == 0) {
    ai = ak = 0;
    return;
  }
  const int lo = min(min(r[0], r[1]), r[2]);
  ai = lo;
  if (n == 3) ai = r[0] + r[1] + r[2] - lo - max(max(r[0], r[1]), r[2]);  // the median (rows are small)
  int col = kNone;
#pragma unroll
  for (int j = 0; j < 3; j++)
    if (live[j] && i[j] == ai) col = min(col, k[j]);
  ak = col;
}

// block prologue: anchor the box one row and two columns below the given
// cell (the tile's anchor, or a bunch block's smallest cell) and clear it
template <class T>
__device__ __forceinline__ void box_init(double *box, int anchor_i, int anchor_k, T &s, const GlobalSink &g) {
  for (int t = threadIdx.x; t < T::kSize; t += blockDim.x) box[t] = 0.0;
  s.g = g;
  s.box = box;
  s.bi = anchor_i - 1;
  s.bk = anchor_k - 2;
  s.ic = -1;
  s.kc = -1;
  s.p00 = s.p10 = s.p01 = s.p11 = 0.0;
  s.r00 = s.r10 = s.r01 = s.r11 = 0.0;
  s.z0 = s.z1 = s.x0 = s.x1 = 0.0;
}

// block epilogue: add the box to the grid (after __syncthreads)
template <class T>
__device__ __forceinline__ void box_flush(const double *box, const T &s) {
  for (int t = threadIdx.x; t < T::kSize; t += blockDim.x) {
    double v = box[t];
    if (v == 0.0) continue;
    int comp = t / (T::kR * T::kK);
    int rem = t - comp * T::kR * T::kK;
    int li = rem / T::kK, lk = rem - li * T::kK;
    s.g.put(comp, s.bi + li, s.bk + lk, v);
  }
}

}  // namespace pdp3

namespace pdp3 {
// smallest (i, k) over the block of two per-thread candidates (INT_MAX when
// the thread saw no living particle); every thread gets the result
__device__ __forceinline__ void block_min2(int &i, int &k) {
  __shared__ int mi[8], mk[8];
  for (int off = 32; off > 0; off >>= 1) {
    i = min(i, __shfl_xor(i, off));
    k = min(k, __shfl_xor(k, off));
  }
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    mi[wave] = i;
    mk[wave] = k;
  }
  __syncthreads();
  i = mi[0];
  k = mk[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
    i = min(i, mi[w]);
    k = min(k, mk[w]);
  }
}
}  // namespace pdp3
