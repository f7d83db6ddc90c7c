// Test-only probe: evaluates device libm functions and the library's
// correctly rounded reciprocal square root on a batch of inputs, so the
// parity tests can measure how often the GPU differs from the host's glibc.
#include <hip/hip_runtime.h>

#include "../../pdp3_amd/csrc/crmath.hpp"

namespace pdp3 {
// double-double helpers and a correctly rounded log, used here only to show
// how far the device log is from the correctly rounded value
struct DD {
  double hi, lo;
};
__device__ __forceinline__ DD dd_add(DD a, DD b) {
  double s, e;
  two_sum(a.hi, b.hi, s, e);
  e += a.lo + b.lo;
  double h = s + e;
  return DD{h, e - (h - s)};
}
__device__ __forceinline__ DD dd_mul(DD a, DD b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e += a.hi * b.lo + a.lo * b.hi;
  double h = p + e;
  return DD{h, e - (h - p)};
}
__device__ __forceinline__ DD dd_div(DD a, DD b) {
  double q1 = a.hi / b.hi;
  DD r = dd_add(a, DD{-(q1 * b.hi), -fma(q1, b.hi, -(q1 * b.hi))});
  r = dd_add(r, DD{-(q1 * b.lo), 0.0});
  double q2 = r.hi / b.hi;
  double h = q1 + q2;
  return DD{h, q2 - (h - q1)};
}

// log(x) for finite x > 0: x = m 2^e with m in [sqrt(1/2), sqrt(2)),
// log m = 2 atanh(s), s = (m-1)/(m+1), summed in double-double, plus e ln 2
__device__ inline double log_cr(double x) {
  int e;
  double m = frexp(x, &e);  // m in [0.5, 1)
  if (m < 0.70710678118654752) {
    m *= 2.0;
    e -= 1;
  }
  DD num{m - 1.0, 0.0};  // exact
  double dh, dl;
  two_sum(m, 1.0, dh, dl);
  DD s = dd_div(num, DD{dh, dl});
  DD s2 = dd_mul(s, s);
  // atanh(s)/s = sum_j s^(2j) / (2j+1); |s| <= 0.1716, 24 terms reach 2^-120
  DD acc{1.0 / 49.0, 0.0};
  for (int j = 23; j >= 0; j--) {
    acc = dd_mul(acc, s2);
    double c = 1.0 / (2 * j + 1);
    DD cj{c, fma(-c, (double)(2 * j + 1), 1.0) / (2 * j + 1)};
    acc = dd_add(acc, cj);
  }
  DD lm = dd_mul(dd_mul(acc, s), DD{2.0, 0.0});
  const DD ln2{6.93147180559945286227e-01, 2.31904681384629955842e-17};
  DD res = dd_add(dd_mul(ln2, DD{(double)e, 0.0}), lm);
  return res.hi + res.lo;
}

}  // namespace pdp3

__global__ void k_probe(const double *x, int n, double *out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = x[i];
  out[5 * i + 0] = pow(v, -0.5);
  out[5 * i + 1] = log(v);
  out[5 * i + 2] = sqrt(v);
  out[5 * i + 3] = pdp3::rsqrt_cr(v);
  out[5 * i + 4] = pdp3::log_cr(v);
}

extern "C" int mathprobe_run(const double *x, int n, double *out) {
  double *d = nullptr;
  if (hipMalloc(&d, sizeof(double) * 6 * (size_t)n) != hipSuccess) return -1;
  (void)hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice);
  hipLaunchKernelGGL(k_probe, dim3((n + 255) / 256), dim3(256), 0, 0, d, n, d + n);
  hipError_t e = hipMemcpy(out, d + n, sizeof(double) * 5 * (size_t)n, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? 0 : -2;
}
