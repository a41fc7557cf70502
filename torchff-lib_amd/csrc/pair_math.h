// Pair arithmetic shared by the fused and the serving kernels.
#pragma once

#include <hip/hip_runtime.h>

// erfc(u) and exp(-u^2) for u >= 0.
// float: exp(-u^2) = 2^(-u^2 log2 e) through the hardware base-2 exponential (relative error
// below 7e-7 for u <= 3 from the rounding of the argument); erfc(u) exp(u^2) = sum_k c_k t^k,
// t = 1 / (1 + 0.39 u) (hardware reciprocal), a degree-8 polynomial fitted to the scaled
// complementary error function on [0, 3.6] by iteratively reweighted least squares (relative
// error 1.3e-9 in exact arithmetic, about 1.5e-6 evaluated in float32, against
// scipy.special.erfc).  The fit is kept this accurate because rigid molecules put every
// intramolecular pair at the same u, where a polynomial error would add up coherently over the
// exclusion corrections.  double: the correctly rounded library functions.
__device__ __forceinline__ void erfc_exp(float u, float& ec, float& em) {
  em = __builtin_amdgcn_exp2f(u * u * -1.44269502f);
  const float t = __builtin_amdgcn_rcpf(fmaf(0.39f, u, 1.0f));
  float p = 9.27117256e-02f;
  p = fmaf(p, t, -5.29840954e-01f);
  p = fmaf(p, t, 1.06552963e+00f);
  p = fmaf(p, t, -8.61721388e-01f);
  p = fmaf(p, t, 7.81784359e-01f);
  p = fmaf(p, t, -4.21555312e-02f);
  p = fmaf(p, t, 2.82117988e-01f);
  p = fmaf(p, t, 2.10995424e-01f);
  p = fmaf(p, t, 5.78745155e-04f);
  ec = p * em;
}
__device__ __forceinline__ void erfc_exp(double u, double& ec, double& em) {
  em = exp(-u * u);
  ec = erfc(u);
}
// hardware reciprocal square root (1 ulp; Lennard-Jones raises the error to 12 ulp of s^12)
__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_t(double x) { return 1.0 / sqrt(x); }
