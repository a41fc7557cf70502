// Shared smooth-PME device code: B-spline weights, grid anchoring, mixed-radix
// Stockham FFT passes over LDS-resident lines (with the fused influence-function
// pass), used by the drop-in PME entry points and the fused cluster path.
#pragma once

#include <math.h>

#include "tff_common.h"

namespace tffpme {
namespace {

template <typename T>
struct cx {
  T x, y;
};

constexpr int FFT_THREADS = 256;
constexpr int MAX_STAGES = 24;
constexpr size_t FFT_LDS_BUDGET = 60 * 1024;  // dynamic LDS per workgroup, below the 64 KiB launch limit

struct FftPlan {
  int K;
  int nst;
  int radix[MAX_STAGES];
};

struct PmeGeom {
  int K[3];
  int order;
  double alpha;
  double inv[9];  // B^-1, row-major
  double vol;
};

bool make_plan(int K, FftPlan& p) {
  p.K = K;
  p.nst = 0;
  int m = K;
  while (m % 4 == 0 && p.nst < MAX_STAGES) { p.radix[p.nst++] = 4; m /= 4; }
  while (m % 2 == 0 && p.nst < MAX_STAGES) { p.radix[p.nst++] = 2; m /= 2; }
  while (m % 3 == 0 && p.nst < MAX_STAGES) { p.radix[p.nst++] = 3; m /= 3; }
  while (m % 5 == 0 && p.nst < MAX_STAGES) { p.radix[p.nst++] = 5; m /= 5; }
  return m == 1 && K >= 1;
}

// ------------------------------------------------------------------ splines
// v[k] = M_P(t + k), dv[k] = M_P'(t + k), k = 0..P-1, t in [0, 1)
template <typename T, int P>
__device__ __forceinline__ void spline_weights(T t, T* v, T* dv) {
  T w[P];
  w[0] = (T)1;
#pragma unroll
  for (int k = 1; k < P; ++k) w[k] = (T)0;
#pragma unroll
  for (int n = 2; n <= P; ++n) {
    if (n == P) {
#pragma unroll
      for (int k = 0; k < P; ++k) dv[k] = w[k] - (k > 0 ? w[k - 1] : (T)0);
    }
    const T div = (T)1 / (T)(n - 1);
#pragma unroll
    for (int k = n - 1; k >= 0; --k) {
      T a = (k < n - 1) ? w[k] : (T)0;
      T b = (k > 0) ? w[k - 1] : (T)0;
      w[k] = ((t + (T)k) * a + ((T)n - t - (T)k) * b) * div;
    }
  }
#pragma unroll
  for (int k = 0; k < P; ++k) v[k] = w[k];
}

// as spline_weights, plus d2v[k] = M_P''(t + k) = M_{P-2}(t+k) - 2 M_{P-2}(t+k-1) + M_{P-2}(t+k-2)
template <typename T, int P>
__device__ __forceinline__ void spline_weights2(T t, T* v, T* dv, T* d2v) {
  T w[P];
  w[0] = (T)1;
#pragma unroll
  for (int k = 1; k < P; ++k) w[k] = (T)0;
#pragma unroll
  for (int n = 2; n <= P; ++n) {
    if (n == P - 1) {
#pragma unroll
      for (int k = 0; k < P; ++k)
        d2v[k] = w[k] - (T)2 * (k > 0 ? w[k - 1] : (T)0) + (k > 1 ? w[k - 2] : (T)0);
    }
    if (n == P) {
#pragma unroll
      for (int k = 0; k < P; ++k) dv[k] = w[k] - (k > 0 ? w[k - 1] : (T)0);
    }
    const T div = (T)1 / (T)(n - 1);
#pragma unroll
    for (int k = n - 1; k >= 0; --k) {
      T a = (k < n - 1) ? w[k] : (T)0;
      T b = (k > 0) ? w[k - 1] : (T)0;
      w[k] = ((t + (T)k) * a + ((T)n - t - (T)k) * b) * div;
    }
  }
#pragma unroll
  for (int k = 0; k < P; ++k) v[k] = w[k];
}

// as spline_weights2, plus d3v[k] = third derivative of M_P at t + k = M_{P-3}(t+k)
// - 3 M_{P-3}(t+k-1) + 3 M_{P-3}(t+k-2) - M_{P-3}(t+k-3); needs P >= 4
template <typename T, int P>
__device__ __forceinline__ void spline_weights3(T t, T* v, T* dv, T* d2v, T* d3v) {
  static_assert(P >= 4, "third derivative needs spline order >= 4");
  T w[P];
  w[0] = (T)1;
#pragma unroll
  for (int k = 1; k < P; ++k) w[k] = (T)0;
#pragma unroll
  for (int n = 2; n <= P; ++n) {
    if (n == P - 2) {
#pragma unroll
      for (int k = 0; k < P; ++k)
        d3v[k] = w[k] - (T)3 * (k > 0 ? w[k - 1] : (T)0) + (T)3 * (k > 1 ? w[k - 2] : (T)0) -
                 (k > 2 ? w[k - 3] : (T)0);
    }
    if (n == P - 1) {
#pragma unroll
      for (int k = 0; k < P; ++k)
        d2v[k] = w[k] - (T)2 * (k > 0 ? w[k - 1] : (T)0) + (k > 1 ? w[k - 2] : (T)0);
    }
    if (n == P) {
#pragma unroll
      for (int k = 0; k < P; ++k) dv[k] = w[k] - (k > 0 ? w[k - 1] : (T)0);
    }
    const T div = (T)1 / (T)(n - 1);
#pragma unroll
    for (int k = n - 1; k >= 0; --k) {
      T a = (k < n - 1) ? w[k] : (T)0;
      T b = (k > 0) ? w[k - 1] : (T)0;
      w[k] = ((t + (T)k) * a + ((T)n - t - (T)k) * b) * div;
    }
  }
#pragma unroll
  for (int k = 0; k < P; ++k) v[k] = w[k];
}

// grid base (already wrapped into [0, K)) and spline parameter t for one axis
__device__ __forceinline__ void anchor_axis(double s, int K, int P, int& base, double& t) {
  double shift = (P & 1) ? 0.5 : 0.0;
  double m0 = ceil(s - shift);
  t = m0 - s + 0.5 * P - (double)(P / 2);
  // (m0 - P / 2) mod K in double: integer-valued operands below 2^52, so the product and the
  // difference are exact; the quotient may be off by one, fixed by the two corrections
  const double b = m0 - (double)(P / 2);
  const double Kd = (double)K;
  double r = b - Kd * floor(b / Kd);
  r = r < 0.0 ? r + Kd : r;
  r = r >= Kd ? r - Kd : r;
  base = (int)r;
}

// ------------------------------------------------------------------ FFT
template <typename T, int DIR, int R>
__device__ __forceinline__ void butterfly(cx<T>* v) {
  // X_k = sum_r v_r exp(DIR 2 pi i r k / R), DIR = -1 forward, +1 inverse
  if constexpr (R == 2) {
    cx<T> a = v[0], b = v[1];
    v[0] = {a.x + b.x, a.y + b.y};
    v[1] = {a.x - b.x, a.y - b.y};
  } else if constexpr (R == 4) {
    cx<T> s02 = {v[0].x + v[2].x, v[0].y + v[2].y}, d02 = {v[0].x - v[2].x, v[0].y - v[2].y};
    cx<T> s13 = {v[1].x + v[3].x, v[1].y + v[3].y}, d13 = {v[1].x - v[3].x, v[1].y - v[3].y};
    // DIR * i * d13
    cx<T> id13 = {-(T)DIR * d13.y, (T)DIR * d13.x};
    v[0] = {s02.x + s13.x, s02.y + s13.y};
    v[2] = {s02.x - s13.x, s02.y - s13.y};
    v[1] = {d02.x + id13.x, d02.y + id13.y};
    v[3] = {d02.x - id13.x, d02.y - id13.y};
  } else if constexpr (R == 3) {
    const T s3 = (T)0.86602540378443864676;
    cx<T> t1 = {v[1].x + v[2].x, v[1].y + v[2].y};
    cx<T> d = {v[1].x - v[2].x, v[1].y - v[2].y};
    cx<T> m = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
    cx<T> id = {-(T)DIR * s3 * d.y, (T)DIR * s3 * d.x};
    v[0] = {v[0].x + t1.x, v[0].y + t1.y};
    v[1] = {m.x + id.x, m.y + id.y};
    v[2] = {m.x - id.x, m.y - id.y};
  } else {  // R == 5
    const T c1 = (T)0.30901699437494742410, c2 = (T)-0.80901699437494742410;
    const T s1 = (T)0.95105651629515357212, s2 = (T)0.58778525229247312917;
    cx<T> a1 = {v[1].x + v[4].x, v[1].y + v[4].y}, b1 = {v[1].x - v[4].x, v[1].y - v[4].y};
    cx<T> a2 = {v[2].x + v[3].x, v[2].y + v[3].y}, b2 = {v[2].x - v[3].x, v[2].y - v[3].y};
    cx<T> x0 = v[0];
    cx<T> m1 = {x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y};
    cx<T> m2 = {x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y};
    // n1 = DIR * i * (s1 b1 + s2 b2), n2 = DIR * i * (s2 b1 - s1 b2)
    cx<T> p1 = {s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y};
    cx<T> p2 = {s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y};
    cx<T> n1 = {-(T)DIR * p1.y, (T)DIR * p1.x};
    cx<T> n2 = {-(T)DIR * p2.y, (T)DIR * p2.x};
    v[0] = {x0.x + a1.x + a2.x, x0.y + a1.y + a2.y};
    v[1] = {m1.x + n1.x, m1.y + n1.y};
    v[4] = {m1.x - n1.x, m1.y - n1.y};
    v[2] = {m2.x + n2.x, m2.y + n2.y};
    v[3] = {m2.x - n2.x, m2.y - n2.y};
  }
}

// x / d and x % d for 0 <= x < 2^22, 1 <= d <= 4096 through a float reciprocal: (x + 0.5) / d is at
// least 0.5 / d from an integer and the float error is below 2^-22 (x + 0.5) / d, so the
// truncation is exact.
__device__ __forceinline__ int fdiv(int x, float inv_d) { return (int)(((float)x + 0.5f) * inv_d); }

// One Stockham stage of radix R over nl lines of length K (line stride LS) held in LDS.
// nl is a power of two (nl = 1 << nl_shift); work item w -> line w & (nl - 1), index w >> nl_shift.
template <typename T, int DIR, int R>
__device__ __forceinline__ void fft_stage(const cx<T>* __restrict__ src, cx<T>* __restrict__ dst,
                                          const cx<T>* __restrict__ tw, int K, int LS, int nl_shift, int Ns) {
  const int KR = K / R;
  const int items = KR << nl_shift;
  const int tstep = K / (Ns * R);
  const float inv_ns = 1.0f / (float)Ns;
  for (int w = threadIdx.x; w < items; w += blockDim.x) {
    const int line = w & ((1 << nl_shift) - 1);
    const int j = w >> nl_shift;
    const cx<T>* in = src + line * LS;
    cx<T>* o = dst + line * LS;
    const int jq = fdiv(j, inv_ns);
    const int jm = j - jq * Ns;
    cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      cx<T> x = in[j + r * KR];
      if (r > 0) {
        cx<T> t = tw[jm * r * tstep];
        T ty = DIR < 0 ? t.y : -t.y;
        x = {x.x * t.x - x.y * ty, x.x * ty + x.y * t.x};
      }
      v[r] = x;
    }
    butterfly<T, DIR, R>(v);
    const int idxD = jq * Ns * R + jm;
#pragma unroll
    for (int r = 0; r < R; ++r) o[idxD + r * Ns] = v[r];
  }
}

// Stockham transform of nl lines of length K (line stride LS) held in LDS.
// Returns the buffer holding the result (a or b).
template <typename T, int DIR>
__device__ cx<T>* fft_lines(cx<T>* a, cx<T>* b, const cx<T>* tw, int K, int LS, int nl_shift, const FftPlan& plan) {
  cx<T>* src = a;
  cx<T>* dst = b;
  int Ns = 1;
  for (int s = 0; s < plan.nst; ++s) {
    const int R = plan.radix[s];
    if (R == 4)
      fft_stage<T, DIR, 4>(src, dst, tw, K, LS, nl_shift, Ns);
    else if (R == 2)
      fft_stage<T, DIR, 2>(src, dst, tw, K, LS, nl_shift, Ns);
    else if (R == 3)
      fft_stage<T, DIR, 3>(src, dst, tw, K, LS, nl_shift, Ns);
    else
      fft_stage<T, DIR, 5>(src, dst, tw, K, LS, nl_shift, Ns);
    __syncthreads();
    cx<T>* t = src;
    src = dst;
    dst = t;
    Ns *= R;
  }
  return src;
}

enum PassMode { PASS_Z_FWD = 0, PASS_Y_FWD = 1, PASS_X_CONV = 2, PASS_Y_INV = 3, PASS_Z_INV = 4 };

struct PassArgs {
  int A, K, C;   // element (a, k, c) at (a*K + k)*C + c
  int AB, CB;    // tile: AB values of a, CB values of c (powers of two, one of them 1)
  int ab_shift, cb_shift;
  int LS;        // LDS line stride
};

struct ConvArgs {
  int K[3];
  double inv[9];
  double vol;
  double alpha;
  int grid_fixed;         // forward z pass input: 0 = float64 grid, 1 = int64 fixed-point grid
  double grid_inv_scale;  // fixed-point scale of that grid (value = integer * grid_inv_scale)
  int nz;                 // z extent of the spectrum: K[2] (full) or K[2] / 2 + 1 (real-input half)
  int half;               // 1: half spectrum along z (energy weight 2 on kz in (0, K[2] / 2))
  int clear_in;           // forward z pass: zero each fixed-point grid value after reading it
};

// cos(pi x), sin(pi x) in the transform's precision
template <typename T>
__device__ __forceinline__ void twiddle_pi(double x, T& c, T& s) {
  if constexpr (sizeof(T) == 4) {
    float sn, cs;
    sincospif((float)x, &sn, &cs);
    c = cs;
    s = sn;
  } else {
    double sn, cs;
    sincospi(x, &sn, &cs);
    c = cs;
    s = sn;
  }
}

// tw[m] = exp(-2 pi i m / K), m < K (LDS; the caller synchronises before use)
template <typename T>
__device__ __forceinline__ void fill_twiddles(cx<T>* tw, int K) {
  for (int m = threadIdx.x; m < K; m += blockDim.x) {
    T c, s;
    twiddle_pi(-2.0 * (double)m / (double)K, c, s);
    tw[m] = {c, s};
  }
}

template <typename T>
__device__ __forceinline__ T signed_m(int m, int K) {
  return (T)(m <= K / 2 ? m : m - K);
}

// One tile of a pass (the body of k_fft_pass): tile index and k-space energy slot explicit, so
// that batched launches over many grids can drive it.
template <typename T, int MODE>
__device__ __forceinline__ void fft_pass_tile(const void* __restrict__ in_v, void* __restrict__ out_v,
                                              const PassArgs& pa, const FftPlan& plan, const ConvArgs& cv,
                                              const T* __restrict__ modx, const T* __restrict__ mody,
                                              const T* __restrict__ modz, int modz_half, int tile,
                                              double* __restrict__ e_slot) {
  extern __shared__ unsigned char lds_raw[];
  const int K = pa.K, LS = pa.LS;
  const int nl = pa.AB * pa.CB;
  cx<T>* tw = reinterpret_cast<cx<T>*>(lds_raw);
  cx<T>* bufA = tw + K;
  cx<T>* bufB = bufA + nl * LS;
  // twiddles exp(-2 pi i m / K)
  fill_twiddles(tw, K);
  const int ntc = (pa.C + pa.CB - 1) / pa.CB;
  const int a0 = (tile / ntc) * pa.AB;
  const int c0 = (tile % ntc) * pa.CB;
  const int nelem = nl * K;
  const int nl_shift = pa.ab_shift + pa.cb_shift;
  const float inv_k = 1.0f / (float)K;
  // element idx of the tile -> (ab, k, cb): idx = (ab * K + k) * CB + cb
  auto split = [&](int idx, int& ab, int& k, int& cb) {
    cb = idx & (pa.CB - 1);
    const int rest = idx >> pa.cb_shift;
    ab = pa.AB == 1 ? 0 : fdiv(rest, inv_k);
    k = rest - ab * K;
  };
  // load
  for (int idx = threadIdx.x; idx < nelem; idx += blockDim.x) {
    int ab, k, cb;
    split(idx, ab, k, cb);
    int a = a0 + ab, c = c0 + cb;
    cx<T> v = {(T)0, (T)0};
    if (a < pa.A && c < pa.C) {
      long long g = ((long long)a * K + k) * pa.C + c;
      if (MODE == PASS_Z_FWD) {
        if (cv.grid_fixed) {
          long long* gi = reinterpret_cast<long long*>(const_cast<void*>(in_v));
          v.x = (T)((double)gi[g] * cv.grid_inv_scale);
          if (cv.clear_in) gi[g] = 0;  // the value is read by this thread only
        } else
          v.x = (T)reinterpret_cast<const double*>(in_v)[g];
      } else {
        v = reinterpret_cast<const cx<T>*>(in_v)[g];
      }
    }
    bufA[(ab * pa.CB + cb) * LS + k] = v;
  }
  __syncthreads();
  cx<T>* res;
  if (MODE == PASS_Z_FWD || MODE == PASS_Y_FWD || MODE == PASS_X_CONV)
    res = fft_lines<T, -1>(bufA, bufB, tw, K, LS, nl_shift, plan);
  else
    res = fft_lines<T, 1>(bufA, bufB, tw, K, LS, nl_shift, plan);
  if (MODE == PASS_X_CONV) {
    // lines are (y, z) = c; along the line k = mx
    double e_acc = 0.0;
    const T two_pi = (T)6.283185307179586;
    const float inv_nz = 1.0f / (float)cv.nz;
    for (int idx = threadIdx.x; idx < nelem; idx += blockDim.x) {
      int line = fdiv(idx, inv_k);
      int k = idx - line * K;
      int c = c0 + line;  // AB == 1 for this pass
      if (c >= pa.C) continue;
      int my = fdiv(c, inv_nz), mz = c - my * cv.nz;
      T fx = signed_m<T>(k, cv.K[0]), fy = signed_m<T>(my, cv.K[1]), fz = signed_m<T>(mz, cv.K[2]);
      T kx = two_pi * (fx * (T)cv.inv[0] + fy * (T)cv.inv[1] + fz * (T)cv.inv[2]);
      T ky = two_pi * (fx * (T)cv.inv[3] + fy * (T)cv.inv[4] + fz * (T)cv.inv[5]);
      T kz = two_pi * (fx * (T)cv.inv[6] + fy * (T)cv.inv[7] + fz * (T)cv.inv[8]);
      T ksq = kx * kx + ky * ky + kz * kz;
      cx<T>& s = res[line * LS + k];
      T G = 0;
      if (k != 0 || my != 0 || mz != 0) {
        int zi = modz_half ? (mz <= cv.K[2] / 2 ? mz : cv.K[2] - mz) : mz;
        T th = modx[k] * mody[my] * modz[zi];
        G = (T)(4.0 * 3.14159265358979323846 / cv.vol) / ksq * exp(-ksq / (T)(4.0 * cv.alpha * cv.alpha)) / (th * th);
      }
      // half spectrum: kz in (0, K/2) stands for kz and -kz
      const double wz = cv.half && mz != 0 && 2 * mz != cv.K[2] ? 1.0 : 0.5;
      e_acc += wz * (double)G * ((double)s.x * (double)s.x + (double)s.y * (double)s.y);
      s.x *= G;
      s.y *= G;
    }
    __syncthreads();
    cx<T>* other = (res == bufA) ? bufB : bufA;
    res = fft_lines<T, 1>(res, other, tw, K, LS, nl_shift, plan);
    double tot = tff_block_sum(e_acc);
    if (threadIdx.x == 0) *e_slot = tot;
  }
  // store
  for (int idx = threadIdx.x; idx < nelem; idx += blockDim.x) {
    int ab, k, cb;
    split(idx, ab, k, cb);
    int a = a0 + ab, c = c0 + cb;
    if (a < pa.A && c < pa.C) {
      long long g = ((long long)a * K + k) * pa.C + c;
      cx<T> v = res[(ab * pa.CB + cb) * LS + k];
      if (MODE == PASS_Z_INV)
        reinterpret_cast<T*>(out_v)[g] = v.x;
      else
        reinterpret_cast<cx<T>*>(out_v)[g] = v;
    }
  }
}

template <typename T, int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_pass(const void* __restrict__ in_v, void* __restrict__ out_v,
                                                          PassArgs pa, FftPlan plan, ConvArgs cv,
                                                          const T* __restrict__ modx, const T* __restrict__ mody,
                                                          const T* __restrict__ modz, int modz_half,
                                                          double* __restrict__ e_part) {
  fft_pass_tile<T, MODE>(in_v, out_v, pa, plan, cv, modx, mody, modz, modz_half, blockIdx.x,
                         MODE == PASS_X_CONV ? e_part + blockIdx.x : nullptr);
}

// Real-to-complex transform along the contiguous (z) axis of a real grid of nlines lines of
// K2 reals, K2 even: a line is taken as M = K2 / 2 complex values z_n = x_2n + i x_2n+1,
// transformed by an M-point FFT (Z), and untangled into X_k = E_k + w^k O_k for k = 0..M with
// E_k = (Z_k + conj Z_M-k) / 2, O_k = (Z_k - conj Z_M-k) / 2i, w = exp(-2 pi i / K2) (Z_M = Z_0).
// Output: nlines x (M + 1) complex.  Input: int64 fixed point (grid_fixed) or float64.
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_z_r2c(const void* __restrict__ in_v, cx<T>* __restrict__ out,
                                                           int nlines, int K2, int AB, int ab_shift, FftPlan plan,
                                                           int grid_fixed, double grid_inv_scale) {
  extern __shared__ unsigned char lds_raw[];
  const int M = K2 >> 1, LS = M + 1;
  cx<T>* tw = reinterpret_cast<cx<T>*>(lds_raw);
  cx<T>* bufA = tw + M;
  cx<T>* bufB = bufA + AB * LS;
  fill_twiddles(tw, M);
  const int l0 = blockIdx.x * AB;
  const float inv_m = 1.0f / (float)M, inv_m1 = 1.0f / (float)(M + 1);
  for (int idx = threadIdx.x; idx < AB * M; idx += blockDim.x) {
    const int ab = fdiv(idx, inv_m), n = idx - ab * M, line = l0 + ab;
    cx<T> v = {(T)0, (T)0};
    if (line < nlines) {
      const long long g = (long long)line * K2 + 2 * n;
      if (grid_fixed) {
        const long long* gi = reinterpret_cast<const long long*>(in_v);
        v = {(T)((double)gi[g] * grid_inv_scale), (T)((double)gi[g + 1] * grid_inv_scale)};
      } else {
        const double* gd = reinterpret_cast<const double*>(in_v);
        v = {(T)gd[g], (T)gd[g + 1]};
      }
    }
    bufA[ab * LS + n] = v;
  }
  __syncthreads();
  const cx<T>* res = fft_lines<T, -1>(bufA, bufB, tw, M, LS, ab_shift, plan);
  for (int idx = threadIdx.x; idx < AB * (M + 1); idx += blockDim.x) {
    const int ab = fdiv(idx, inv_m1), k = idx - ab * (M + 1), line = l0 + ab;
    if (line >= nlines) continue;
    const cx<T> zk = res[ab * LS + (k == M ? 0 : k)];
    const cx<T> zm = res[ab * LS + (k == 0 ? 0 : M - k)];
    const T ex = (T)0.5 * (zk.x + zm.x), ey = (T)0.5 * (zk.y - zm.y);
    const T ox = (T)0.5 * (zk.y + zm.y), oy = (T)-0.5 * (zk.x - zm.x);
    T c, sn;
    twiddle_pi(-2.0 * (double)k / (double)K2, c, sn);
    out[(long long)line * (M + 1) + k] = {ex + c * ox - sn * oy, ey + c * oy + sn * ox};
  }
}

// Inverse of k_fft_z_r2c without normalisation: X_k (k = 0..M) of real lines -> x_n =
// sum over the full spectrum of X_k exp(2 pi i k n / K2), via Z_k = (X_k + conj X_M-k)
// + i exp(2 pi i k / K2) (X_k - conj X_M-k), z = M-point inverse FFT of Z, x_2n + i x_2n+1 = z_n.
template <typename T>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_z_c2r(const cx<T>* __restrict__ in, T* __restrict__ out,
                                                           int nlines, int K2, int AB, int ab_shift, FftPlan plan) {
  extern __shared__ unsigned char lds_raw[];
  const int M = K2 >> 1, LS = M + 1;
  cx<T>* tw = reinterpret_cast<cx<T>*>(lds_raw);
  cx<T>* bufA = tw + M;
  cx<T>* bufB = bufA + AB * LS;
  fill_twiddles(tw, M);
  const int l0 = blockIdx.x * AB;
  const float inv_m = 1.0f / (float)M, inv_m1 = 1.0f / (float)(M + 1);
  for (int idx = threadIdx.x; idx < AB * (M + 1); idx += blockDim.x) {
    const int ab = fdiv(idx, inv_m1), k = idx - ab * (M + 1), line = l0 + ab;
    bufA[ab * LS + k] = line < nlines ? in[(long long)line * (M + 1) + k] : cx<T>{(T)0, (T)0};
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < AB * M; idx += blockDim.x) {
    const int ab = fdiv(idx, inv_m), k = idx - ab * M;
    const cx<T> xk = bufA[ab * LS + k], xm = bufA[ab * LS + M - k];
    const T ax = xk.x + xm.x, ay = xk.y - xm.y;
    const T bx = xk.x - xm.x, by = xk.y + xm.y;
    T c, sn;
    twiddle_pi(2.0 * (double)k / (double)K2, c, sn);
    const T wx = c * bx - sn * by, wy = c * by + sn * bx;
    bufB[ab * LS + k] = {ax - wy, ay + wx};
  }
  __syncthreads();
  const cx<T>* res = fft_lines<T, 1>(bufB, bufA, tw, M, LS, ab_shift, plan);
  for (int idx = threadIdx.x; idx < AB * M; idx += blockDim.x) {
    const int ab = fdiv(idx, inv_m), n = idx - ab * M, line = l0 + ab;
    if (line >= nlines) continue;
    const cx<T> z = res[ab * LS + n];
    out[(long long)line * K2 + 2 * n] = z.x;
    out[(long long)line * K2 + 2 * n + 1] = z.y;
  }
}

struct ZPlanCfg {
  int AB, ab_shift, blocks;
  size_t lds;
};
// tiles of the real z passes: AB lines (a power of two, at most 16) of M + 1 complex, two buffers
ZPlanCfg zcfg(int nlines, int K2, size_t celem) {
  ZPlanCfg c{};
  const int M = K2 / 2;
  const size_t per_line = 2 * (size_t)(M + 1) * celem;
  int lines = (int)((FFT_LDS_BUDGET - (size_t)M * celem) / per_line);
  if (lines > 16) lines = 16;
  if (lines < 1) lines = 1;
  c.AB = 1;
  while (c.AB * 2 <= lines) c.AB *= 2;
  c.ab_shift = 0;
  while ((1 << c.ab_shift) < c.AB) ++c.ab_shift;
  c.blocks = (nlines + c.AB - 1) / c.AB;
  c.lds = (size_t)M * celem + 2 * (size_t)c.AB * (M + 1) * celem;
  return c;
}

struct PassCfg {
  PassArgs pa;
  int tiles;
  size_t lds;
};

PassCfg pass_cfg(int A, int K, int C, size_t celem) {
  PassCfg c{};
  int LS = K + 1;
  size_t per_line = 2 * (size_t)LS * celem;
  int lines = (int)((FFT_LDS_BUDGET - (size_t)K * celem) / per_line);
  if (lines < 1) lines = 1;
  if (lines > 16) lines = 16;  // more, smaller tiles: >= 512 workgroups for a 96^3 grid
  int p2 = 1;
  while (p2 * 2 <= lines) p2 *= 2;  // a power of two: line indices by mask and shift
  lines = p2;
  c.pa.A = A;
  c.pa.K = K;
  c.pa.C = C;
  c.pa.LS = LS;
  if (C == 1) {
    c.pa.AB = lines < A ? lines : A;
    c.pa.CB = 1;
  } else {
    c.pa.AB = 1;
    c.pa.CB = lines < C ? lines : C;
  }
  // AB / CB: the largest power of two not above lines and A / C
  while (c.pa.AB & (c.pa.AB - 1)) c.pa.AB &= c.pa.AB - 1;
  while (c.pa.CB & (c.pa.CB - 1)) c.pa.CB &= c.pa.CB - 1;
  c.pa.ab_shift = 0;
  while ((1 << c.pa.ab_shift) < c.pa.AB) ++c.pa.ab_shift;
  c.pa.cb_shift = 0;
  while ((1 << c.pa.cb_shift) < c.pa.CB) ++c.pa.cb_shift;
  int ntc = (C + c.pa.CB - 1) / c.pa.CB;
  int nta = (A + c.pa.AB - 1) / c.pa.AB;
  c.tiles = ntc * nta;
  c.lds = (size_t)K * celem + 2 * (size_t)c.pa.AB * c.pa.CB * LS * celem;
  return c;
}

PmeGeom make_geom(const TffBox& box, const int K[3], int order, double alpha) {
  PmeGeom g;
  for (int a = 0; a < 3; ++a) g.K[a] = K[a];
  for (int k = 0; k < 9; ++k) g.inv[k] = box.inv[k];
  g.vol = box.det;
  g.order = order;
  g.alpha = alpha;
  return g;
}


}  // namespace
}  // namespace tffpme
