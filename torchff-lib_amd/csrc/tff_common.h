// Shared device/host helpers for the torchff MI355X (gfx950) library.
//
// Conventions used by every kernel in this library:
//  * box: 3x3 row-major lattice vectors (rows a, b, c), as torchff passes it
//    (csrc/common/pbc.cuh:6-18 of the reference); its inverse is formed on the
//    host by cofactors in float64 (pbc.cuh:38-64) and travels by value.
//  * minimum image: s = d . B^-1, s -= round(s), d = s . B, evaluated in
//    float64 with floating-point contraction disabled, so the cutoff
//    predicate is bit-identical to a plain IEEE evaluation on the CPU.
//  * wavefronts are 64 lanes (CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define TFF_WAVE 64

enum TffDtype { TFF_F32 = 0, TFF_F64 = 1 };

enum TffStatus {
  TFF_OK = 0,
  TFF_ERR_ARG = 1,
  TFF_ERR_WORKSPACE = 2,
  TFF_ERR_LAUNCH = 3,
  TFF_ERR_UNSUPPORTED = 4,
  TFF_FULL = 5,  // a capacity would be exceeded (nothing was changed)
};

struct TffBox {
  double b[9];
  double inv[9];
  double det;
  int periodic;
  int ortho;  // periodic with every off-diagonal entry of the box and of its inverse zero
};

// Host: cofactor inverse with the reference's formula, contraction off.
static inline TffBox tff_make_box(const double* m) {
#pragma clang fp contract(off)
  TffBox B;
  for (int k = 0; k < 9; ++k) B.b[k] = m ? m[k] : 0.0;
  B.periodic = m != nullptr;
  B.ortho = 0;
  if (!m) {
    for (int k = 0; k < 9; ++k) B.inv[k] = 0.0;
    B.det = 0.0;
    return B;
  }
  double m00 = m[0], m01 = m[1], m02 = m[2];
  double m10 = m[3], m11 = m[4], m12 = m[5];
  double m20 = m[6], m21 = m[7], m22 = m[8];
  double c00 = m11 * m22 - m12 * m21;
  double c01 = m12 * m20 - m10 * m22;
  double c02 = m10 * m21 - m11 * m20;
  double c10 = m02 * m21 - m01 * m22;
  double c11 = m00 * m22 - m02 * m20;
  double c12 = m01 * m20 - m00 * m21;
  double c20 = m01 * m12 - m02 * m11;
  double c21 = m02 * m10 - m00 * m12;
  double c22 = m00 * m11 - m01 * m10;
  double det = m00 * c00 + m01 * c01 + m02 * c02;
  double id = 1.0 / det;
  B.inv[0] = c00 * id; B.inv[1] = c10 * id; B.inv[2] = c20 * id;
  B.inv[3] = c01 * id; B.inv[4] = c11 * id; B.inv[5] = c21 * id;
  B.inv[6] = c02 * id; B.inv[7] = c12 * id; B.inv[8] = c22 * id;
  B.det = det;
  B.ortho = 1;
  for (int k = 0; k < 9; ++k)
    if (k % 4 != 0 && (B.b[k] != 0.0 || B.inv[k] != 0.0)) B.ortho = 0;
  return B;
}

// Minimum-image displacement, float64, fixed operation order, no FMA.  Orthorhombic boxes take the
// same operations without the zero off-diagonal terms: adding an exact zero leaves a sum unchanged
// (up to the sign of a zero, which the squared distance loses), so the result is the general one.
__device__ __forceinline__ void tff_min_image_d(const TffBox& B, double& x, double& y, double& z) {
#pragma clang fp contract(off)
  if (!B.periodic) return;
  if (B.ortho) {
    double sa = x * B.inv[0], sb = y * B.inv[4], sc = z * B.inv[8];
    sa = sa - round(sa);
    sb = sb - round(sb);
    sc = sc - round(sc);
    x = sa * B.b[0];
    y = sb * B.b[4];
    z = sc * B.b[8];
    return;
  }
  double sa = x * B.inv[0] + y * B.inv[3] + z * B.inv[6];
  double sb = x * B.inv[1] + y * B.inv[4] + z * B.inv[7];
  double sc = x * B.inv[2] + y * B.inv[5] + z * B.inv[8];
  sa = sa - round(sa);
  sb = sb - round(sb);
  sc = sc - round(sc);
  x = sa * B.b[0] + sb * B.b[3] + sc * B.b[6];
  y = sa * B.b[1] + sb * B.b[4] + sc * B.b[7];
  z = sa * B.b[2] + sb * B.b[5] + sc * B.b[8];
}

__device__ __forceinline__ double tff_dist2_d(const TffBox& B, double dx, double dy, double dz) {
#pragma clang fp contract(off)
  tff_min_image_d(B, dx, dy, dz);
  return dx * dx + dy * dy + dz * dz;
}

// Minimum image in the working precision T (used by energy kernels, not by predicates).
template <typename T>
__device__ __forceinline__ void tff_min_image(const TffBox& B, T& x, T& y, T& z) {
  if (!B.periodic) return;
  if (B.ortho) {  // the general form without its zero terms (same values)
    T sa = x * (T)B.inv[0], sb = y * (T)B.inv[4], sc = z * (T)B.inv[8];
    sa -= rint(sa);
    sb -= rint(sb);
    sc -= rint(sc);
    x = sa * (T)B.b[0];
    y = sb * (T)B.b[4];
    z = sc * (T)B.b[8];
    return;
  }
  T sa = x * (T)B.inv[0] + y * (T)B.inv[3] + z * (T)B.inv[6];
  T sb = x * (T)B.inv[1] + y * (T)B.inv[4] + z * (T)B.inv[7];
  T sc = x * (T)B.inv[2] + y * (T)B.inv[5] + z * (T)B.inv[8];
  sa -= rint(sa);
  sb -= rint(sb);
  sc -= rint(sc);
  x = sa * (T)B.b[0] + sb * (T)B.b[3] + sc * (T)B.b[6];
  y = sa * (T)B.b[1] + sb * (T)B.b[4] + sc * (T)B.b[7];
  z = sa * (T)B.b[2] + sb * (T)B.b[5] + sc * (T)B.b[8];
}

// CSR exclusion test: is j listed in row i?  (reference csrc/nblist/exclusions.cuh:26-36)
__device__ __forceinline__ bool tff_is_excluded(int64_t i, int64_t j, const int64_t* row_ptr,
                                                const int64_t* col) {
  if (!col) return false;
  int64_t s = row_ptr[i], e = row_ptr[i + 1];
  for (int64_t k = s; k < e; ++k)
    if (col[k] == j) return true;
  return false;
}

// Makes this wave's LDS writes visible to its own lanes (lanes read what other lanes wrote).
__device__ __forceinline__ void tff_wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__device__ __forceinline__ T tff_wave_sum(T v) {
  for (int off = TFF_WAVE / 2; off > 0; off >>= 1) v += __shfl_down(v, off, TFF_WAVE);
  return v;
}

// Block-wide sum of a double; result valid in thread 0.  blockDim.x <= 1024.
__device__ __forceinline__ double tff_block_sum(double v) {
  __shared__ double s_part[1024 / TFF_WAVE];
  v = tff_wave_sum(v);
  int lane = threadIdx.x & (TFF_WAVE - 1);
  int wid = threadIdx.x / TFF_WAVE;
  __syncthreads();
  if (lane == 0) s_part[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
    int nw = (blockDim.x + TFF_WAVE - 1) / TFF_WAVE;
    for (int w = 0; w < nw; ++w) r += s_part[w];
  }
  return r;
}

// Exclusive scan of in[0..m) into out[0..m) (and out[m] = total when write_total) by one workgroup
// of 1024 threads, in tiles of 4096: each thread loads four consecutive entries (coalesced), wave
// scans by shuffles, wave totals through LDS, a running carry between tiles.  Every thread of the
// workgroup calls it; returns the total in every thread.  O: the type of the sums.
template <typename C, typename O>
__device__ __forceinline__ O tff_block_scan_1024(const C* __restrict__ in, long long m, O* __restrict__ out,
                                                 bool write_total) {
  __shared__ O s_w[1024 / TFF_WAVE];
  __shared__ O s_carry;
  const int t = threadIdx.x, lane = t & (TFF_WAVE - 1), wv = t / TFF_WAVE;
  if (t == 0) s_carry = 0;
  __syncthreads();
  for (long long base = 0; base < m; base += 4096) {
    const long long i0 = base + 4 * t;
    O v[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = i0 + k < m ? (O)in[i0 + k] : (O)0;
    const O local = v[0] + v[1] + v[2] + v[3];
    O incl = local;
#pragma unroll
    for (int off = 1; off < TFF_WAVE; off <<= 1) {
      const O x = __shfl_up(incl, off, TFF_WAVE);
      if (lane >= off) incl += x;
    }
    if (lane == TFF_WAVE - 1) s_w[wv] = incl;
    __syncthreads();
    O before = s_carry;
    for (int q = 0; q < wv; ++q) before += s_w[q];
    O run = before + incl - local;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (i0 + k < m) out[i0 + k] = run;
      run += v[k];
    }
    __syncthreads();  // every thread has read s_carry and s_w
    if (t == 1023) s_carry = before + incl;
    __syncthreads();
  }
  const O total = s_carry;
  if (write_total && t == 0) out[m] = total;
  return total;
}

// Deterministic sum of n partials by one workgroup (fixed strided order + tree); valid in thread 0.
__device__ __forceinline__ double tff_sum_partials(const double* __restrict__ part, int n) {
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += blockDim.x) acc += part[k];
  return tff_block_sum(acc);
}

// Fixed-point (int64) accumulation of forces and other per-atom sums: value * 2^FP_SHIFT.
#define TFF_FP_SCALE 4294967296.0  // 2^32
__device__ __forceinline__ unsigned long long tff_to_fixed(double v) {
  return (unsigned long long)(long long)__double2ll_rn(v * TFF_FP_SCALE);
}
__device__ __forceinline__ double tff_from_fixed(long long v) { return (double)v / TFF_FP_SCALE; }

static inline size_t tff_align(size_t x) { return (x + 255) & ~(size_t)255; }

// Thread-local error message of the last failing call.
void tff_set_error(const char* fmt, ...);

// Optional event timing around one kernel launch (no-ops unless tff_prof_enable(1)).
void* tff_prof_begin(const char* name, hipStream_t stream);
void tff_prof_end(void* token, hipStream_t stream);
bool tff_prof_active();  // event timing requested (tff_prof_enable)

#define TFF_CHECK_LAUNCH(name)                                              \
  do {                                                                      \
    hipError_t _e = hipGetLastError();                                      \
    if (_e != hipSuccess) {                                                 \
      tff_set_error("%s: kernel launch failed: %s", name, hipGetErrorString(_e)); \
      return TFF_ERR_LAUNCH;                                                \
    }                                                                       \
  } while (0)
