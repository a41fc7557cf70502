// Smooth-PME reciprocal space for point charges on gfx950, float32 or float64.
//
// Replaces the reference's pme_long_range op for rank 0
// (csrc/pme/pme.cu:15-248: spread_q_kernel, torch rfftn, pme_convolution_fused_kernel,
// irfftn, 1/2 sum(Phi Q), self term; backward interpolate_kernel), generalised from
// the hard-wired spline order 6 to orders 3..8.
//
// Pipeline (one K0 x K1 x K2 grid, layout (x, y, z) with z fastest):
//   1. spread      one thread per atom: s = K (r . B^-1), m0 = ceil(s - (p mod 2)/2),
//                  weights M_p(t + k), t = m0 - s + p/2 - floor(p/2) in [0, 1), onto
//                  grid points m0 - floor(p/2) + k (k = 0..p-1), float64 atomics.
//   2. FFT z       (forward; reads the float64 charge grid, writes complex T)
//   3. FFT y       (forward)
//   4. FFT x  +  influence function G(k) = 4 pi exp(-k^2/4a^2) / (V k^2 |theta|^2)
//             +  k-space energy 1/2 sum G |S|^2 in float64  +  inverse FFT x,
//                all in LDS without writing the spectrum back.
//   5. FFT y       (inverse)
//   6. FFT z       (inverse; writes the real potential grid Phi in T)
//   7. gather      one thread per atom: phi_i = sum_g Phi W_i, dE/dr_i = q_i grad phi_i,
//                  dE/dq_i = phi_i - 2 a q_i / sqrt(pi).
// Every FFT pass is a mixed-radix (2, 3, 4, 5) Stockham transform of whole lines
// held in LDS (ping-pong buffers, padded line stride), batched so that global
// loads and stores are coalesced along the contiguous axis.  Energy partials are
// per workgroup and reduced in a fixed order.

#include <math.h>

#include "pme_core.h"
#include "tff_common.h"
#include "torchff_hip.h"

namespace {

using namespace tffpme;

template <typename T, int P>
__global__ __launch_bounds__(256) void k_spread(const T* __restrict__ coords, const T* __restrict__ q, long long n, PmeGeom g,
                         double* __restrict__ grid) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r0 = (double)coords[3 * i], r1 = (double)coords[3 * i + 1], r2 = (double)coords[3 * i + 2];
  T Mv[3][P], dM[3][P];
  int base[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double s = (r0 * g.inv[a] + r1 * g.inv[3 + a] + r2 * g.inv[6 + a]) * (double)g.K[a];
    double t;
    anchor_axis(s, g.K[a], P, base[a], t);
    spline_weights<T, P>((T)t, Mv[a], dM[a]);
  }
  const T qi = q[i];
  const int K1 = g.K[1], K2 = g.K[2];
  int gx = base[0];
  for (int a = 0; a < P; ++a) {
    int gy = base[1];
    const T wa = qi * Mv[0][a];
    for (int b = 0; b < P; ++b) {
      const T wab = wa * Mv[1][b];
      double* row = grid + ((long long)gx * K1 + gy) * K2;
      int gz = base[2];
#pragma unroll
      for (int c = 0; c < P; ++c) {
        atomicAdd(&row[gz], (double)(wab * Mv[2][c]));
        gz = gz + 1 == K2 ? 0 : gz + 1;
      }
      gy = gy + 1 == K1 ? 0 : gy + 1;
    }
    gx = gx + 1 == g.K[0] ? 0 : gx + 1;
  }
}

template <typename T, int P>
__global__ __launch_bounds__(256) void k_gather(const T* __restrict__ phi_grid, const T* __restrict__ coords, const T* __restrict__ q,
                         long long n, PmeGeom g, double scale, T* __restrict__ epot, T* __restrict__ dcoords) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r0 = (double)coords[3 * i], r1 = (double)coords[3 * i + 1], r2 = (double)coords[3 * i + 2];
  T Mv[3][P], dM[3][P];
  int base[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double s = (r0 * g.inv[a] + r1 * g.inv[3 + a] + r2 * g.inv[6 + a]) * (double)g.K[a];
    double t;
    anchor_axis(s, g.K[a], P, base[a], t);
    spline_weights<T, P>((T)t, Mv[a], dM[a]);
  }
  const int K1 = g.K[1], K2 = g.K[2];
  T phi = 0, du0 = 0, du1 = 0, du2 = 0;
  int gx = base[0];
  for (int a = 0; a < P; ++a) {
    int gy = base[1];
    for (int b = 0; b < P; ++b) {
      const T* row = phi_grid + ((long long)gx * K1 + gy) * K2;
      int gz = base[2];
      T s0 = 0, s1 = 0;
#pragma unroll
      for (int c = 0; c < P; ++c) {
        T v = row[gz];
        s0 += v * Mv[2][c];
        s1 += v * dM[2][c];
        gz = gz + 1 == K2 ? 0 : gz + 1;
      }
      phi += Mv[0][a] * Mv[1][b] * s0;
      du0 += dM[0][a] * Mv[1][b] * s0;
      du1 += Mv[0][a] * dM[1][b] * s0;
      du2 += Mv[0][a] * Mv[1][b] * s1;
      gy = gy + 1 == K1 ? 0 : gy + 1;
    }
    gx = gx + 1 == g.K[0] ? 0 : gx + 1;
  }
  // weights depend on t = m0 - s + const, so d/ds = -d/dt; ds_a/dr_j = K_a inv[j][a]
  const double qi = (double)q[i];
  const double d0 = -(double)du0 * g.K[0], d1 = -(double)du1 * g.K[1], d2 = -(double)du2 * g.K[2];
  const double gxr = d0 * g.inv[0] + d1 * g.inv[1] + d2 * g.inv[2];
  const double gyr = d0 * g.inv[3] + d1 * g.inv[4] + d2 * g.inv[5];
  const double gzr = d0 * g.inv[6] + d1 * g.inv[7] + d2 * g.inv[8];
  const double self = -2.0 * g.alpha * 0.5641895835477563 * qi;  // 1/sqrt(pi)
  if (epot) epot[i] = (T)(scale * ((double)phi + self));
  if (dcoords) {
    dcoords[3 * i] = (T)(scale * qi * gxr);
    dcoords[3 * i + 1] = (T)(scale * qi * gyr);
    dcoords[3 * i + 2] = (T)(scale * qi * gzr);
  }
}

__global__ void k_self_energy_parts(const void* q, int dtype, long long n, double alpha, double* part) {
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double qi = dtype == TFF_F64 ? ((const double*)q)[i] : (double)((const float*)q)[i];
    acc += qi * qi;
  }
  double tot = tff_block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = -alpha * 0.5641895835477563 * tot;
}

template <typename T>
__global__ __launch_bounds__(256) void k_pme_energy_final(const double* __restrict__ parts, int n, double scale,
                                                          T* __restrict__ out) {
  double s = tff_sum_partials(parts, n);
  if (threadIdx.x == 0) out[0] = (T)(s * scale);
}

__global__ void k_zero_grid(double* p, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) p[i] = 0.0;
}

constexpr int SELF_BLOCKS = 64;

struct PmeWs {
  double* grid;      // K^3 float64 charge grid
  void* work;        // K^3 complex T
  double* parts;     // energy partials
  size_t bytes;
  int nparts;
};

PmeWs carve(void* base, int dtype, const int K[3]) {
  PmeWs w{};
  size_t n = (size_t)K[0] * K[1] * K[2];
  size_t celem = dtype == TFF_F64 ? 16 : 8;
  PassCfg px = pass_cfg(1, K[0], K[1] * K[2], celem);
  w.nparts = px.tiles + SELF_BLOCKS;
  size_t off = 0;
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* r = p ? p + off : nullptr;
    off += tff_align(b);
    return r;
  };
  w.grid = (double*)take(sizeof(double) * n);
  w.work = (void*)take(celem * n);
  w.parts = (double*)take(sizeof(double) * (size_t)w.nparts);
  w.bytes = off;
  return w;
}

template <typename T>
int launch_spread(int order, const T* coords, const T* q, long long n, const PmeGeom& g, double* grid,
                  hipStream_t st) {
  unsigned nb = (unsigned)((n + 255) / 256);
  if (n == 0) return TFF_OK;
  void* tok = tff_prof_begin("pme_spread", st);
  switch (order) {
    case 3: hipLaunchKernelGGL((k_spread<T, 3>), dim3(nb), dim3(256), 0, st, coords, q, n, g, grid); break;
    case 4: hipLaunchKernelGGL((k_spread<T, 4>), dim3(nb), dim3(256), 0, st, coords, q, n, g, grid); break;
    case 5: hipLaunchKernelGGL((k_spread<T, 5>), dim3(nb), dim3(256), 0, st, coords, q, n, g, grid); break;
    case 6: hipLaunchKernelGGL((k_spread<T, 6>), dim3(nb), dim3(256), 0, st, coords, q, n, g, grid); break;
    case 7: hipLaunchKernelGGL((k_spread<T, 7>), dim3(nb), dim3(256), 0, st, coords, q, n, g, grid); break;
    case 8: hipLaunchKernelGGL((k_spread<T, 8>), dim3(nb), dim3(256), 0, st, coords, q, n, g, grid); break;
    default: tff_set_error("PME: unsupported spline order %d (3..8)", order); return TFF_ERR_UNSUPPORTED;
  }
  tff_prof_end(tok, st);
  TFF_CHECK_LAUNCH("k_spread");
  return TFF_OK;
}

template <typename T>
int launch_gather(int order, const T* phi, const T* coords, const T* q, long long n, const PmeGeom& g, double scale,
                  T* epot, T* dcoords, hipStream_t st) {
  unsigned nb = (unsigned)((n + 255) / 256);
  if (n == 0) return TFF_OK;
  void* tok = tff_prof_begin("pme_gather", st);
#define TFF_G(P) \
  hipLaunchKernelGGL((k_gather<T, P>), dim3(nb), dim3(256), 0, st, phi, coords, q, n, g, scale, epot, dcoords)
  switch (order) {
    case 3: TFF_G(3); break;
    case 4: TFF_G(4); break;
    case 5: TFF_G(5); break;
    case 6: TFF_G(6); break;
    case 7: TFF_G(7); break;
    case 8: TFF_G(8); break;
    default: tff_set_error("PME: unsupported spline order %d (3..8)", order); return TFF_ERR_UNSUPPORTED;
  }
#undef TFF_G
  tff_prof_end(tok, st);
  TFF_CHECK_LAUNCH("k_gather");
  return TFF_OK;
}

template <typename T, int MODE>
int launch_pass(const PassCfg& c, const FftPlan& plan, const ConvArgs& cv, const void* in, void* out, const T* mx,
                const T* my, const T* mz, int mz_half, double* parts, hipStream_t st) {
  void* tok = tff_prof_begin(MODE == PASS_X_CONV ? "pme_fft_conv" : "pme_fft", st);
  hipLaunchKernelGGL((k_fft_pass<T, MODE>), dim3(c.tiles), dim3(FFT_THREADS), c.lds, st, in, out, c.pa, plan, cv, mx,
                     my, mz, mz_half, parts);
  tff_prof_end(tok, st);
  TFF_CHECK_LAUNCH("k_fft_pass");
  return TFF_OK;
}

template <typename T>
int pme_energy(const T* coords, const T* q, long long n, const TffBox& box, const int K[3], int order, double alpha,
               const T* mx, const T* my, const T* mz, int mz_half, double scale, T* energy, T* phi_out,
               const PmeWs& w, hipStream_t st) {
  FftPlan pl[3];
  for (int a = 0; a < 3; ++a)
    if (!make_plan(K[a], pl[a])) {
      tff_set_error("PME: grid size %d is not of the form 2^a 3^b 5^c", K[a]);
      return TFF_ERR_UNSUPPORTED;
    }
  const long long ng = (long long)K[0] * K[1] * K[2];
  PmeGeom g = make_geom(box, K, order, alpha);
  {
    long long b = (ng + 255) / 256;
    if (b > 8192) b = 8192;
    hipLaunchKernelGGL(k_zero_grid, dim3((unsigned)b), dim3(256), 0, st, w.grid, ng);
  }
  int rc = launch_spread<T>(order, coords, q, n, g, w.grid, st);
  if (rc) return rc;
  const size_t ce = sizeof(cx<T>);
  ConvArgs cv{};
  for (int a = 0; a < 3; ++a) cv.K[a] = K[a];
  for (int k = 0; k < 9; ++k) cv.inv[k] = box.inv[k];
  cv.vol = box.det;
  cv.alpha = alpha;
  cv.grid_fixed = 0;
  cv.grid_inv_scale = 1.0;
  cv.nz = K[2];
  cv.half = 0;
  PassCfg cz = pass_cfg(K[0] * K[1], K[2], 1, ce);
  PassCfg cy = pass_cfg(K[0], K[1], K[2], ce);
  PassCfg cxp = pass_cfg(1, K[0], K[1] * K[2], ce);
  if ((rc = launch_pass<T, PASS_Z_FWD>(cz, pl[2], cv, w.grid, w.work, mx, my, mz, mz_half, w.parts, st))) return rc;
  if ((rc = launch_pass<T, PASS_Y_FWD>(cy, pl[1], cv, w.work, w.work, mx, my, mz, mz_half, w.parts, st))) return rc;
  if ((rc = launch_pass<T, PASS_X_CONV>(cxp, pl[0], cv, w.work, w.work, mx, my, mz, mz_half, w.parts, st))) return rc;
  if ((rc = launch_pass<T, PASS_Y_INV>(cy, pl[1], cv, w.work, w.work, mx, my, mz, mz_half, w.parts, st))) return rc;
  if ((rc = launch_pass<T, PASS_Z_INV>(cz, pl[2], cv, w.work, phi_out, mx, my, mz, mz_half, w.parts, st))) return rc;
  hipLaunchKernelGGL(k_self_energy_parts, dim3(SELF_BLOCKS), dim3(256), 0, st, (const void*)q,
                     (int)(sizeof(T) == 8 ? TFF_F64 : TFF_F32), n, alpha, w.parts + cxp.tiles);
  if (energy)
    hipLaunchKernelGGL(k_pme_energy_final<T>, dim3(1), dim3(256), 0, st, w.parts, cxp.tiles + SELF_BLOCKS, scale,
                       energy);
  TFF_CHECK_LAUNCH("pme energy");
  return TFF_OK;
}

int check_geom(const double* box_host, const int K[3], int order) {
  if (!box_host) {
    tff_set_error("PME requires a periodic box");
    return TFF_ERR_ARG;
  }
  if (order < 3 || order > 8) {
    tff_set_error("PME: unsupported spline order %d (3..8)", order);
    return TFF_ERR_UNSUPPORTED;
  }
  for (int a = 0; a < 3; ++a) {
    if (K[a] < order || K[a] > 4096) {
      tff_set_error("PME: grid size %d must be in [order, 4096]", K[a]);
      return TFF_ERR_ARG;
    }
    FftPlan p;
    if (!make_plan(K[a], p)) {
      tff_set_error("PME: grid size %d is not of the form 2^a 3^b 5^c", K[a]);
      return TFF_ERR_UNSUPPORTED;
    }
  }
  return TFF_OK;
}

}  // namespace

extern "C" size_t tff_pme_workspace_bytes(int dtype, long long natoms, int K0, int K1, int K2) {
  (void)natoms;
  int K[3] = {K0, K1, K2};
  return carve(nullptr, dtype, K).bytes;
}

extern "C" int tff_pme_energy(int dtype, const void* coords, const void* charges, long long natoms,
                              const double* box_host, int K0, int K1, int K2, int order, double alpha,
                              const void* xmoduli, const void* ymoduli, const void* zmoduli, int zmoduli_half,
                              double scale, void* energy_out, void* phi_grid_out, void* workspace,
                              size_t workspace_bytes, hipStream_t stream) {
  int K[3] = {K0, K1, K2};
  int rc = check_geom(box_host, K, order);
  if (rc) return rc;
  if (natoms < 0 || (natoms > 0 && (!coords || !charges)) || !xmoduli || !ymoduli || !zmoduli || !phi_grid_out) {
    tff_set_error("tff_pme_energy: invalid arguments");
    return TFF_ERR_ARG;
  }
  size_t need = tff_pme_workspace_bytes(dtype, natoms, K0, K1, K2);
  if (workspace_bytes < need || !workspace) {
    tff_set_error("tff_pme_energy: workspace too small (%zu < %zu)", workspace_bytes, need);
    return TFF_ERR_WORKSPACE;
  }
  TffBox box = tff_make_box(box_host);
  PmeWs w = carve(workspace, dtype, K);
  if (dtype == TFF_F64)
    return pme_energy<double>((const double*)coords, (const double*)charges, natoms, box, K, order, alpha,
                              (const double*)xmoduli, (const double*)ymoduli, (const double*)zmoduli, zmoduli_half,
                              scale, (double*)energy_out, (double*)phi_grid_out, w, stream);
  if (dtype == TFF_F32)
    return pme_energy<float>((const float*)coords, (const float*)charges, natoms, box, K, order, alpha,
                             (const float*)xmoduli, (const float*)ymoduli, (const float*)zmoduli, zmoduli_half,
                             scale, (float*)energy_out, (float*)phi_grid_out, w, stream);
  tff_set_error("tff_pme_energy: unsupported dtype %d", dtype);
  return TFF_ERR_ARG;
}

extern "C" int tff_pme_gather(int dtype, const void* phi_grid, const void* coords, const void* charges,
                              long long natoms, const double* box_host, int K0, int K1, int K2, int order,
                              double alpha, double scale, void* epot_out, void* dcoords_out, hipStream_t stream) {
  int K[3] = {K0, K1, K2};
  int rc = check_geom(box_host, K, order);
  if (rc) return rc;
  if (natoms < 0 || (natoms > 0 && (!coords || !charges || !phi_grid))) {
    tff_set_error("tff_pme_gather: invalid arguments");
    return TFF_ERR_ARG;
  }
  TffBox box = tff_make_box(box_host);
  PmeGeom g = make_geom(box, K, order, alpha);
  if (dtype == TFF_F64)
    return launch_gather<double>(order, (const double*)phi_grid, (const double*)coords, (const double*)charges, natoms,
                                 g, scale, (double*)epot_out, (double*)dcoords_out, stream);
  if (dtype == TFF_F32)
    return launch_gather<float>(order, (const float*)phi_grid, (const float*)coords, (const float*)charges, natoms, g,
                                scale, (float*)epot_out, (float*)dcoords_out, stream);
  tff_set_error("tff_pme_gather: unsupported dtype %d", dtype);
  return TFF_ERR_ARG;
}
