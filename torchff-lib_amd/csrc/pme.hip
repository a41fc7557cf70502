// Smooth-PME reciprocal space for point charges, dipoles and quadrupoles on gfx950, float32 or
// float64.
//
// Replaces the reference's pme_long_range / pme_long_range_all ops (ranks 0..2; the multipole
// spread and interpolation are described at k_spread / k_gather below) for rank 0
// (csrc/pme/pme.cu:15-248: spread_q_kernel, torch rfftn, pme_convolution_fused_kernel,
// irfftn, 1/2 sum(Phi Q), self term; backward interpolate_kernel), generalised from
// the hard-wired spline order 6 to orders 3..8.
//
// Pipeline (one K0 x K1 x K2 grid, layout (x, y, z) with z fastest):
//   1. spread      s = K (r . B^-1), m0 = ceil(s - (p mod 2)/2), weights M_p(t + k),
//                  t = m0 - s + p/2 - floor(p/2) in [0, 1), onto grid points m0 - floor(p/2) + k
//                  (k = 0..p-1): atoms binned by the 8^3 brick of their stencil base, one
//                  workgroup per brick summing the reaching atoms in LDS (float64) and writing
//                  its brick.
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

// Per-atom multipoles.  Level 0: charges only; level 1: charges and dipoles; level 2: also
// Cartesian quadrupoles theta (3 x 3 row-major per atom).  An atom contributes to the charge grid
//   Q(m) += q W(m) + p . grad_r W(m) + (1/3) theta : grad_r grad_r W(m)
// (pme_ref.py:145-235).  With u = m0 - s + const, d/ds_a = -d/du_a and ds_a/dr_x = R_ax =
// K_a (B^-1)_{xa}, so p . grad_r W = -sum_a pu_a dW/du_a with pu = R p, and
// theta : grad grad W = sum_ab (R theta R^T)_ab d2W/du_a du_b.
struct Multipoles {
  const void* q;   // N, may be NULL (zero) at levels >= 1
  const void* p;   // N x 3, NULL = zero
  const void* th;  // N x 9, NULL = zero
};

template <typename T>
__device__ __forceinline__ double mp_load(const void* a, long long k) {
  return (double)((const T*)a)[k];
}

// lattice-frame dipole pu_a = sum_x R_ax p_x and quadrupole coefficients tc = (1/3) R theta R^T
// symmetrised: tc[0..2] diagonal aa, tc[3] = ab01, tc[4] = 02, tc[5] = 12 (off-diagonals doubled,
// as each appears twice in the contraction)
template <typename T>
__device__ __forceinline__ void lattice_multipoles(const Multipoles& mp, long long i, const PmeGeom& g, double pu[3],
                                                   double tc[6]) {
  double R[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int x = 0; x < 3; ++x) R[a][x] = (double)g.K[a] * g.inv[3 * x + a];
#pragma unroll
  for (int a = 0; a < 3; ++a) pu[a] = 0.0;
  if (mp.p) {
    const double p0 = mp_load<T>(mp.p, 3 * i), p1 = mp_load<T>(mp.p, 3 * i + 1), p2 = mp_load<T>(mp.p, 3 * i + 2);
#pragma unroll
    for (int a = 0; a < 3; ++a) pu[a] = R[a][0] * p0 + R[a][1] * p1 + R[a][2] * p2;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) tc[k] = 0.0;
  if (mp.th) {
    double th[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) th[k] = mp_load<T>(mp.th, 9 * i + k);
    double tR[3][3];  // theta R^T: tR[x][b] = sum_y th[x][y] R[b][y]
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int b = 0; b < 3; ++b) tR[x][b] = th[3 * x] * R[b][0] + th[3 * x + 1] * R[b][1] + th[3 * x + 2] * R[b][2];
    double U[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) U[a][b] = R[a][0] * tR[0][b] + R[a][1] * tR[1][b] + R[a][2] * tR[2][b];
    const double third = 1.0 / 3.0;
    tc[0] = U[0][0] * third;
    tc[1] = U[1][1] * third;
    tc[2] = U[2][2] * third;
    tc[3] = (U[0][1] + U[1][0]) * third;
    tc[4] = (U[0][2] + U[2][0]) * third;
    tc[5] = (U[1][2] + U[2][1]) * third;
  }
}

// Owner-brick spread.  The grid is cut into bricks of SB^3 points; the atoms are binned by the
// brick of their stencil base (a counting sort: k_brick_bin, k_brick_scan, k_brick_scatter); one
// workgroup per brick then takes the atoms of every brick whose stencils can reach it (per axis the
// bricks holding bases in [start - P + 1, end], at most 4), adds their in-brick contributions in LDS
// (float64) and writes its brick once.  No global atomics, and no clearing of the grid: every
// point is written by its owner.
constexpr int SB = 8;
constexpr int SB_SRC = 4;  // source bricks per axis
// a brick in LDS: x fastest, rows of SB + 1 and planes of (SB + 1) SB + 1 doubles, so that the
// atomics of lanes at other x, y or z fall on different banks
constexpr int SB_Y = SB + 1, SB_Z = SB_Y * SB + 1;

__device__ __forceinline__ void atom_base(const double* r, const PmeGeom& g, int a, int P, int& base, double& t) {
  const double s = (r[0] * g.inv[a] + r[1] * g.inv[3 + a] + r[2] * g.inv[6 + a]) * (double)g.K[a];
  anchor_axis(s, g.K[a], P, base, t);
}

// atomicAdd(&arr[id], 1) for every active lane, one atomic per distinct id in the wave: the lanes
// are grouped by id with ballots (no memory traffic), then the first lane of every group adds the
// group's size, all groups in one atomic instruction; returns the lane's old value + its rank in
// its group
__device__ __forceinline__ int wave_agg_add(int* __restrict__ arr, int id, bool active) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  unsigned long long pending = __builtin_amdgcn_ballot_w64(active);
  int leader_of_mine = lane, rank = 0, size = 0;
  while (pending) {
    const int leader = __builtin_ctzll(pending);
    const int lid = __shfl(id, leader, TFF_WAVE);
    const bool mine = active && id == lid;
    const unsigned long long m = __builtin_amdgcn_ballot_w64(mine);
    if (mine) {
      leader_of_mine = leader;
      rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    }
    if (lane == leader) size = __builtin_popcountll(m);
    pending &= ~m;
  }
  int old = 0;
  if (active && lane == leader_of_mine) old = atomicAdd(&arr[id], size);
  old = __shfl(old, leader_of_mine, TFF_WAVE);
  return old + rank;
}

template <typename T>
__global__ __launch_bounds__(256) void k_brick_bin(const T* __restrict__ coords, long long n, PmeGeom g,
                                                   int* __restrict__ brick_of, int* __restrict__ count) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const bool active = i < n;
  int id = 0;
  if (active) {
    const double r[3] = {(double)coords[3 * i], (double)coords[3 * i + 1], (double)coords[3 * i + 2]};
    int bb[3];
    for (int a = 0; a < 3; ++a) {
      int base;
      double t;
      atom_base(r, g, a, g.order, base, t);
      bb[a] = base / SB;
    }
    const int nb1 = (g.K[1] + SB - 1) / SB, nb2 = (g.K[2] + SB - 1) / SB;
    id = (bb[0] * nb1 + bb[1]) * nb2 + bb[2];
    brick_of[i] = id;
  }
  wave_agg_add(count, id, active);
}

// exclusive scan of count[0..m) into start[0..m], one workgroup of 1024 threads
__global__ __launch_bounds__(1024) void k_brick_scan(const int* __restrict__ count, int m, int* __restrict__ start) {
  __shared__ int s_tot[1024];
  const int t = threadIdx.x;
  const int per = (m + 1023) / 1024;
  const int b = t * per, e = b + per < m ? b + per : m;
  int sum = 0;
  for (int k = b; k < e; ++k) sum += count[k];
  s_tot[t] = sum;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int v = t >= o ? s_tot[t - o] : 0;
    __syncthreads();
    s_tot[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? s_tot[t - 1] : 0;
  for (int k = b; k < e; ++k) {
    start[k] = run;
    run += count[k];
  }
  if (t == 1023) start[m] = s_tot[1023];
}

__global__ __launch_bounds__(256) void k_brick_scatter(long long n, const int* __restrict__ brick_of,
                                                       const int* __restrict__ start, int* __restrict__ cursor,
                                                       int* __restrict__ list) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const bool active = i < n;
  const int id = active ? brick_of[i] : 0;
  const int pos = wave_agg_add(cursor, id, active);
  if (active) list[start[id] + pos] = (int)i;
}

__global__ void k_zero_ints(int* __restrict__ p, int m) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < m; k += gridDim.x * blockDim.x) p[k] = 0;
}

template <typename T, int P, int LVL>
__global__ __launch_bounds__(256) void k_spread(const T* __restrict__ coords, Multipoles mp, PmeGeom g,
                                                const int* __restrict__ start, const int* __restrict__ list,
                                                double* __restrict__ grid) {
  __shared__ double s_b[SB_Z * SB];
  const int K0 = g.K[0], K1 = g.K[1], K2 = g.K[2];
  const int nb[3] = {(K0 + SB - 1) / SB, (K1 + SB - 1) / SB, (K2 + SB - 1) / SB};
  const int bid = blockIdx.x;
  const int B[3] = {bid / (nb[1] * nb[2]), (bid / nb[2]) % nb[1], bid % nb[2]};
  int B0[3], Bn[3];
  for (int a = 0; a < 3; ++a) {
    B0[a] = B[a] * SB;
    Bn[a] = g.K[a] - B0[a] < SB ? g.K[a] - B0[a] : SB;
  }
  for (int k = threadIdx.x; k < SB_Z * SB; k += blockDim.x) s_b[k] = 0.0;
  // source bricks: per axis those holding a stencil base in [B0 - P + 1, B0 + Bn - 1] (periodic);
  // their atom ranges are concatenated so that every thread takes atoms of the whole set
  __shared__ int s_a0[SB_SRC * SB_SRC * SB_SRC], s_pref[SB_SRC * SB_SRC * SB_SRC + 1], s_nc;
  if (threadIdx.x == 0) {
    int src[3][SB_SRC], ns[3];
    for (int a = 0; a < 3; ++a) {
      ns[a] = 0;
      for (int v = B0[a] - (P - 1); v <= B0[a] + Bn[a] - 1; ++v) {
        int w = v % g.K[a];
        w = w < 0 ? w + g.K[a] : w;
        const int sb = w / SB;
        bool seen = false;
        for (int k = 0; k < ns[a]; ++k) seen |= src[a][k] == sb;
        if (!seen && ns[a] < SB_SRC) src[a][ns[a]++] = sb;
      }
    }
    int nc = 0, tot = 0;
    for (int sx = 0; sx < ns[0]; ++sx)
      for (int sy = 0; sy < ns[1]; ++sy)
        for (int sz = 0; sz < ns[2]; ++sz) {
          const int sb = (src[0][sx] * nb[1] + src[1][sy]) * nb[2] + src[2][sz];
          s_a0[nc] = start[sb];
          s_pref[nc] = tot;
          tot += start[sb + 1] - start[sb];
          ++nc;
        }
    s_pref[nc] = tot;
    s_nc = nc;
  }
  __syncthreads();
  constexpr int PD = LVL >= 2 ? P : 1;
  const int ncomb = s_nc, total = s_pref[ncomb];
  {
        for (int k = (int)threadIdx.x; k < total; k += blockDim.x) {
          int c = 0;
          while (s_pref[c + 1] <= k) ++c;
          const int i = list[s_a0[c] + (k - s_pref[c])];
          const double r[3] = {(double)coords[3 * i], (double)coords[3 * i + 1], (double)coords[3 * i + 2]};
          T Mv[3][P], dM[3][P], d2M[3][PD];
          int loc[3][P];
          for (int a = 0; a < 3; ++a) {
            int base;
            double t;
            atom_base(r, g, a, P, base, t);
            if constexpr (LVL >= 2)
              spline_weights2<T, P>((T)t, Mv[a], dM[a], d2M[a]);
            else
              spline_weights<T, P>((T)t, Mv[a], dM[a]);
            int gp = base;
#pragma unroll
            for (int c2 = 0; c2 < P; ++c2) {
              const int l = gp - B0[a];
              loc[a][c2] = l >= 0 && l < Bn[a] ? l : -1;
              gp = gp + 1 == g.K[a] ? 0 : gp + 1;
            }
          }
          if constexpr (LVL == 0) {
            const T qi = ((const T*)mp.q)[i];
            for (int ka = 0; ka < P; ++ka) {
              if (loc[0][ka] < 0) continue;
              const T wa = qi * Mv[0][ka];
              for (int kb = 0; kb < P; ++kb) {
                if (loc[1][kb] < 0) continue;
                const T wab = wa * Mv[1][kb];
                double* row = s_b + loc[1][kb] * SB_Y + loc[0][ka];
#pragma unroll
                for (int kc = 0; kc < P; ++kc)
                  if (loc[2][kc] >= 0) atomicAdd(&row[loc[2][kc] * SB_Z], (double)(wab * Mv[2][kc]));
              }
            }
          } else {
            double pud[3], tcd[6];
            lattice_multipoles<T>(mp, i, g, pud, tcd);
            const T qi = mp.q ? ((const T*)mp.q)[i] : (T)0;
            const T pu0 = (T)pud[0], pu1 = (T)pud[1], pu2 = (T)pud[2];
            T tc[6];
#pragma unroll
            for (int k2 = 0; k2 < 6; ++k2) tc[k2] = (T)tcd[k2];
            for (int ka = 0; ka < P; ++ka) {
              if (loc[0][ka] < 0) continue;
              for (int kb = 0; kb < P; ++kb) {
                if (loc[1][kb] < 0) continue;
                const T xy = Mv[0][ka] * Mv[1][kb], dxy = dM[0][ka] * Mv[1][kb], xdy = Mv[0][ka] * dM[1][kb];
                T c0 = qi * xy - pu0 * dxy - pu1 * xdy;
                T c1 = -pu2 * xy;
                T c2 = (T)0;
                if constexpr (LVL >= 2) {
                  c0 += tc[0] * d2M[0][ka] * Mv[1][kb] + tc[1] * Mv[0][ka] * d2M[1][kb] + tc[3] * dM[0][ka] * dM[1][kb];
                  c1 += tc[4] * dxy + tc[5] * xdy;
                  c2 = tc[2] * xy;
                }
                double* row = s_b + loc[1][kb] * SB_Y + loc[0][ka];
#pragma unroll
                for (int kc = 0; kc < P; ++kc) {
                  if (loc[2][kc] < 0) continue;
                  T v = c0 * Mv[2][kc] + c1 * dM[2][kc];
                  if constexpr (LVL >= 2) v += c2 * d2M[2][kc];
                  atomicAdd(&row[loc[2][kc] * SB_Z], (double)v);
                }
              }
            }
          }
        }
      }
  __syncthreads();
  for (int k = threadIdx.x; k < SB * SB * SB; k += blockDim.x) {
    const int lz = k % SB, ly = (k / SB) % SB, lx = k / (SB * SB);
    if (lx < Bn[0] && ly < Bn[1] && lz < Bn[2])
      grid[((long long)(B0[0] + lx) * K1 + (B0[1] + ly)) * K2 + (B0[2] + lz)] = s_b[lz * SB_Z + ly * SB_Y + lx];
  }
}

// Interpolation of the potential grid at each atom.  With phi_i = sum_m W_i(m) Phi(m) and its
// lattice derivatives g_a, h_ab, c_abc (u-frame), the Cartesian derivatives are
// grad phi = -R^T g, grad grad phi = R^T h R, grad grad grad phi = -R^T R^T R^T c, and
//   dE/dq = phi - 2 a q/sqrt(pi)                                  (potential)
//   -dE/dp = -grad phi + 4 a^3/(3 sqrt(pi)) p                     (field)
//   -3 dE/dtheta = -grad grad phi + 16 a^5/(15 sqrt(pi)) theta    (field gradient)
//   dE/dr = q grad phi + (grad grad phi) p + (1/3) theta : grad grad grad phi
// (pme_ref.py:334-469 conventions; csrc/pme/kernels.cuh:191-470 for the charges-only case).
// Levels: 0 = first derivatives, 1 = + Hessian, 2 = + third derivatives.
template <typename T, int P, int LVL>
__global__ __launch_bounds__(256) void k_gather(const T* __restrict__ phi_grid, const T* __restrict__ coords,
                                                Multipoles mp, long long n, PmeGeom g, double scale,
                                                T* __restrict__ epot, T* __restrict__ field, T* __restrict__ fgrad,
                                                T* __restrict__ dcoords) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r0 = (double)coords[3 * i], r1 = (double)coords[3 * i + 1], r2 = (double)coords[3 * i + 2];
  constexpr int P2 = LVL >= 1 ? P : 1, P3 = LVL >= 2 ? P : 1;
  T Mv[3][P], dM[3][P], d2M[3][P2], d3M[3][P3];
  int base[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double s = (r0 * g.inv[a] + r1 * g.inv[3 + a] + r2 * g.inv[6 + a]) * (double)g.K[a];
    double t;
    anchor_axis(s, g.K[a], P, base[a], t);
    if constexpr (LVL >= 2)
      spline_weights3<T, P>((T)t, Mv[a], dM[a], d2M[a], d3M[a]);
    else if constexpr (LVL == 1)
      spline_weights2<T, P>((T)t, Mv[a], dM[a], d2M[a]);
    else
      spline_weights<T, P>((T)t, Mv[a], dM[a]);
  }
  const int K1 = g.K[1], K2 = g.K[2];
  T phi = 0, du0 = 0, du1 = 0, du2 = 0;
  T h00 = 0, h11 = 0, h22 = 0, h01 = 0, h02 = 0, h12 = 0;
  T c000 = 0, c111 = 0, c222 = 0, c001 = 0, c002 = 0, c011 = 0, c022 = 0, c112 = 0, c122 = 0, c012 = 0;
  int gx = base[0];
  for (int a = 0; a < P; ++a) {
    int gy = base[1];
    for (int b = 0; b < P; ++b) {
      const T* row = phi_grid + ((long long)gx * K1 + gy) * K2;
      int gz = base[2];
      T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
#pragma unroll
      for (int c = 0; c < P; ++c) {
        T v = row[gz];
        s0 += v * Mv[2][c];
        s1 += v * dM[2][c];
        if constexpr (LVL >= 1) s2 += v * d2M[2][c];
        if constexpr (LVL >= 2) s3 += v * d3M[2][c];
        gz = gz + 1 == K2 ? 0 : gz + 1;
      }
      phi += Mv[0][a] * Mv[1][b] * s0;
      du0 += dM[0][a] * Mv[1][b] * s0;
      du1 += Mv[0][a] * dM[1][b] * s0;
      du2 += Mv[0][a] * Mv[1][b] * s1;
      if constexpr (LVL >= 1) {
        h00 += d2M[0][a] * Mv[1][b] * s0;
        h11 += Mv[0][a] * d2M[1][b] * s0;
        h22 += Mv[0][a] * Mv[1][b] * s2;
        h01 += dM[0][a] * dM[1][b] * s0;
        h02 += dM[0][a] * Mv[1][b] * s1;
        h12 += Mv[0][a] * dM[1][b] * s1;
      }
      if constexpr (LVL >= 2) {
        c000 += d3M[0][a] * Mv[1][b] * s0;
        c111 += Mv[0][a] * d3M[1][b] * s0;
        c222 += Mv[0][a] * Mv[1][b] * s3;
        c001 += d2M[0][a] * dM[1][b] * s0;
        c002 += d2M[0][a] * Mv[1][b] * s1;
        c011 += dM[0][a] * d2M[1][b] * s0;
        c022 += dM[0][a] * Mv[1][b] * s2;
        c112 += Mv[0][a] * d2M[1][b] * s1;
        c122 += Mv[0][a] * dM[1][b] * s2;
        c012 += dM[0][a] * dM[1][b] * s1;
      }
      gy = gy + 1 == K1 ? 0 : gy + 1;
    }
    gx = gx + 1 == g.K[0] ? 0 : gx + 1;
  }
  constexpr double RPI = 0.5641895835477563;  // 1/sqrt(pi)
  const double al = g.alpha;
  if constexpr (LVL == 0) {
    // weights depend on t = m0 - s + const, so d/ds = -d/dt; ds_a/dr_j = K_a inv[j][a]
    const double qi = (double)((const T*)mp.q)[i];
    const double d0 = -(double)du0 * g.K[0], d1 = -(double)du1 * g.K[1], d2 = -(double)du2 * g.K[2];
    const double gxr = d0 * g.inv[0] + d1 * g.inv[1] + d2 * g.inv[2];
    const double gyr = d0 * g.inv[3] + d1 * g.inv[4] + d2 * g.inv[5];
    const double gzr = d0 * g.inv[6] + d1 * g.inv[7] + d2 * g.inv[8];
    const double self = -2.0 * al * RPI * qi;
    if (epot) epot[i] = (T)(scale * ((double)phi + self));
    if (field) {
      field[3 * i] = (T)(-scale * gxr);
      field[3 * i + 1] = (T)(-scale * gyr);
      field[3 * i + 2] = (T)(-scale * gzr);
    }
    if (dcoords) {
      dcoords[3 * i] = (T)(scale * qi * gxr);
      dcoords[3 * i + 1] = (T)(scale * qi * gyr);
      dcoords[3 * i + 2] = (T)(scale * qi * gzr);
    }
  } else {
    double R[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int x = 0; x < 3; ++x) R[a][x] = (double)g.K[a] * g.inv[3 * x + a];
    const double qi = mp.q ? mp_load<T>(mp.q, i) : 0.0;
    double pu[3], tc[6];
    lattice_multipoles<T>(mp, i, g, pu, tc);
    const double gu[3] = {(double)du0, (double)du1, (double)du2};
    const double H[3][3] = {{(double)h00, (double)h01, (double)h02},
                            {(double)h01, (double)h11, (double)h12},
                            {(double)h02, (double)h12, (double)h22}};
    // lattice-frame vector v with dE/dr = R^T v:
    //   v_a = -q g_a + sum_b h_ab pu_b - sum_bc c_abc (1/3)(R theta R^T)_bc
    double v[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) v[a] = -qi * gu[a] + H[a][0] * pu[0] + H[a][1] * pu[1] + H[a][2] * pu[2];
    if constexpr (LVL >= 2) {
      // tc holds (1/3) U with off-diagonals doubled, so sum_bc c_abc U_bc / 3 =
      // c_a00 tc0 + c_a11 tc1 + c_a22 tc2 + c_a01 tc3 + c_a02 tc4 + c_a12 tc5
      const double C0[6] = {(double)c000, (double)c011, (double)c022, (double)c001, (double)c002, (double)c012};
      const double C1[6] = {(double)c001, (double)c111, (double)c122, (double)c011, (double)c012, (double)c112};
      const double C2[6] = {(double)c002, (double)c112, (double)c222, (double)c012, (double)c022, (double)c122};
      double s0 = 0, s1 = 0, s2 = 0;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        s0 += C0[k] * tc[k];
        s1 += C1[k] * tc[k];
        s2 += C2[k] * tc[k];
      }
      v[0] -= s0;
      v[1] -= s1;
      v[2] -= s2;
    }
    double grad[3];
#pragma unroll
    for (int x = 0; x < 3; ++x) grad[x] = -(R[0][x] * gu[0] + R[1][x] * gu[1] + R[2][x] * gu[2]);
    if (epot) epot[i] = (T)(scale * ((double)phi - 2.0 * al * RPI * qi));
    if (field) {
      const double cf = 4.0 * al * al * al * RPI / 3.0;
#pragma unroll
      for (int x = 0; x < 3; ++x) field[3 * i + x] = (T)(scale * (-grad[x] + (mp.p ? cf * mp_load<T>(mp.p, 3 * i + x) : 0.0)));
    }
    if (fgrad) {
      // grad grad phi = R^T H R
      const double cg = 16.0 * al * al * al * al * al * RPI / 15.0;
      double HR[3][3];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int y = 0; y < 3; ++y) HR[a][y] = H[a][0] * R[0][y] + H[a][1] * R[1][y] + H[a][2] * R[2][y];
#pragma unroll
      for (int x = 0; x < 3; ++x)
#pragma unroll
        for (int y = 0; y < 3; ++y) {
          const double hxy = R[0][x] * HR[0][y] + R[1][x] * HR[1][y] + R[2][x] * HR[2][y];
          fgrad[9 * i + 3 * x + y] =
              (T)(scale * (-hxy + (mp.th ? cg * mp_load<T>(mp.th, 9 * i + 3 * x + y) : 0.0)));
        }
    }
    if (dcoords) {
#pragma unroll
      for (int x = 0; x < 3; ++x) dcoords[3 * i + x] = (T)(scale * (R[0][x] * v[0] + R[1][x] * v[1] + R[2][x] * v[2]));
    }
  }
}

// self energy -a/sqrt(pi) sum (q^2 + (2 a^2/3) |p|^2 + (8 a^4/45) theta : theta)
// (csrc/ewald/self_contribution.cuh:4-110)
template <typename T>
__global__ void k_self_energy_parts(Multipoles mp, long long n, double alpha, double* part) {
  double acc = 0.0;
  const double a2 = alpha * alpha;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double e = 0.0;
    if (mp.q) {
      const double qi = mp_load<T>(mp.q, i);
      e = qi * qi;
    }
    if (mp.p) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) {
        const double v = mp_load<T>(mp.p, 3 * i + k);
        s += v * v;
      }
      e += 2.0 * a2 / 3.0 * s;
    }
    if (mp.th) {
      double s = 0.0;
      for (int k = 0; k < 9; ++k) {
        const double v = mp_load<T>(mp.th, 9 * i + k);
        s += v * v;
      }
      e += 8.0 * a2 * a2 / 45.0 * s;
    }
    acc += e;
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


constexpr int SELF_BLOCKS = 64;

struct PmeWs {
  double* grid;      // K^3 float64 charge grid
  void* work;        // K^3 complex T
  double* parts;     // energy partials
  int* brick_of;     // natoms: brick of each atom's stencil base
  int* bcount;       // bricks
  int* bstart;       // bricks + 1
  int* bcursor;      // bricks
  int* blist;        // natoms, by brick
  int nbricks;
  size_t bytes;
  int nparts;
};

int brick_count(const int K[3]) {
  return ((K[0] + SB - 1) / SB) * ((K[1] + SB - 1) / SB) * ((K[2] + SB - 1) / SB);
}

PmeWs carve(void* base, int dtype, long long natoms, const int K[3]) {
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
  const size_t na = natoms > 0 ? (size_t)natoms : 1;
  w.nbricks = brick_count(K);
  w.brick_of = (int*)take(4 * na);
  w.bcount = (int*)take(4 * (size_t)w.nbricks);
  w.bstart = (int*)take(4 * ((size_t)w.nbricks + 1));
  w.bcursor = (int*)take(4 * (size_t)w.nbricks);
  w.blist = (int*)take(4 * na);
  w.bytes = off;
  return w;
}

// kernel level of a multipole set: 0 = charges only, 1 = dipoles (or charges absent), 2 = quadrupoles
int spread_level(const Multipoles& mp) { return mp.th ? 2 : (mp.p || !mp.q) ? 1 : 0; }

#define TFF_ORDER_DISPATCH(CALL)                                                            \
  switch (order) {                                                                          \
    case 3: CALL(3); break;                                                                 \
    case 4: CALL(4); break;                                                                 \
    case 5: CALL(5); break;                                                                 \
    case 6: CALL(6); break;                                                                 \
    case 7: CALL(7); break;                                                                 \
    case 8: CALL(8); break;                                                                 \
    default: tff_set_error("PME: unsupported spline order %d (3..8)", order); return TFF_ERR_UNSUPPORTED; \
  }

template <typename T, int L>
int spread_level_launch(int order, unsigned nbricks, const T* coords, const Multipoles& mp, const PmeGeom& g,
                        const PmeWs& w, double* grid, hipStream_t st) {
#define TFF_S(P)                                                                                            \
  hipLaunchKernelGGL((k_spread<T, P, L>), dim3(nbricks), dim3(256), 0, st, coords, mp, g, (const int*)w.bstart, \
                     (const int*)w.blist, grid)
  TFF_ORDER_DISPATCH(TFF_S)
#undef TFF_S
  return TFF_OK;
}

// bins the atoms by brick, then one workgroup per brick writes it (the whole grid, every call)
template <typename T>
int launch_spread(int order, const T* coords, const Multipoles& mp, long long n, const PmeGeom& g, const PmeWs& w,
                  double* grid, hipStream_t st) {
  void* tok = tff_prof_begin("pme_spread", st);
  const int nbk = w.nbricks;
  {
    int zb = (nbk + 255) / 256;
    zb = zb > 4096 ? 4096 : (zb < 1 ? 1 : zb);
    hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)zb), dim3(256), 0, st, w.bcount, nbk);
    hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)zb), dim3(256), 0, st, w.bcursor, nbk);
  }
  if (n > 0) {
    const unsigned nb = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_brick_bin<T>, dim3(nb), dim3(256), 0, st, coords, n, g, w.brick_of, w.bcount);
    hipLaunchKernelGGL(k_brick_scan, dim3(1), dim3(1024), 0, st, (const int*)w.bcount, nbk, w.bstart);
    hipLaunchKernelGGL(k_brick_scatter, dim3(nb), dim3(256), 0, st, n, (const int*)w.brick_of,
                       (const int*)w.bstart, w.bcursor, w.blist);
  } else {
    hipLaunchKernelGGL(k_brick_scan, dim3(1), dim3(1024), 0, st, (const int*)w.bcount, nbk, w.bstart);
  }
  const int lvl = spread_level(mp);
  const unsigned nbr = (unsigned)nbk;
  int rc = lvl == 0   ? spread_level_launch<T, 0>(order, nbr, coords, mp, g, w, grid, st)
           : lvl == 1 ? spread_level_launch<T, 1>(order, nbr, coords, mp, g, w, grid, st)
                      : spread_level_launch<T, 2>(order, nbr, coords, mp, g, w, grid, st);
  if (rc) return rc;
  tff_prof_end(tok, st);
  TFF_CHECK_LAUNCH("k_spread");
  return TFF_OK;
}

template <typename T, int P, int L>
void gather_one(unsigned nb, const T* phi, const T* coords, const Multipoles& mp, long long n, const PmeGeom& g,
                double scale, T* epot, T* field, T* fgrad, T* dcoords, hipStream_t st) {
  if constexpr (L < 2 || P >= 4)  // third derivatives need order >= 4 (checked by the callers)
    hipLaunchKernelGGL((k_gather<T, P, L>), dim3(nb), dim3(256), 0, st, phi, coords, mp, n, g, scale, epot, field,
                       fgrad, dcoords);
}

template <typename T, int L>
int gather_level_launch(int order, unsigned nb, const T* phi, const T* coords, const Multipoles& mp, long long n,
                        const PmeGeom& g, double scale, T* epot, T* field, T* fgrad, T* dcoords, hipStream_t st) {
#define TFF_G(P) gather_one<T, P, L>(nb, phi, coords, mp, n, g, scale, epot, field, fgrad, dcoords, st)
  TFF_ORDER_DISPATCH(TFF_G)
#undef TFF_G
  return TFF_OK;
}

template <typename T>
int launch_gather(int order, const T* phi, const T* coords, const Multipoles& mp, long long n, const PmeGeom& g,
                  double scale, T* epot, T* field, T* fgrad, T* dcoords, hipStream_t st) {
  unsigned nb = (unsigned)((n + 255) / 256);
  if (n == 0) return TFF_OK;
  void* tok = tff_prof_begin("pme_gather", st);
  const int lvl = mp.th ? 2 : (mp.p || !mp.q || fgrad) ? 1 : 0;
  int rc = lvl == 0 ? gather_level_launch<T, 0>(order, nb, phi, coords, mp, n, g, scale, epot, field, fgrad, dcoords, st)
           : lvl == 1
               ? gather_level_launch<T, 1>(order, nb, phi, coords, mp, n, g, scale, epot, field, fgrad, dcoords, st)
               : gather_level_launch<T, 2>(order, nb, phi, coords, mp, n, g, scale, epot, field, fgrad, dcoords, st);
  if (rc) return rc;
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
int pme_energy(const T* coords, const Multipoles& mp, long long n, const TffBox& box, const int K[3], int order, double alpha,
               const T* mx, const T* my, const T* mz, int mz_half, double scale, T* energy, T* phi_out,
               const PmeWs& w, hipStream_t st) {
  FftPlan pl[3];
  for (int a = 0; a < 3; ++a)
    if (!make_plan(K[a], pl[a])) {
      tff_set_error("PME: grid size %d is not of the form 2^a 3^b 5^c", K[a]);
      return TFF_ERR_UNSUPPORTED;
    }
  PmeGeom g = make_geom(box, K, order, alpha);
  int rc = launch_spread<T>(order, coords, mp, n, g, w, w.grid, st);
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
  hipLaunchKernelGGL(k_self_energy_parts<T>, dim3(SELF_BLOCKS), dim3(256), 0, st, mp, n, alpha,
                     w.parts + cxp.tiles);
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
  int K[3] = {K0, K1, K2};
  return carve(nullptr, dtype, natoms, K).bytes;
}

namespace {

int pme_energy_entry(const char* name, bool need_q, int dtype, const void* coords, const Multipoles& mp,
                     long long natoms,
                     const double* box_host, int K0, int K1, int K2, int order, double alpha, const void* xmoduli,
                     const void* ymoduli, const void* zmoduli, int zmoduli_half, double scale, void* energy_out,
                     void* phi_grid_out, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  int K[3] = {K0, K1, K2};
  int rc = check_geom(box_host, K, order);
  if (rc) return rc;
  if (natoms < 0 || (natoms > 0 && (!coords || (need_q && !mp.q))) || !xmoduli || !ymoduli || !zmoduli ||
      !phi_grid_out) {
    tff_set_error("%s: invalid arguments", name);
    return TFF_ERR_ARG;
  }
  if (mp.th && order < 4) {
    tff_set_error("%s: quadrupoles need spline order >= 4 (got %d)", name, order);
    return TFF_ERR_UNSUPPORTED;
  }
  size_t need = tff_pme_workspace_bytes(dtype, natoms, K0, K1, K2);
  if (workspace_bytes < need || !workspace) {
    tff_set_error("%s: workspace too small (%zu < %zu)", name, workspace_bytes, need);
    return TFF_ERR_WORKSPACE;
  }
  TffBox box = tff_make_box(box_host);
  PmeWs w = carve(workspace, dtype, natoms, K);
  if (dtype == TFF_F64)
    return pme_energy<double>((const double*)coords, mp, natoms, box, K, order, alpha, (const double*)xmoduli,
                              (const double*)ymoduli, (const double*)zmoduli, zmoduli_half, scale,
                              (double*)energy_out, (double*)phi_grid_out, w, stream);
  if (dtype == TFF_F32)
    return pme_energy<float>((const float*)coords, mp, natoms, box, K, order, alpha, (const float*)xmoduli,
                             (const float*)ymoduli, (const float*)zmoduli, zmoduli_half, scale, (float*)energy_out,
                             (float*)phi_grid_out, w, stream);
  tff_set_error("%s: unsupported dtype %d", name, dtype);
  return TFF_ERR_ARG;
}

int pme_gather_entry(const char* name, bool need_q, int dtype, const void* phi_grid, const void* coords,
                     const Multipoles& mp,
                     long long natoms, const double* box_host, int K0, int K1, int K2, int order, double alpha,
                     double scale, void* epot_out, void* field_out, void* fgrad_out, void* dcoords_out,
                     hipStream_t stream) {
  int K[3] = {K0, K1, K2};
  int rc = check_geom(box_host, K, order);
  if (rc) return rc;
  if (natoms < 0 || (natoms > 0 && (!coords || !phi_grid || (need_q && !mp.q)))) {
    tff_set_error("%s: invalid arguments", name);
    return TFF_ERR_ARG;
  }
  if (mp.th && order < 4) {
    tff_set_error("%s: quadrupoles need spline order >= 4 (got %d)", name, order);
    return TFF_ERR_UNSUPPORTED;
  }
  TffBox box = tff_make_box(box_host);
  PmeGeom g = make_geom(box, K, order, alpha);
  if (dtype == TFF_F64)
    return launch_gather<double>(order, (const double*)phi_grid, (const double*)coords, mp, natoms, g, scale,
                                 (double*)epot_out, (double*)field_out, (double*)fgrad_out, (double*)dcoords_out,
                                 stream);
  if (dtype == TFF_F32)
    return launch_gather<float>(order, (const float*)phi_grid, (const float*)coords, mp, natoms, g, scale,
                                (float*)epot_out, (float*)field_out, (float*)fgrad_out, (float*)dcoords_out, stream);
  tff_set_error("%s: unsupported dtype %d", name, dtype);
  return TFF_ERR_ARG;
}

}  // namespace

extern "C" int tff_pme_energy(int dtype, const void* coords, const void* charges, long long natoms,
                              const double* box_host, int K0, int K1, int K2, int order, double alpha,
                              const void* xmoduli, const void* ymoduli, const void* zmoduli, int zmoduli_half,
                              double scale, void* energy_out, void* phi_grid_out, void* workspace,
                              size_t workspace_bytes, hipStream_t stream) {
  const Multipoles mp{charges, nullptr, nullptr};
  return pme_energy_entry("tff_pme_energy", true, dtype, coords, mp, natoms, box_host, K0, K1, K2, order, alpha,
                          xmoduli,
                          ymoduli, zmoduli, zmoduli_half, scale, energy_out, phi_grid_out, workspace, workspace_bytes,
                          stream);
}

extern "C" int tff_pme_gather(int dtype, const void* phi_grid, const void* coords, const void* charges,
                              long long natoms, const double* box_host, int K0, int K1, int K2, int order,
                              double alpha, double scale, void* epot_out, void* dcoords_out, hipStream_t stream) {
  const Multipoles mp{charges, nullptr, nullptr};
  return pme_gather_entry("tff_pme_gather", true, dtype, phi_grid, coords, mp, natoms, box_host, K0, K1, K2, order,
                          alpha,
                          scale, epot_out, nullptr, nullptr, dcoords_out, stream);
}

extern "C" int tff_pme_energy_multipole(int dtype, const void* coords, const void* charges, const void* dipoles,
                                        const void* quadrupoles, long long natoms, const double* box_host, int K0,
                                        int K1, int K2, int order, double alpha, const void* xmoduli,
                                        const void* ymoduli, const void* zmoduli, int zmoduli_half, double scale,
                                        void* energy_out, void* phi_grid_out, void* workspace,
                                        size_t workspace_bytes, hipStream_t stream) {
  const Multipoles mp{charges, dipoles, quadrupoles};
  return pme_energy_entry("tff_pme_energy_multipole", !dipoles && !quadrupoles, dtype, coords, mp, natoms, box_host,
                          K0, K1, K2, order, alpha,
                          xmoduli, ymoduli, zmoduli, zmoduli_half, scale, energy_out, phi_grid_out, workspace,
                          workspace_bytes, stream);
}

extern "C" int tff_pme_gather_multipole(int dtype, const void* phi_grid, const void* coords, const void* charges,
                                        const void* dipoles, const void* quadrupoles, long long natoms,
                                        const double* box_host, int K0, int K1, int K2, int order, double alpha,
                                        double scale, void* epot_out, void* field_out, void* field_grad_out,
                                        void* dcoords_out, hipStream_t stream) {
  const Multipoles mp{charges, dipoles, quadrupoles};
  return pme_gather_entry("tff_pme_gather_multipole", !dipoles && !quadrupoles, dtype, phi_grid, coords, mp, natoms,
                          box_host, K0, K1, K2,
                          order, alpha, scale, epot_out, field_out, field_grad_out, dcoords_out, stream);
}
