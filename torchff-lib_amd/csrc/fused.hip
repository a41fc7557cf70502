// Fused energy + forces for fixed-charge LJ + Ewald (real space and PME) on gfx950.
//
// One call evaluates, for an orthorhombic periodic box,
//   E = sum_{i<j, r<=rc, not excluded} [ k q_i q_j erfc(a r)/r + 4 eps_ij s^6 (s^6 - 1) ]
//     + sum_{i<j excluded, r<=rc} k q_i q_j (erfc(a r) - 1)/r
//     + k (E_recip + E_self)                      (smooth PME, any order 3..8)
// with Lorentz-Berthelot mixing, and F = -dE/dr.  It is the composition of the
// reference's MultipolarInteraction(rank 0, with pairs_excl), Vdw(LennardJones)
// and PME(rank 0) (examples/tip3p/model.py:100-123), evaluated without an explicit
// pair list.  Every excluded pair (i != j) of the CSR exclusion list receives the
// erfc - 1 correction, which for rigid water is exactly the reference's
// coulomb_excl_pairs set (examples/tip3p/md.py:152-158).
//
// Data layout: atoms are binned into x-y columns of about one cluster width and
// fine z bins, counting-sorted, and cut per column into clusters of 32
// consecutive atoms (padded).  Cluster arrays (TC4 x, y, z, q and TC2 sigma/2,
// sqrt(eps)) are cluster-major, 32 slots per cluster.  Cluster pairs within
// the list cutoff (bounding-box test) are found per column neighbourhood; the
// pair kernel gives each 64-lane wavefront one 32 x 32 tile at a time: lane l
// owns i-slot l & 31 and half l >> 5 of the j-slots, which rotate through the
// tile so that every j accumulator in LDS is written by one lane per step.
// Forces and energies accumulate as int64 fixed point (bitwise independent of
// the order of summation); the reciprocal-space energy is a float64 k-space sum.
// TC = float is the mixed-precision mode, TC = double the all-float64 mode.

#include "pme_core.h"
#include "tff_common.h"
#include "torchff_hip.h"

namespace {

using namespace tffpme;

constexpr int CL = 32;          // atoms per cluster
constexpr int MAXEX = 8;        // exclusion partners per atom kept for the tile masks
constexpr int FB = 256;         // block size of the per-atom kernels
constexpr int TILE_BLOCK = 256; // 4 waves
constexpr int TILE_WAVES = TILE_BLOCK / TFF_WAVE;
constexpr double F_SCALE = 4294967296.0;        // 2^32: forces and energies (kJ/mol, nm)
constexpr double Q_SCALE = 1099511627776.0;     // 2^40: PME charge grid

template <typename T> struct V4 { T x, y, z, w; };
template <typename T> struct V2 { T x, y; };

struct Geo {
  double L[3], invL[3];
  int ncol[2];
  int nzb;
  double colw[2], zbw;
  int nkeys;
  int ncl_max;
  long long pair_cap;
  int maxn;  // cluster-pair segment length per i-cluster
  int reach[2];
};

__device__ __forceinline__ long long to_fx(double v, double s) { return __double2ll_rn(v * s); }

// ---------------------------------------------------------------- 1. binning
template <typename TI>
__global__ __launch_bounds__(FB) void k_bin_atoms(const TI* __restrict__ coords, long long n, Geo g,
                                                  int* __restrict__ key_of, int* __restrict__ key_count) {
  long long i = blockIdx.x * (long long)FB + threadIdx.x;
  if (i >= n) return;
  double w[3];
  for (int a = 0; a < 3; ++a) {
    double f = (double)coords[3 * i + a] * g.invL[a];
    f -= floor(f);
    w[a] = f * g.L[a];
  }
  int cx = (int)(w[0] / g.colw[0]); cx = cx < 0 ? 0 : (cx >= g.ncol[0] ? g.ncol[0] - 1 : cx);
  int cy = (int)(w[1] / g.colw[1]); cy = cy < 0 ? 0 : (cy >= g.ncol[1] ? g.ncol[1] - 1 : cy);
  int cz = (int)(w[2] / g.zbw);     cz = cz < 0 ? 0 : (cz >= g.nzb ? g.nzb - 1 : cz);
  int key = (cx * g.ncol[1] + cy) * g.nzb + cz;
  key_of[i] = key;
  atomicAdd(&key_count[key], 1);
}

// exclusive scan (single workgroup); also counts clusters per column and their offsets
__global__ __launch_bounds__(1024) void k_scan_keys(const int* __restrict__ cnt, int m, int* __restrict__ start) {
  __shared__ int s[1024];
  int t = threadIdx.x;
  int per = (m + 1023) / 1024;
  int b = t * per, e = b + per < m ? b + per : m;
  int sum = 0;
  for (int k = b; k < e; ++k) sum += cnt[k];
  s[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? s[t - 1] : 0;
  for (int k = b; k < e; ++k) {
    start[k] = run;
    run += cnt[k];
  }
  if (t == 1023) start[m] = s[1023];
}

// clusters per column -> cluster offsets per column (single workgroup scan over columns)
__global__ __launch_bounds__(1024) void k_scan_columns(const int* __restrict__ key_start, int ncols, int nzb,
                                                       int* __restrict__ col_cl_start, int* __restrict__ n_clusters) {
  __shared__ int s[1024];
  int t = threadIdx.x;
  int per = (ncols + 1023) / 1024;
  int b = t * per, e = b + per < ncols ? b + per : ncols;
  int sum = 0;
  for (int c = b; c < e; ++c) {
    int na = key_start[(c + 1) * nzb] - key_start[c * nzb];
    sum += (na + CL - 1) / CL;
  }
  s[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? s[t - 1] : 0;
  for (int c = b; c < e; ++c) {
    col_cl_start[c] = run;
    int na = key_start[(c + 1) * nzb] - key_start[c * nzb];
    run += (na + CL - 1) / CL;
  }
  if (t == 1023) {
    col_cl_start[ncols] = s[1023];
    *n_clusters = s[1023];
  }
}

__global__ __launch_bounds__(FB) void k_scatter_atoms(long long n, const int* __restrict__ key_of,
                                                      const int* __restrict__ key_start, int* __restrict__ cursor,
                                                      int* __restrict__ sorted_orig, int* __restrict__ pos_of) {
  long long i = blockIdx.x * (long long)FB + threadIdx.x;
  if (i >= n) return;
  int k = key_of[i];
  int p = key_start[k] + atomicAdd(&cursor[k], 1);
  sorted_orig[p] = (int)i;
  pos_of[i] = p;
}

// slot index (cluster * 32 + s) of sorted position p
__device__ __forceinline__ int slot_of_pos(int p, int col, const int* key_start, const int* col_cl_start, int nzb) {
  int cs = key_start[col * nzb];
  int r = p - cs;
  return (col_cl_start[col] + r / CL) * CL + (r % CL);
}

__device__ __forceinline__ int column_of_key(int key, int nzb) { return key / nzb; }

// ---------------------------------------------------------------- 2. cluster arrays
// One thread per slot of the padded cluster arrays (ncl * 32 slots).
template <typename TI, typename TC>
__global__ __launch_bounds__(FB) void k_fill_clusters(const TI* __restrict__ coords, const TI* __restrict__ q,
                                                      const TI* __restrict__ sig, const TI* __restrict__ eps,
                                                      Geo g, const int* __restrict__ ncl_ptr,
                                                      const int* __restrict__ col_cl_start,
                                                      const int* __restrict__ key_start,
                                                      const int* __restrict__ sorted_orig,
                                                      V4<TC>* __restrict__ xyzq, V2<TC>* __restrict__ lj,
                                                      int* __restrict__ slot_orig, int* __restrict__ cl_col) {
  int sl = blockIdx.x * FB + threadIdx.x;
  int ncl = *ncl_ptr;
  if (sl >= ncl * CL) return;
  int c = sl / CL, s = sl % CL;
  // column of cluster c: binary search in col_cl_start
  int lo = 0, hi = g.ncol[0] * g.ncol[1];
  while (hi - lo > 1) {
    int mid = (lo + hi) / 2;
    if (col_cl_start[mid] <= c) lo = mid; else hi = mid;
  }
  int col = lo;
  if (s == 0) cl_col[c] = col;
  int cs = key_start[col * g.nzb], ce = key_start[(col + 1) * g.nzb];
  int p = cs + (c - col_cl_start[col]) * CL + s;
  if (p < ce) {
    int i = sorted_orig[p];
    double w[3];
    for (int a = 0; a < 3; ++a) {
      double f = (double)coords[3 * i + a] * g.invL[a];
      f -= floor(f);
      w[a] = f * g.L[a];
    }
    xyzq[sl] = {(TC)w[0], (TC)w[1], (TC)w[2], (TC)q[i]};
    lj[sl] = {(TC)(0.5 * (double)sig[i]), (TC)sqrt((double)eps[i] > 0 ? (double)eps[i] : 0.0)};
    slot_orig[sl] = i;
  } else {
    xyzq[sl] = {(TC)0, (TC)0, (TC)0, (TC)0};
    lj[sl] = {(TC)0, (TC)0};
    slot_orig[sl] = -1;
  }
}

// exclusion partners of every slot as packed (cluster << 5 | slot), -1 = none
__global__ __launch_bounds__(FB) void k_fill_exclusions(long long n, const int* __restrict__ pos_of,
                                                        const int* __restrict__ key_of, Geo g,
                                                        const int* __restrict__ key_start,
                                                        const int* __restrict__ col_cl_start,
                                                        const int64_t* __restrict__ row_ptr,
                                                        const int64_t* __restrict__ col,
                                                        int* __restrict__ excl_part, int* __restrict__ status) {
  long long i = blockIdx.x * (long long)FB + threadIdx.x;
  if (i >= n) return;
  int si = slot_of_pos(pos_of[i], column_of_key(key_of[i], g.nzb), key_start, col_cl_start, g.nzb);
  int k = 0;
  if (col) {
    for (long long e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      long long j = col[e];
      if (j == i || j < 0 || j >= n) continue;
      if (k == MAXEX) {
        atomicOr(status, TFF_FUSED_EXCL_OVERFLOW);
        break;
      }
      excl_part[si * MAXEX + k++] =
          slot_of_pos(pos_of[j], column_of_key(key_of[j], g.nzb), key_start, col_cl_start, g.nzb);
    }
  }
  for (; k < MAXEX; ++k) excl_part[si * MAXEX + k] = -1;
}

// padding slots carry no exclusions
__global__ void k_clear_pad_exclusions(const int* __restrict__ ncl_ptr, const int* __restrict__ slot_orig,
                                       int* __restrict__ excl_part) {
  int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= *ncl_ptr * CL) return;
  if (slot_orig[sl] < 0)
    for (int k = 0; k < MAXEX; ++k) excl_part[sl * MAXEX + k] = -1;
}

// cluster bounding boxes (centre, half extent): one wave per cluster, lanes 0..31
template <typename TC>
__global__ __launch_bounds__(FB) void k_bbox(const int* __restrict__ ncl_ptr, const V4<TC>* __restrict__ xyzq,
                                             const int* __restrict__ slot_orig, float4* __restrict__ bb_c,
                                             float4* __restrict__ bb_h) {
  int c = (blockIdx.x * FB + threadIdx.x) / TFF_WAVE;
  int lane = threadIdx.x & (TFF_WAVE - 1);
  if (c >= *ncl_ptr) return;
  int sl = c * CL + (lane & (CL - 1));
  bool v = slot_orig[sl] >= 0;
  V4<TC> p = xyzq[sl];
  float mn[3], mx[3];
  float pv[3] = {(float)p.x, (float)p.y, (float)p.z};
  for (int a = 0; a < 3; ++a) {
    mn[a] = v ? pv[a] : 3.0e38f;
    mx[a] = v ? pv[a] : -3.0e38f;
    for (int off = 16; off > 0; off >>= 1) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, TFF_WAVE));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, TFF_WAVE));
    }
  }
  if (lane == 0) {
    bb_c[c] = make_float4(0.5f * (mn[0] + mx[0]), 0.5f * (mn[1] + mx[1]), 0.5f * (mn[2] + mx[2]), 0.f);
    // widen by a relative epsilon so float rounding of the test never drops a pair
    bb_h[c] = make_float4(0.5f * (mx[0] - mn[0]) * 1.0001f + 1e-6f, 0.5f * (mx[1] - mn[1]) * 1.0001f + 1e-6f,
                          0.5f * (mx[2] - mn[2]) * 1.0001f + 1e-6f, 0.f);
  }
}

// ---------------------------------------------------------------- 3. cluster pairs
// One wave per i-cluster; lanes sweep the clusters of the neighbouring columns in a fixed
// order and write the accepted j-clusters (cj >= ci) into the i-cluster's own segment.
__global__ __launch_bounds__(FB) void k_cluster_pairs(const int* __restrict__ ncl_ptr, const int* __restrict__ cl_col,
                                                      const int* __restrict__ col_cl_start,
                                                      const float4* __restrict__ bb_c,
                                                      const float4* __restrict__ bb_h, Geo g, float rlist2,
                                                      int* __restrict__ seg, int* __restrict__ seg_count,
                                                      int* __restrict__ status) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int w = threadIdx.x / TFF_WAVE;
  const int ci = blockIdx.x * (FB / TFF_WAVE) + w;
  const int ncl = *ncl_ptr;
  if (ci >= ncl) return;
  const int col = cl_col[ci];
  const int cx = col / g.ncol[1], cy = col % g.ncol[1];
  const float4 c0 = bb_c[ci], h0 = bb_h[ci];
  const float Lx = (float)g.L[0], Ly = (float)g.L[1], Lz = (float)g.L[2];
  const float iLx = (float)g.invL[0], iLy = (float)g.invL[1], iLz = (float)g.invL[2];
  int nseg = 0;
  int* my = seg + (long long)ci * g.maxn;
  const int nx = 2 * g.reach[0] + 1 >= g.ncol[0] ? g.ncol[0] : 2 * g.reach[0] + 1;
  const int ny = 2 * g.reach[1] + 1 >= g.ncol[1] ? g.ncol[1] : 2 * g.reach[1] + 1;
  const int x0 = nx == g.ncol[0] ? 0 : cx - g.reach[0];
  const int y0 = ny == g.ncol[1] ? 0 : cy - g.reach[1];
  for (int ix = 0; ix < nx; ++ix) {
    int xx = x0 + ix;
    xx += xx < 0 ? g.ncol[0] : (xx >= g.ncol[0] ? -g.ncol[0] : 0);
    for (int iy = 0; iy < ny; ++iy) {
      int yy = y0 + iy;
      yy += yy < 0 ? g.ncol[1] : (yy >= g.ncol[1] ? -g.ncol[1] : 0);
      int ncol_id = xx * g.ncol[1] + yy;
      int cs = col_cl_start[ncol_id], ce = col_cl_start[ncol_id + 1];
      for (int b = cs; b < ce; b += TFF_WAVE) {
        int cj = b + lane;
        bool ok = false;
        if (cj < ce && cj >= ci) {
          float4 c1 = bb_c[cj], h1 = bb_h[cj];
          float dx = c0.x - c1.x, dy = c0.y - c1.y, dz = c0.z - c1.z;
          dx -= Lx * rintf(dx * iLx);
          dy -= Ly * rintf(dy * iLy);
          dz -= Lz * rintf(dz * iLz);
          float gx = fmaxf(fabsf(dx) - h0.x - h1.x, 0.f);
          float gy = fmaxf(fabsf(dy) - h0.y - h1.y, 0.f);
          float gz = fmaxf(fabsf(dz) - h0.z - h1.z, 0.f);
          ok = gx * gx + gy * gy + gz * gz <= rlist2;
        }
        unsigned long long m = __ballot(ok);
        if (ok) {
          int k = nseg + __popcll(m & ((1ull << lane) - 1ull));
          if (k < g.maxn) my[k] = cj;
        }
        nseg += __popcll(m);
      }
    }
  }
  if (lane == 0) {
    if (nseg > g.maxn) atomicOr(status, TFF_FUSED_PAIR_OVERFLOW);
    seg_count[ci] = nseg < g.maxn ? nseg : g.maxn;
  }
}

// exclusive scan of the per-cluster pair counts (single workgroup), total into *ncp
__global__ __launch_bounds__(1024) void k_scan_counts(const int* __restrict__ ncl_ptr, const int* __restrict__ cnt,
                                                      int* __restrict__ start, int* __restrict__ ncp) {
  __shared__ int s[1024];
  const int m = *ncl_ptr;
  int t = threadIdx.x;
  int per = (m + 1023) / 1024;
  int b = t * per, e = b + per < m ? b + per : m;
  int sum = 0;
  for (int k = b; k < e; ++k) sum += cnt[k];
  s[t] = sum;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? s[t - off] : 0;
    __syncthreads();
    s[t] += v;
    __syncthreads();
  }
  int run = t > 0 ? s[t - 1] : 0;
  for (int k = b; k < e; ++k) {
    start[k] = run;
    run += cnt[k];
  }
  if (t == 1023) *ncp = s[1023];
}

// segments -> one compact cluster-pair list in (ci, segment order): deterministic
__global__ __launch_bounds__(FB) void k_compact_pairs(const int* __restrict__ ncl_ptr, const int* __restrict__ seg,
                                                      const int* __restrict__ seg_count,
                                                      const int* __restrict__ seg_start, int maxn,
                                                      int2* __restrict__ pairs) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int ci = (blockIdx.x * FB + threadIdx.x) / TFF_WAVE;
  if (ci >= *ncl_ptr) return;
  const int n = seg_count[ci], o = seg_start[ci];
  for (int k = lane; k < n; k += TFF_WAVE) pairs[o + k] = make_int2(ci, seg[(long long)ci * maxn + k]);
}

// order the atoms of each sort key by their original index (the scatter used atomics)
__global__ __launch_bounds__(FB) void k_sort_keys(int nkeys, const int* __restrict__ key_start,
                                                  int* __restrict__ sorted_orig, int* __restrict__ pos_of) {
  int k = blockIdx.x * FB + threadIdx.x;
  if (k >= nkeys) return;
  const int b = key_start[k], e = key_start[k + 1];
  for (int p = b + 1; p < e; ++p) {
    int v = sorted_orig[p];
    int q = p - 1;
    while (q >= b && sorted_orig[q] > v) {
      sorted_orig[q + 1] = sorted_orig[q];
      --q;
    }
    sorted_orig[q + 1] = v;
  }
  for (int p = b; p < e; ++p) pos_of[sorted_orig[p]] = p;
}

// ---------------------------------------------------------------- 4. tile kernel
template <typename TC>
struct TileArgs {
  TC L[3], invL[3];
  TC rc2;
  TC alpha;
  TC coul;
};

template <typename TC>
__device__ __forceinline__ void flush_i(int ci, const TC* f, int lane, const int* slot_orig,
                                        unsigned long long* facc) {
  // lanes 0..31 hold the combined i-slot forces
  if (lane < CL) {
    int sl = ci * CL + lane;
    if (slot_orig[sl] >= 0) {
      atomicAdd(&facc[3 * sl + 0], (unsigned long long)to_fx((double)f[0], F_SCALE));
      atomicAdd(&facc[3 * sl + 1], (unsigned long long)to_fx((double)f[1], F_SCALE));
      atomicAdd(&facc[3 * sl + 2], (unsigned long long)to_fx((double)f[2], F_SCALE));
    }
  }
}

template <typename TC>
__global__ __launch_bounds__(TILE_BLOCK) void k_tiles(const int2* __restrict__ cpairs, const int* __restrict__ ncp_ptr,
                                                      long long cap, const V4<TC>* __restrict__ xyzq,
                                                      const V2<TC>* __restrict__ lj, const int* __restrict__ slot_orig,
                                                      const int* __restrict__ excl_part, TileArgs<TC> ta,
                                                      unsigned long long* __restrict__ facc,
                                                      unsigned long long* __restrict__ eacc) {
  __shared__ V4<TC> s_xyzq[TILE_WAVES][CL];
  __shared__ V2<TC> s_lj[TILE_WAVES][CL];
  __shared__ TC s_fj[TILE_WAVES][2][3][CL];
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int w = threadIdx.x / TFF_WAVE;
  const int is = lane & (CL - 1);
  const int half = lane >> 5;
  long long ncp = *ncp_ptr;
  if (ncp > cap) ncp = cap;
  const long long nw = (long long)gridDim.x * TILE_WAVES;
  const long long gw = (long long)blockIdx.x * TILE_WAVES + w;
  // contiguous chunk of tiles per wave so consecutive tiles share the i-cluster
  const long long per = (ncp + nw - 1) / nw;
  const long long t0 = gw * per;
  long long t1 = t0 + per;
  if (t1 > ncp) t1 = ncp;
  const TC two_over_sqrt_pi = (TC)1.1283791670955126;
  TC e_lj = 0, e_c = 0;
  int cur_i = -1;
  TC fi[3] = {0, 0, 0};
  V4<TC> pi = {0, 0, 0, 0};
  V2<TC> li = {0, 0};
  bool vi = false;
  int ex[MAXEX];
  for (long long t = t0; t < t1; ++t) {
    const int2 cp = cpairs[t];
    const int ci = cp.x, cj = cp.y;
    if (ci != cur_i) {
      if (cur_i >= 0) {
        // combine the two halves of each i-slot and flush
        for (int a = 0; a < 3; ++a) fi[a] += __shfl_xor(fi[a], 32, TFF_WAVE);
        flush_i<TC>(cur_i, fi, lane, slot_orig, facc);
      }
      cur_i = ci;
      fi[0] = fi[1] = fi[2] = 0;
      const int sl = ci * CL + is;
      pi = xyzq[sl];
      li = lj[sl];
      vi = slot_orig[sl] >= 0;
#pragma unroll
      for (int k = 0; k < MAXEX; ++k) ex[k] = excl_part[sl * MAXEX + k];
    }
    // stage the j cluster and clear the j accumulators
    if (lane < CL) {
      s_xyzq[w][lane] = xyzq[cj * CL + lane];
      V2<TC> l = lj[cj * CL + lane];
      if (slot_orig[cj * CL + lane] < 0) l.y = (TC)-1;  // marks padding
      s_lj[w][lane] = l;
    }
    for (int a = 0; a < 3; ++a) s_fj[w][half][a][is] = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    unsigned int exmask = 0;
#pragma unroll
    for (int k = 0; k < MAXEX; ++k)
      if (ex[k] >= 0 && (ex[k] >> 5) == cj) exmask |= 1u << (ex[k] & 31);
    const bool diag = ci == cj;
#pragma unroll 4
    for (int k = 0; k < CL / 2; ++k) {
      const int js = (is + 16 * half + k) & (CL - 1);
      const V4<TC> pj = s_xyzq[w][js];
      const V2<TC> lJ = s_lj[w][js];
      TC dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      dx -= ta.L[0] * rint(dx * ta.invL[0]);
      dy -= ta.L[1] * rint(dy * ta.invL[1]);
      dz -= ta.L[2] * rint(dz * ta.invL[2]);
      const TC r2 = dx * dx + dy * dy + dz * dz;
      bool ok = vi && lJ.y >= (TC)0 && r2 <= ta.rc2 && (!diag || js > is);
      TC fs = 0;
      if (ok) {
        const bool excl = (exmask >> js) & 1u;
        const TC rinv = (TC)1 / sqrt(r2);
        const TC rinv2 = rinv * rinv;
        const TC r = r2 * rinv;
        const TC u = ta.alpha * r;
        const TC ec = erfc(u) - (excl ? (TC)1 : (TC)0);
        const TC ex2 = exp(-u * u) * two_over_sqrt_pi * u;
        const TC qq = ta.coul * pi.w * pj.w;
        e_c += qq * ec * rinv;
        fs = qq * (ec + ex2) * rinv * rinv2;
        if (!excl) {
          const TC sg = li.x + lJ.x;
          const TC ep = li.y * lJ.y;
          const TC s2 = sg * sg * rinv2;
          const TC s6 = s2 * s2 * s2;
          e_lj += (TC)4 * ep * s6 * (s6 - (TC)1);
          fs += (TC)24 * ep * s6 * ((TC)2 * s6 - (TC)1) * rinv2;
        }
      }
      // F_i += fs d, F_j -= fs d   (fs = -dE/dr / r)
      fi[0] += fs * dx;
      fi[1] += fs * dy;
      fi[2] += fs * dz;
      s_fj[w][half][0][js] -= fs * dx;
      s_fj[w][half][1][js] -= fs * dy;
      s_fj[w][half][2][js] -= fs * dz;
      // the next step's lane for this j-slot is another lane: keep the LDS read-modify-writes of
      // consecutive steps in program order (a wave's LDS operations execute in issue order)
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < CL) {
      const int sl = cj * CL + lane;
      if (slot_orig[sl] >= 0) {
        for (int a = 0; a < 3; ++a) {
          TC v = s_fj[w][0][a][lane] + s_fj[w][1][a][lane];
          if (v != (TC)0) atomicAdd(&facc[3 * sl + a], (unsigned long long)to_fx((double)v, F_SCALE));
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (cur_i >= 0) {
    for (int a = 0; a < 3; ++a) fi[a] += __shfl_xor(fi[a], 32, TFF_WAVE);
    flush_i<TC>(cur_i, fi, lane, slot_orig, facc);
  }
  // energies: wave sums in double, then fixed point
  double el = tff_wave_sum((double)e_lj), ecl = tff_wave_sum((double)e_c);
  if (lane == 0 && t1 > t0) {
    atomicAdd(&eacc[0], (unsigned long long)to_fx(el, F_SCALE));
    atomicAdd(&eacc[1], (unsigned long long)to_fx(ecl, F_SCALE));
  }
}

// ---------------------------------------------------------------- 5. PME on cluster slots
template <typename TC, int P>
__global__ __launch_bounds__(FB) void k_spread_slots(const int* __restrict__ ncl_ptr, const V4<TC>* __restrict__ xyzq,
                                                     const int* __restrict__ slot_orig, PmeGeom g,
                                                     unsigned long long* __restrict__ grid) {
  int sl = blockIdx.x * FB + threadIdx.x;
  if (sl >= *ncl_ptr * CL || slot_orig[sl] < 0) return;
  V4<TC> p = xyzq[sl];
  TC Mv[3][P], dM[3][P];
  int base[3];
  const double r[3] = {(double)p.x, (double)p.y, (double)p.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double s = (r[0] * g.inv[a] + r[1] * g.inv[3 + a] + r[2] * g.inv[6 + a]) * (double)g.K[a];
    double t;
    anchor_axis(s, g.K[a], P, base[a], t);
    spline_weights<TC, P>((TC)t, Mv[a], dM[a]);
  }
  const TC qi = p.w;
  const int K1 = g.K[1], K2 = g.K[2];
  int gx = base[0];
#pragma unroll 1
  for (int a = 0; a < P; ++a) {
    int gy = base[1];
    const TC wa = qi * Mv[0][a];
#pragma unroll
    for (int b = 0; b < P; ++b) {
      const TC wab = wa * Mv[1][b];
      unsigned long long* row = grid + ((long long)gx * K1 + gy) * K2;
      int gz = base[2];
#pragma unroll
      for (int c = 0; c < P; ++c) {
        atomicAdd(&row[gz], (unsigned long long)to_fx((double)(wab * Mv[2][c]), Q_SCALE));
        gz = gz + 1 == K2 ? 0 : gz + 1;
      }
      gy = gy + 1 == K1 ? 0 : gy + 1;
    }
    gx = gx + 1 == g.K[0] ? 0 : gx + 1;
  }
}

template <typename TC, int P>
__global__ __launch_bounds__(FB) void k_gather_slots(const int* __restrict__ ncl_ptr, const TC* __restrict__ phi_grid,
                                                     const V4<TC>* __restrict__ xyzq, const int* __restrict__ slot_orig,
                                                     PmeGeom g, double coul, unsigned long long* __restrict__ facc) {
  int sl = blockIdx.x * FB + threadIdx.x;
  if (sl >= *ncl_ptr * CL || slot_orig[sl] < 0) return;
  V4<TC> p = xyzq[sl];
  TC Mv[3][P], dM[3][P];
  int base[3];
  const double r[3] = {(double)p.x, (double)p.y, (double)p.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    double s = (r[0] * g.inv[a] + r[1] * g.inv[3 + a] + r[2] * g.inv[6 + a]) * (double)g.K[a];
    double t;
    anchor_axis(s, g.K[a], P, base[a], t);
    spline_weights<TC, P>((TC)t, Mv[a], dM[a]);
  }
  const int K1 = g.K[1], K2 = g.K[2];
  TC du0 = 0, du1 = 0, du2 = 0;
  int gx = base[0];
#pragma unroll 1
  for (int a = 0; a < P; ++a) {
    int gy = base[1];
    TC t0 = 0, t1 = 0, t2 = 0;
#pragma unroll
    for (int b = 0; b < P; ++b) {
      const TC* row = phi_grid + ((long long)gx * K1 + gy) * K2;
      int gz = base[2];
      TC s0 = 0, s1 = 0;
#pragma unroll
      for (int c = 0; c < P; ++c) {
        TC v = row[gz];
        s0 += v * Mv[2][c];
        s1 += v * dM[2][c];
        gz = gz + 1 == K2 ? 0 : gz + 1;
      }
      t0 += Mv[1][b] * s0;
      t1 += dM[1][b] * s0;
      t2 += Mv[1][b] * s1;
      gy = gy + 1 == K1 ? 0 : gy + 1;
    }
    du0 += dM[0][a] * t0;
    du1 += Mv[0][a] * t1;
    du2 += Mv[0][a] * t2;
    gx = gx + 1 == g.K[0] ? 0 : gx + 1;
  }
  // dE/dr = k q grad phi; weights depend on -s, ds_a/dr_j = K_a inv[j][a]
  const double kq = coul * (double)p.w;
  const double d0 = -(double)du0 * g.K[0], d1 = -(double)du1 * g.K[1], d2 = -(double)du2 * g.K[2];
  const double gxr = d0 * g.inv[0] + d1 * g.inv[1] + d2 * g.inv[2];
  const double gyr = d0 * g.inv[3] + d1 * g.inv[4] + d2 * g.inv[5];
  const double gzr = d0 * g.inv[6] + d1 * g.inv[7] + d2 * g.inv[8];
  // force = -dE/dr, added to this slot's accumulator (one thread per slot: no atomics needed)
  facc[3 * sl + 0] += (unsigned long long)to_fx(-kq * gxr, F_SCALE);
  facc[3 * sl + 1] += (unsigned long long)to_fx(-kq * gyr, F_SCALE);
  facc[3 * sl + 2] += (unsigned long long)to_fx(-kq * gzr, F_SCALE);
}

// ---------------------------------------------------------------- 6. finalize
template <typename TI>
__global__ __launch_bounds__(FB) void k_forces_out(const int* __restrict__ ncl_ptr, const int* __restrict__ slot_orig,
                                                   const unsigned long long* __restrict__ facc, TI* __restrict__ forces) {
  int sl = blockIdx.x * FB + threadIdx.x;
  if (sl >= *ncl_ptr * CL) return;
  int i = slot_orig[sl];
  if (i < 0) return;
  for (int a = 0; a < 3; ++a) forces[3 * i + a] = (TI)((double)(long long)facc[3 * sl + a] / F_SCALE);
}

template <typename TI>
__global__ void k_self_q2(const TI* __restrict__ q, long long n, double* __restrict__ part) {
  double acc = 0.0;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double v = (double)q[i];
    acc += v * v;
  }
  double t = tff_block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_energies_out(const unsigned long long* __restrict__ eacc,
                                                      const double* __restrict__ kparts, int nk,
                                                      const double* __restrict__ qparts, int nq, double coul,
                                                      double alpha, double* __restrict__ out) {
  double ek = tff_sum_partials(kparts, nk);
  double q2 = tff_sum_partials(qparts, nq);
  if (threadIdx.x != 0) return;
  out[0] = (double)(long long)eacc[0] / F_SCALE;
  out[1] = (double)(long long)eacc[1] / F_SCALE;
  out[2] = coul * ek;
  out[3] = -coul * alpha * 0.5641895835477563 * q2;
}

__global__ void k_zero_u64(unsigned long long* p, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) p[i] = 0ull;
}

__global__ void k_zero_i32(int* p, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) p[i] = 0;
}

// ---------------------------------------------------------------- host side
constexpr int QBLOCKS = 64;

bool make_geo(const TffBox& box, long long n, double list_cutoff, Geo& g) {
  const double* b = box.b;
  if (!box.periodic || b[1] != 0 || b[2] != 0 || b[3] != 0 || b[5] != 0 || b[6] != 0 || b[7] != 0) return false;
  for (int a = 0; a < 3; ++a) {
    g.L[a] = b[4 * a];
    if (!(g.L[a] > 0)) return false;
    g.invL[a] = 1.0 / g.L[a];
  }
  const double vol = g.L[0] * g.L[1] * g.L[2];
  const double rho = (double)(n > 0 ? n : 1) / vol;
  const double a = cbrt((double)CL / rho);  // edge of a cube holding one cluster
  for (int k = 0; k < 2; ++k) {
    int nc = (int)floor(g.L[k] / a);
    if (nc < 1) nc = 1;
    if (nc > 1024) nc = 1024;
    g.ncol[k] = nc;
    g.colw[k] = g.L[k] / nc;
    // clusters |d| columns apart are at least (|d| - 1) column widths apart
    g.reach[k] = (int)floor(list_cutoff / g.colw[k]) + 1;
  }
  int nzb = (int)floor(g.L[2] / (a / 8.0));
  if (nzb < 1) nzb = 1;
  if (nzb > 4096) nzb = 4096;
  g.nzb = nzb;
  g.zbw = g.L[2] / nzb;
  g.nkeys = g.ncol[0] * g.ncol[1] * g.nzb;
  g.ncl_max = (int)((n + CL - 1) / CL) + g.ncol[0] * g.ncol[1];
  // cluster pairs per i-cluster (half list): clusters whose cube of edge a lies within the
  // dilated cube of edge a + 2 rlist, with a 2.5x margin for cluster shapes, in multiples of 64
  double dil = (a + 2.0 * list_cutoff) / a;
  int maxn = (int)ceil(dil * dil * dil * 1.25 / 64.0) * 64;
  if (maxn < 128) maxn = 128;
  if (maxn > g.ncl_max) maxn = ((g.ncl_max + 63) / 64) * 64;
  g.maxn = maxn;
  g.pair_cap = (long long)g.ncl_max * maxn;
  return true;
}

struct FusedWs {
  int *key_of, *key_count, *key_start, *cursor, *sorted_orig, *pos_of, *col_cl_start, *ncl, *cl_col, *slot_orig,
      *excl_part, *ncp, *seg, *seg_count, *seg_start;
  void *xyzq, *lj;
  float4 *bb_c, *bb_h;
  int2* cpairs;
  unsigned long long *facc, *eacc, *grid;
  void* work;
  double *kparts, *qparts;
  size_t bytes;
  int nk;
};

FusedWs carve(void* base, long long n, const Geo& g, const int K[3], size_t tc) {
  FusedWs w{};
  size_t off = 0;
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* r = p ? p + off : nullptr;
    off += tff_align(b > 0 ? b : 1);
    return r;
  };
  size_t nn = (size_t)(n > 0 ? n : 1);
  size_t slots = (size_t)g.ncl_max * CL;
  size_t ng = (size_t)K[0] * K[1] * K[2];
  int ncols = g.ncol[0] * g.ncol[1];
  w.key_of = (int*)take(4 * nn);
  w.key_count = (int*)take(4 * (size_t)(g.nkeys + 1));
  w.key_start = (int*)take(4 * (size_t)(g.nkeys + 1));
  w.cursor = (int*)take(4 * (size_t)(g.nkeys + 1));
  w.sorted_orig = (int*)take(4 * nn);
  w.pos_of = (int*)take(4 * nn);
  w.col_cl_start = (int*)take(4 * (size_t)(ncols + 1));
  w.ncl = (int*)take(16);
  w.ncp = (int*)take(16);
  w.cl_col = (int*)take(4 * (size_t)g.ncl_max);
  w.slot_orig = (int*)take(4 * slots);
  w.excl_part = (int*)take(4 * slots * MAXEX);
  w.xyzq = take(4 * tc * slots);
  w.lj = take(2 * tc * slots);
  w.bb_c = (float4*)take(16 * (size_t)g.ncl_max);
  w.bb_h = (float4*)take(16 * (size_t)g.ncl_max);
  w.cpairs = (int2*)take(8 * (size_t)g.pair_cap);
  w.seg = (int*)take(4 * (size_t)g.pair_cap);
  w.seg_count = (int*)take(4 * (size_t)g.ncl_max);
  w.seg_start = (int*)take(4 * (size_t)g.ncl_max);
  w.facc = (unsigned long long*)take(8 * 3 * slots);
  w.eacc = (unsigned long long*)take(8 * 4);
  w.grid = (unsigned long long*)take(8 * ng);
  w.work = take(2 * tc * ng);
  PassCfg px = pass_cfg(1, K[0], K[1] * K[2], 2 * tc);
  w.nk = px.tiles;
  w.kparts = (double*)take(8 * (size_t)w.nk);
  w.qparts = (double*)take(8 * QBLOCKS);
  w.bytes = off;
  return w;
}

template <typename TC>
int launch_pme_slots(int order, const FusedWs& w, const PmeGeom& pg, int nslots, const TC* phi, double coul,
                     bool spread, hipStream_t st) {
  unsigned nb = (unsigned)((nslots + FB - 1) / FB);
#define TFF_SP(P)                                                                                               \
  if (spread)                                                                                                   \
    hipLaunchKernelGGL((k_spread_slots<TC, P>), dim3(nb), dim3(FB), 0, st, w.ncl, (const V4<TC>*)w.xyzq,        \
                       w.slot_orig, pg, w.grid);                                                                 \
  else                                                                                                          \
    hipLaunchKernelGGL((k_gather_slots<TC, P>), dim3(nb), dim3(FB), 0, st, w.ncl, phi, (const V4<TC>*)w.xyzq,   \
                       w.slot_orig, pg, coul, w.facc)
  switch (order) {
    case 3: TFF_SP(3); break;
    case 4: TFF_SP(4); break;
    case 5: TFF_SP(5); break;
    case 6: TFF_SP(6); break;
    case 7: TFF_SP(7); break;
    case 8: TFF_SP(8); break;
    default: tff_set_error("fused: unsupported spline order %d (3..8)", order); return TFF_ERR_UNSUPPORTED;
  }
#undef TFF_SP
  return TFF_OK;
}

template <typename TI, typename TC>
int run_fused(const TI* coords, const TI* q, const TI* sig, const TI* eps, long long n, const TffBox& box,
              const Geo& g, const int64_t* row_ptr, const int64_t* col, double cutoff, double list_cutoff,
              double alpha, double coul, const int K[3], int order, const TC* mx, const TC* my, const TC* mz,
              double* energies, TI* forces, int* status, const FusedWs& w, hipStream_t st) {
  const unsigned nbA = (unsigned)((n + FB - 1) / FB);
  const int nslots = g.ncl_max * CL;
  const unsigned nbS = (unsigned)((nslots + FB - 1) / FB);
  const int ncols = g.ncol[0] * g.ncol[1];
  auto zero32 = [&](int* p, long long m) {
    long long b = (m + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_zero_i32, dim3((unsigned)b), dim3(256), 0, st, p, m);
  };
  auto zero64 = [&](unsigned long long* p, long long m) {
    long long b = (m + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)b), dim3(256), 0, st, p, m);
  };
  zero32(w.key_count, g.nkeys + 1);
  zero32(w.cursor, g.nkeys + 1);
  zero32(w.ncp, 4);
  zero32(status, 1);
  zero64(w.facc, 3LL * nslots);
  zero64(w.eacc, 4);
  zero64(w.grid, (long long)K[0] * K[1] * K[2]);
  // 1. sort
  hipLaunchKernelGGL(k_bin_atoms<TI>, dim3(nbA), dim3(FB), 0, st, coords, n, g, w.key_of, w.key_count);
  hipLaunchKernelGGL(k_scan_keys, dim3(1), dim3(1024), 0, st, w.key_count, g.nkeys, w.key_start);
  hipLaunchKernelGGL(k_scan_columns, dim3(1), dim3(1024), 0, st, w.key_start, ncols, g.nzb, w.col_cl_start, w.ncl);
  hipLaunchKernelGGL(k_scatter_atoms, dim3(nbA), dim3(FB), 0, st, n, w.key_of, w.key_start, w.cursor,
                     w.sorted_orig, w.pos_of);
  hipLaunchKernelGGL(k_sort_keys, dim3((g.nkeys + FB - 1) / FB), dim3(FB), 0, st, g.nkeys, w.key_start,
                     w.sorted_orig, w.pos_of);
  // 2. cluster arrays
  zero32(w.slot_orig, nslots);
  hipLaunchKernelGGL((k_fill_clusters<TI, TC>), dim3(nbS), dim3(FB), 0, st, coords, q, sig, eps, g, w.ncl,
                     w.col_cl_start, w.key_start, w.sorted_orig, (V4<TC>*)w.xyzq, (V2<TC>*)w.lj, w.slot_orig,
                     w.cl_col);
  hipLaunchKernelGGL(k_clear_pad_exclusions, dim3(nbS), dim3(FB), 0, st, w.ncl, w.slot_orig, w.excl_part);
  hipLaunchKernelGGL(k_fill_exclusions, dim3(nbA), dim3(FB), 0, st, n, w.pos_of, w.key_of, g, w.key_start,
                     w.col_cl_start, row_ptr, col, w.excl_part, status);
  const unsigned nbW = (unsigned)(((long long)g.ncl_max * TFF_WAVE + FB - 1) / FB);
  hipLaunchKernelGGL(k_bbox<TC>, dim3(nbW), dim3(FB), 0, st, w.ncl, (const V4<TC>*)w.xyzq, w.slot_orig, w.bb_c,
                     w.bb_h);
  // 3. cluster pairs
  {
    void* tok = tff_prof_begin("fused_cluster_pairs", st);
    float rl2 = (float)(list_cutoff * list_cutoff);
    hipLaunchKernelGGL(k_cluster_pairs, dim3(nbW), dim3(FB), 0, st, w.ncl, w.cl_col, w.col_cl_start, w.bb_c,
                       w.bb_h, g, rl2, w.seg, w.seg_count, status);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, w.ncl, w.seg_count, w.seg_start, w.ncp);
    hipLaunchKernelGGL(k_compact_pairs, dim3(nbW), dim3(FB), 0, st, w.ncl, w.seg, w.seg_count, w.seg_start, g.maxn,
                       w.cpairs);
    tff_prof_end(tok, st);
  }
  // 4. tiles
  {
    TileArgs<TC> ta;
    for (int a = 0; a < 3; ++a) {
      ta.L[a] = (TC)g.L[a];
      ta.invL[a] = (TC)g.invL[a];
    }
    ta.rc2 = (TC)(cutoff * cutoff);
    ta.alpha = (TC)alpha;
    ta.coul = (TC)coul;
    long long want = (g.pair_cap + TILE_WAVES - 1) / TILE_WAVES;
    long long blocks = want < 4096 ? want : 4096;
    if (blocks < 1) blocks = 1;
    void* tok = tff_prof_begin("fused_tiles", st);
    hipLaunchKernelGGL(k_tiles<TC>, dim3((unsigned)blocks), dim3(TILE_BLOCK), 0, st, w.cpairs, w.ncp, g.pair_cap,
                       (const V4<TC>*)w.xyzq, (const V2<TC>*)w.lj, w.slot_orig, w.excl_part, ta, w.facc, w.eacc);
    tff_prof_end(tok, st);
  }
  // 5. PME
  PmeGeom pg = make_geom(box, K, order, alpha);
  FftPlan pl[3];
  for (int a = 0; a < 3; ++a) make_plan(K[a], pl[a]);
  {
    void* tok = tff_prof_begin("fused_spread", st);
    int rc = launch_pme_slots<TC>(order, w, pg, nslots, nullptr, coul, true, st);
    tff_prof_end(tok, st);
    if (rc) return rc;
  }
  ConvArgs cv;
  for (int a = 0; a < 3; ++a) cv.K[a] = K[a];
  for (int k = 0; k < 9; ++k) cv.inv[k] = box.inv[k];
  cv.vol = box.det;
  cv.alpha = alpha;
  cv.grid_fixed = 1;
  cv.grid_inv_scale = 1.0 / Q_SCALE;
  const size_t ce = 2 * sizeof(TC);
  PassCfg cz = pass_cfg(K[0] * K[1], K[2], 1, ce);
  PassCfg cy = pass_cfg(K[0], K[1], K[2], ce);
  PassCfg cxp = pass_cfg(1, K[0], K[1] * K[2], ce);
  TC* phi = (TC*)w.grid;  // the real potential grid reuses the charge grid's memory after the z pass
  {
    void* tok = tff_prof_begin("fused_fft", st);
    hipLaunchKernelGGL((k_fft_pass<TC, PASS_Z_FWD>), dim3(cz.tiles), dim3(FFT_THREADS), cz.lds, st, w.grid, w.work,
                       cz.pa, pl[2], cv, mx, my, mz, 1, w.kparts);
    hipLaunchKernelGGL((k_fft_pass<TC, PASS_Y_FWD>), dim3(cy.tiles), dim3(FFT_THREADS), cy.lds, st, w.work, w.work,
                       cy.pa, pl[1], cv, mx, my, mz, 1, w.kparts);
    hipLaunchKernelGGL((k_fft_pass<TC, PASS_X_CONV>), dim3(cxp.tiles), dim3(FFT_THREADS), cxp.lds, st, w.work,
                       w.work, cxp.pa, pl[0], cv, mx, my, mz, 1, w.kparts);
    hipLaunchKernelGGL((k_fft_pass<TC, PASS_Y_INV>), dim3(cy.tiles), dim3(FFT_THREADS), cy.lds, st, w.work, w.work,
                       cy.pa, pl[1], cv, mx, my, mz, 1, w.kparts);
    hipLaunchKernelGGL((k_fft_pass<TC, PASS_Z_INV>), dim3(cz.tiles), dim3(FFT_THREADS), cz.lds, st, w.work, phi,
                       cz.pa, pl[2], cv, mx, my, mz, 1, w.kparts);
    tff_prof_end(tok, st);
  }
  {
    void* tok = tff_prof_begin("fused_gather", st);
    int rc = launch_pme_slots<TC>(order, w, pg, nslots, phi, coul, false, st);
    tff_prof_end(tok, st);
    if (rc) return rc;
  }
  // 6. outputs
  hipLaunchKernelGGL(k_forces_out<TI>, dim3(nbS), dim3(FB), 0, st, w.ncl, w.slot_orig, w.facc, forces);
  hipLaunchKernelGGL(k_self_q2<TI>, dim3(QBLOCKS), dim3(256), 0, st, q, n, w.qparts);
  hipLaunchKernelGGL(k_energies_out, dim3(1), dim3(256), 0, st, w.eacc, w.kparts, w.nk, w.qparts, QBLOCKS, coul,
                     alpha, energies);
  TFF_CHECK_LAUNCH("fused energy+forces");
  return TFF_OK;
}

}  // namespace

extern "C" size_t tff_fused_workspace_bytes(int compute_dtype, long long natoms, const double* box_host,
                                            double list_cutoff, int K0, int K1, int K2) {
  TffBox box = tff_make_box(box_host);
  Geo g;
  if (!make_geo(box, natoms, list_cutoff, g)) return 0;
  int K[3] = {K0, K1, K2};
  return carve(nullptr, natoms, g, K, compute_dtype == TFF_F64 ? 8 : 4).bytes;
}

extern "C" int tff_fused_energy_forces(int io_dtype, int compute_dtype, const void* coords, const void* charges,
                                       const void* sigma, const void* epsilon, long long natoms,
                                       const double* box_host, const int64_t* excl_row_ptr,
                                       const int64_t* excl_col, double cutoff, double list_cutoff,
                                       double ewald_alpha, double coul_constant, int K0, int K1, int K2, int order,
                                       const void* xmoduli, const void* ymoduli, const void* zmoduli,
                                       double* energies_out, void* forces_out, int32_t* status_out,
                                       void* workspace, size_t workspace_bytes, hipStream_t stream) {
  TffBox box = tff_make_box(box_host);
  Geo g;
  if (!make_geo(box, natoms, list_cutoff, g)) {
    tff_set_error("fused: needs an orthorhombic periodic box (diagonal box matrix)");
    return TFF_ERR_UNSUPPORTED;
  }
  if (natoms <= 0 || !coords || !charges || !sigma || !epsilon || !energies_out || !forces_out || !status_out ||
      !xmoduli || !ymoduli || !zmoduli || !(cutoff > 0) || list_cutoff < cutoff || ewald_alpha < 0) {
    tff_set_error("fused: invalid arguments");
    return TFF_ERR_ARG;
  }
  if ((excl_row_ptr == nullptr) != (excl_col == nullptr)) {
    tff_set_error("fused: exclusion row_ptr and col must be given together");
    return TFF_ERR_ARG;
  }
  int K[3] = {K0, K1, K2};
  for (int a = 0; a < 3; ++a) {
    FftPlan p;
    if (K[a] < order || K[a] > 4096 || !make_plan(K[a], p)) {
      tff_set_error("fused: PME grid size %d must be 2^a 3^b 5^c and >= the spline order", K[a]);
      return TFF_ERR_UNSUPPORTED;
    }
  }
  for (int a = 0; a < 3; ++a)
    if (2.0 * list_cutoff > g.L[a]) {
      tff_set_error("fused: the box (%.4g nm) must exceed twice the list cutoff (%.4g nm) for the minimum image",
                    g.L[a], list_cutoff);
      return TFF_ERR_UNSUPPORTED;
    }
  size_t tc = compute_dtype == TFF_F64 ? 8 : 4;
  size_t need = carve(nullptr, natoms, g, K, tc).bytes;
  if (!workspace || workspace_bytes < need) {
    tff_set_error("fused: workspace too small (%zu < %zu)", workspace_bytes, need);
    return TFF_ERR_WORKSPACE;
  }
  FusedWs w = carve(workspace, natoms, g, K, tc);
#define TFF_RUN(TI, TC)                                                                                          \
  return run_fused<TI, TC>((const TI*)coords, (const TI*)charges, (const TI*)sigma, (const TI*)epsilon, natoms, \
                           box, g, excl_row_ptr, excl_col, cutoff, list_cutoff, ewald_alpha, coul_constant, K,   \
                           order, (const TC*)xmoduli, (const TC*)ymoduli, (const TC*)zmoduli, energies_out,      \
                           (TI*)forces_out, (int*)status_out, w, stream)
  if (io_dtype == TFF_F64 && compute_dtype == TFF_F32) TFF_RUN(double, float);
  if (io_dtype == TFF_F32 && compute_dtype == TFF_F32) TFF_RUN(float, float);
  if (io_dtype == TFF_F64 && compute_dtype == TFF_F64) TFF_RUN(double, double);
  if (io_dtype == TFF_F32 && compute_dtype == TFF_F64) TFF_RUN(float, double);
#undef TFF_RUN
  tff_set_error("fused: unsupported dtype combination");
  return TFF_ERR_ARG;
}
