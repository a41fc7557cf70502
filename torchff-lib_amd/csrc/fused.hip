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

// Reorder the 32 slots of every cluster into four spatially compact groups of 8: sort by x
// and split in halves, sort each half by y and split, order each group by z (padding slots,
// keyed +inf, collect at the end; ties broken by slot).  One wave per cluster, lanes 0..31.
// Also writes slot_of_orig[i], the final slot of atom i.
template <typename TC>
__global__ __launch_bounds__(FB) void k_reorder_clusters(const int* __restrict__ ncl_ptr, V4<TC>* __restrict__ xyzq,
                                                         V2<TC>* __restrict__ lj, int* __restrict__ slot_orig,
                                                         int* __restrict__ slot_of_orig) {
  const int c = (blockIdx.x * FB + threadIdx.x) / TFF_WAVE;
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  if (c >= *ncl_ptr) return;
  const int s = lane & (CL - 1);
  const int sl = c * CL + s;
  const V4<TC> p = xyzq[sl];
  const V2<TC> l = lj[sl];
  const int o = slot_orig[sl];
  const float inf = 3.0e38f;
  const float kx = o >= 0 ? (float)p.x : inf, ky = o >= 0 ? (float)p.y : inf, kz = o >= 0 ? (float)p.z : inf;
  // rank by x over all 32 slots
  int rx = 0;
  for (int t = 0; t < CL; ++t) {
    float vx = __shfl(kx, t, TFF_WAVE);
    rx += (vx < kx || (vx == kx && t < s)) ? 1 : 0;
  }
  const int half = rx >= CL / 2;
  int ry = 0;
  for (int t = 0; t < CL; ++t) {
    float vy = __shfl(ky, t, TFF_WAVE);
    int ht = __shfl(half, t, TFF_WAVE);
    ry += (ht == half && (vy < ky || (vy == ky && t < s))) ? 1 : 0;
  }
  const int grp = 2 * half + (ry >= CL / 4);
  int rz = 0;
  for (int t = 0; t < CL; ++t) {
    float vz = __shfl(kz, t, TFF_WAVE);
    int gt = 2 * __shfl(half, t, TFF_WAVE);
    int rt = __shfl(ry, t, TFF_WAVE);
    gt += rt >= CL / 4;
    rz += (gt == grp && (vz < kz || (vz == kz && t < s))) ? 1 : 0;
  }
  const int ns = c * CL + grp * 8 + rz;
  __builtin_amdgcn_wave_barrier();
  if (lane < CL) {
    xyzq[ns] = p;
    lj[ns] = l;
    slot_orig[ns] = o;
    if (o >= 0) slot_of_orig[o] = ns;
  }
}

// exclusion partners of every slot as packed (cluster << 5 | slot), -1 = none
__global__ __launch_bounds__(FB) void k_fill_exclusions(long long n, const int* __restrict__ slot_of_orig,
                                                        const int64_t* __restrict__ row_ptr,
                                                        const int64_t* __restrict__ col,
                                                        int* __restrict__ excl_part, int* __restrict__ status) {
  long long i = blockIdx.x * (long long)FB + threadIdx.x;
  if (i >= n) return;
  int si = slot_of_orig[i];
  int k = 0;
  if (col) {
    for (long long e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      long long j = col[e];
      if (j == i || j < 0 || j >= n) continue;
      if (k == MAXEX) {
        atomicOr(status, TFF_FUSED_EXCL_OVERFLOW);
        break;
      }
      excl_part[si * MAXEX + k++] = slot_of_orig[j];
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

// cluster bounding boxes (centre, half extent) of the whole cluster and of its four groups of
// 8 slots; an empty group gets half extent -1.  One wave per cluster, lanes 0..31.
template <typename TC>
__global__ __launch_bounds__(FB) void k_bbox(const int* __restrict__ ncl_ptr, const V4<TC>* __restrict__ xyzq,
                                             const int* __restrict__ slot_orig, float4* __restrict__ bb_c,
                                             float4* __restrict__ bb_h, float4* __restrict__ sb_c,
                                             float4* __restrict__ sb_h) {
  int c = (blockIdx.x * FB + threadIdx.x) / TFF_WAVE;
  int lane = threadIdx.x & (TFF_WAVE - 1);
  if (c >= *ncl_ptr) return;
  int sl = c * CL + (lane & (CL - 1));
  bool v = slot_orig[sl] >= 0;
  V4<TC> p = xyzq[sl];
  float mn[3], mx[3], gmn[3], gmx[3];
  float pv[3] = {(float)p.x, (float)p.y, (float)p.z};
  for (int a = 0; a < 3; ++a) {
    mn[a] = v ? pv[a] : 3.0e38f;
    mx[a] = v ? pv[a] : -3.0e38f;
    for (int off = 4; off > 0; off >>= 1) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, TFF_WAVE));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, TFF_WAVE));
    }
    gmn[a] = mn[a];
    gmx[a] = mx[a];
    for (int off = 16; off > 4; off >>= 1) {
      mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, TFF_WAVE));
      mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, TFF_WAVE));
    }
  }
  // widen by a relative epsilon so float rounding of the tests never drops a pair
  if (lane == 0) {
    bb_c[c] = make_float4(0.5f * (mn[0] + mx[0]), 0.5f * (mn[1] + mx[1]), 0.5f * (mn[2] + mx[2]), 0.f);
    bb_h[c] = make_float4(0.5f * (mx[0] - mn[0]) * 1.0001f + 1e-6f, 0.5f * (mx[1] - mn[1]) * 1.0001f + 1e-6f,
                          0.5f * (mx[2] - mn[2]) * 1.0001f + 1e-6f, 0.f);
  }
  if (lane < CL && (lane & 7) == 0) {
    const int gi = c * 4 + lane / 8;
    if (gmx[0] < gmn[0]) {
      sb_c[gi] = make_float4(0.f, 0.f, 0.f, 0.f);
      sb_h[gi] = make_float4(-1.f, -1.f, -1.f, 0.f);
    } else {
      sb_c[gi] = make_float4(0.5f * (gmn[0] + gmx[0]), 0.5f * (gmn[1] + gmx[1]), 0.5f * (gmn[2] + gmx[2]), 0.f);
      sb_h[gi] = make_float4(0.5f * (gmx[0] - gmn[0]) * 1.0001f + 1e-6f, 0.5f * (gmx[1] - gmn[1]) * 1.0001f + 1e-6f,
                             0.5f * (gmx[2] - gmn[2]) * 1.0001f + 1e-6f, 0.f);
    }
  }
}

// ---------------------------------------------------------------- 3. cluster pairs
// One wave per i-cluster; lanes sweep the clusters of the neighbouring columns in a fixed
// order and write the accepted j-clusters (cj >= ci) into the i-cluster's own segment.
__global__ __launch_bounds__(FB) void k_cluster_pairs(const int* __restrict__ ncl_ptr, const int* __restrict__ cl_col,
                                                      const int* __restrict__ col_cl_start,
                                                      const float4* __restrict__ bb_c,
                                                      const float4* __restrict__ bb_h, Geo g, float rlist2,
                                                      const int* __restrict__ excl_part,
                                                      int* __restrict__ seg, int* __restrict__ seg_count,
                                                      int* __restrict__ status) {
  __shared__ int s_xp[FB / TFF_WAVE][CL * MAXEX];
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int w = threadIdx.x / TFF_WAVE;
  const int ci = blockIdx.x * (FB / TFF_WAVE) + w;
  const int ncl = *ncl_ptr;
  if (ci >= ncl) return;
  // clusters (other than ci) holding an exclusion partner of an atom of ci, compacted in LDS
  int nxp = 0;
  for (int k = 0; k < MAXEX; ++k) {
    int v = lane < CL ? excl_part[(ci * CL + lane) * MAXEX + k] : -1;
    int pc = v >= 0 ? (v >> 5) : -1;
    bool keep = pc >= 0 && pc != ci;
    unsigned long long m = __ballot(keep);
    if (keep) s_xp[w][nxp + __popcll(m & ((1ull << lane) - 1ull))] = pc;
    nxp += __popcll(m);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
          // bit 30 marks tiles that contain excluded pairs
          int xf = cj == ci ? 1 : 0;
          for (int t = 0; t < nxp; ++t) xf |= s_xp[w][t] == cj ? 1 : 0;
          int k = nseg + __popcll(m & ((1ull << lane) - 1ull));
          if (k < g.maxn) my[k] = cj | (xf << 30);
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
                                                      int* __restrict__ start, int* __restrict__ ncp,
                                                      int* __restrict__ status) {
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
  if (t == 1023) {
    *ncp = s[1023];
    status[1] = s[1023];
  }
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
// erfc(u) and exp(-u^2) for u >= 0.  float: library exp, hardware reciprocal, and
// erfc(u) exp(u^2) = sum_k c_k t^k, t = 1 / (1 + 0.39 u), a degree-8 polynomial fitted to the
// scaled complementary error function on [0, 3.6] by iteratively reweighted least squares
// (relative error 1.3e-9 in exact arithmetic, about 1.5e-6 evaluated in float32, against
// scipy.special.erfc); double: the correctly rounded library functions.
__device__ __forceinline__ void erfc_exp(float u, float& ec, float& em) {
  em = expf(-u * u);
  const float t = __frcp_rn(1.0f + 0.39f * u);
  float p = 9.27117256e-02f;
  p = p * t - 5.29840954e-01f;
  p = p * t + 1.06552963e+00f;
  p = p * t - 8.61721388e-01f;
  p = p * t + 7.81784359e-01f;
  p = p * t - 4.21555312e-02f;
  p = p * t + 2.82117988e-01f;
  p = p * t + 2.10995424e-01f;
  p = p * t + 5.78745155e-04f;
  ec = p * em;
}
__device__ __forceinline__ void erfc_exp(double u, double& ec, double& em) {
  em = exp(-u * u);
  ec = erfc(u);
}
// hardware reciprocal square root refined by one Newton step (LJ raises its error to the 6th power)
__device__ __forceinline__ float rsqrt_t(float x) {
  const float y = rsqrtf(x);
  return y * (1.5f - 0.5f * x * y * y);
}
__device__ __forceinline__ double rsqrt_t(double x) { return 1.0 / sqrt(x); }

template <typename TC>
struct TileArgs {
  TC L[3], invL[3];
  TC rc2;
  TC alpha;
  TC coul;
};

// One wave per 32 x 32 cluster tile at a time, evaluated as up to 16 sub-tiles of 8 x 8 atoms
// (the clusters' compact groups of 8): lane = (ii = lane & 7, jj = lane >> 3) holds one pair of
// a sub-tile, and only sub-tiles whose group bounding boxes are within the list cutoff are
// evaluated (a wave-uniform loop).  F_i stays in registers across consecutive tiles of the same
// i-cluster; F_j is reduced over the 8 lanes sharing jj and added with int64 atomics per tile.
template <typename TC>
__global__ __launch_bounds__(TILE_BLOCK) void k_tiles(const int2* __restrict__ cpairs, const int* __restrict__ ncp_ptr,
                                                      long long cap, const V4<TC>* __restrict__ xyzq,
                                                      const V2<TC>* __restrict__ lj, const int* __restrict__ slot_orig,
                                                      const int* __restrict__ excl_part, const float4* __restrict__ sb_c,
                                                      const float4* __restrict__ sb_h, TileArgs<TC> ta, float rlist2,
                                                      unsigned long long* __restrict__ facc,
                                                      unsigned long long* __restrict__ eacc) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int w = threadIdx.x / TFF_WAVE;
  const int ii = lane & 7, jj = lane >> 3;
  long long ncp = *ncp_ptr;
  if (ncp > cap) ncp = cap;
  const long long nw = (long long)gridDim.x * TILE_WAVES;
  const long long gw = (long long)blockIdx.x * TILE_WAVES + w;
  const long long per = (ncp + nw - 1) / nw;
  const long long t0 = gw * per;
  long long t1 = t0 + per;
  if (t1 > ncp) t1 = ncp;
  const TC two_over_sqrt_pi = (TC)1.1283791670955126;
  const float fL[3] = {(float)ta.L[0], (float)ta.L[1], (float)ta.L[2]};
  const float fiL[3] = {(float)ta.invL[0], (float)ta.invL[1], (float)ta.invL[2]};
  TC e_lj = 0, e_c = 0;
  int cur_i = -1;
  TC fi[4][3];
  V4<TC> pi[4];
  V2<TC> li[4];
  bool vi[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    fi[g][0] = fi[g][1] = fi[g][2] = 0;
    vi[g] = false;
  }
  auto flush_i = [&](int c) {
    // reduce over the 8 lanes sharing ii (lane bits 3..5), lanes 0..7 write
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        TC v = fi[g][a];
        v += __shfl_xor(v, 8, TFF_WAVE);
        v += __shfl_xor(v, 16, TFF_WAVE);
        v += __shfl_xor(v, 32, TFF_WAVE);
        fi[g][a] = v;
      }
    if (jj == 0) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int sl = c * CL + g * 8 + ii;
        if (vi[g]) {
#pragma unroll
          for (int a = 0; a < 3; ++a)
            if (fi[g][a] != (TC)0) atomicAdd(&facc[3 * sl + a], (unsigned long long)to_fx((double)fi[g][a], F_SCALE));
        }
      }
    }
  };
  for (long long t = t0; t < t1; ++t) {
    const int2 cp = cpairs[t];
    const int ci = cp.x, cj = cp.y & 0x3fffffff;
    const bool xflag = (cp.y >> 30) & 1;
    if (ci != cur_i) {
      if (cur_i >= 0) flush_i(cur_i);
      cur_i = ci;
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int sl = ci * CL + g * 8 + ii;
        pi[g] = xyzq[sl];
        li[g] = lj[sl];
        vi[g] = slot_orig[sl] >= 0;
        fi[g][0] = fi[g][1] = fi[g][2] = 0;
      }
    }
    const bool diag = ci == cj;
    // active sub-tiles: group bounding boxes within the list cutoff (lanes 0..15 test one each)
    bool sub_ok = false;
    if (lane < 16) {
      const int gi = lane >> 2, gj = lane & 3;
      const float4 c0 = sb_c[ci * 4 + gi], h0 = sb_h[ci * 4 + gi];
      const float4 c1 = sb_c[cj * 4 + gj], h1 = sb_h[cj * 4 + gj];
      if (h0.x >= 0.f && h1.x >= 0.f && (!diag || gj >= gi)) {
        float dx = c0.x - c1.x, dy = c0.y - c1.y, dz = c0.z - c1.z;
        dx -= fL[0] * rintf(dx * fiL[0]);
        dy -= fL[1] * rintf(dy * fiL[1]);
        dz -= fL[2] * rintf(dz * fiL[2]);
        const float gx = fmaxf(fabsf(dx) - h0.x - h1.x, 0.f);
        const float gy = fmaxf(fabsf(dy) - h0.y - h1.y, 0.f);
        const float gz = fmaxf(fabsf(dz) - h0.z - h1.z, 0.f);
        sub_ok = gx * gx + gy * gy + gz * gz <= rlist2;
      }
    }
    const unsigned int mask = (unsigned int)(__ballot(sub_ok) & 0xffffull);
    if (mask == 0) continue;
    // exclusion bit masks of this lane's four i-slots against cj (only for flagged tiles)
    unsigned int exm[4] = {0u, 0u, 0u, 0u};
    if (xflag) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int sl = ci * CL + g * 8 + ii;
#pragma unroll
        for (int k = 0; k < MAXEX; ++k) {
          const int pv = excl_part[sl * MAXEX + k];
          if (pv >= 0 && (pv >> 5) == cj) exm[g] |= 1u << (pv & 31);
        }
      }
    }
    V4<TC> pj[4];
    V2<TC> lJ[4];
    bool vj[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int sl = cj * CL + g * 8 + jj;
      pj[g] = xyzq[sl];
      lJ[g] = lj[sl];
      vj[g] = slot_orig[sl] >= 0;
    }
    TC fj[4][3];
#pragma unroll
    for (int g = 0; g < 4; ++g) fj[g][0] = fj[g][1] = fj[g][2] = 0;
#pragma unroll
    for (int b = 0; b < 16; ++b) {
      if (!((mask >> b) & 1u)) continue;
      const int gi = b >> 2, gj = b & 3;
      TC dx = pi[gi].x - pj[gj].x, dy = pi[gi].y - pj[gj].y, dz = pi[gi].z - pj[gj].z;
      dx -= ta.L[0] * rint(dx * ta.invL[0]);
      dy -= ta.L[1] * rint(dy * ta.invL[1]);
      dz -= ta.L[2] * rint(dz * ta.invL[2]);
      const TC r2 = dx * dx + dy * dy + dz * dz;
      const bool ok = vi[gi] && vj[gj] && r2 <= ta.rc2 && (!(diag && gi == gj) || jj > ii);
      TC fs = 0;
      if (ok) {
        const bool excl = (exm[gi] >> (gj * 8 + jj)) & 1u;
        const TC rinv = rsqrt_t(r2);
        const TC rinv2 = rinv * rinv;
        const TC r = r2 * rinv;
        const TC u = ta.alpha * r;
        TC ec, em;
        erfc_exp(u, ec, em);
        ec -= excl ? (TC)1 : (TC)0;
        const TC ex2 = em * two_over_sqrt_pi * u;
        const TC qq = ta.coul * pi[gi].w * pj[gj].w;
        e_c += qq * ec * rinv;
        fs = qq * (ec + ex2) * rinv * rinv2;
        if (!excl) {
          const TC sg = li[gi].x + lJ[gj].x;
          const TC ep = li[gi].y * lJ[gj].y;
          const TC s2 = sg * sg * rinv2;
          const TC s6 = s2 * s2 * s2;
          e_lj += (TC)4 * ep * s6 * (s6 - (TC)1);
          fs += (TC)24 * ep * s6 * ((TC)2 * s6 - (TC)1) * rinv2;
        }
      }
      // F_i += fs d, F_j -= fs d   (fs = -dE/dr / r)
      fi[gi][0] += fs * dx;
      fi[gi][1] += fs * dy;
      fi[gi][2] += fs * dz;
      fj[gj][0] -= fs * dx;
      fj[gj][1] -= fs * dy;
      fj[gj][2] -= fs * dz;
    }
    // reduce F_j over the 8 lanes sharing jj (lane bits 0..2); lanes with ii == 0 write
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        TC v = fj[g][a];
        v += __shfl_xor(v, 1, TFF_WAVE);
        v += __shfl_xor(v, 2, TFF_WAVE);
        v += __shfl_xor(v, 4, TFF_WAVE);
        fj[g][a] = v;
      }
    if (ii == 0) {
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int sl = cj * CL + g * 8 + jj;
        if (vj[g]) {
#pragma unroll
          for (int a = 0; a < 3; ++a)
            if (fj[g][a] != (TC)0) atomicAdd(&facc[3 * sl + a], (unsigned long long)to_fx((double)fj[g][a], F_SCALE));
        }
      }
    }
  }
  if (cur_i >= 0) flush_i(cur_i);
  // energies: wave sums in double, then fixed point
  double el = tff_wave_sum((double)e_lj), ecl = tff_wave_sum((double)e_c);
  if (lane == 0 && t1 > t0) {
    atomicAdd(&eacc[0], (unsigned long long)to_fx(el, F_SCALE));
    atomicAdd(&eacc[1], (unsigned long long)to_fx(ecl, F_SCALE));
  }
}

// ---------------------------------------------------------------- 5. PME on cluster slots
// Owner-computes spreading: one workgroup per BR^3 block ("brick") of grid points.  The
// workgroup lists the clusters whose stencils can reach its brick (column neighbourhood,
// circular interval tests in grid units), accumulates their charges into the brick in LDS
// with int64 fixed-point atomics (order independent), then stores the brick: every grid
// point is written once, so the grid needs neither zeroing nor global atomics.
constexpr int BR = 8;
constexpr int BRICK_MAXC = 1024;

__device__ __forceinline__ bool circ_overlap(int lo, int len, int b0, int b, int K) {
  if (len >= K) return true;
  int d1 = (b0 - lo) % K;
  if (d1 < 0) d1 += K;
  int d2 = (lo - b0) % K;
  if (d2 < 0) d2 += K;
  return d1 < len || d2 < b;
}

// grid-index interval [lo, lo + len) touched by the stencils of atoms with scaled coordinate
// in [slo, shi] (spline order P, anchor of the spline weights in tffpme::anchor_axis)
__device__ __forceinline__ void stencil_interval(float slo, float shi, int P, int& lo, int& len) {
  lo = (int)floorf(slo) - P / 2 - 1;
  int hi = (int)ceilf(shi) + P - P / 2;
  len = hi - lo + 1;
}

template <typename TC, int P>
__global__ __launch_bounds__(256) void k_spread_bricks(const int* __restrict__ ncl_ptr, const V4<TC>* __restrict__ xyzq,
                                                       const int* __restrict__ slot_orig, const float4* __restrict__ bb_c,
                                                       const float4* __restrict__ bb_h,
                                                       const int* __restrict__ col_cl_start, Geo g, PmeGeom pg,
                                                       unsigned long long* __restrict__ grid, int* __restrict__ status) {
  __shared__ unsigned long long s_q[BR * BR * BR];
  __shared__ int s_list[BRICK_MAXC];
  __shared__ int s_n;
  const int K0 = pg.K[0], K1 = pg.K[1], K2 = pg.K[2];
  const int nbx = (K0 + BR - 1) / BR, nby = (K1 + BR - 1) / BR, nbz = (K2 + BR - 1) / BR;
  const int bid = blockIdx.x;
  if (bid >= nbx * nby * nbz) return;
  const int bz = bid % nbz, by = (bid / nbz) % nby, bx = bid / (nbz * nby);
  const int B0[3] = {bx * BR, by * BR, bz * BR};
  const int Bn[3] = {K0 - B0[0] < BR ? K0 - B0[0] : BR, K1 - B0[1] < BR ? K1 - B0[1] : BR,
                     K2 - B0[2] < BR ? K2 - B0[2] : BR};
  for (int k = threadIdx.x; k < BR * BR * BR; k += blockDim.x) s_q[k] = 0ull;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  // candidate clusters: columns whose x-y extent can reach the brick
  const float kl[3] = {(float)(pg.K[0] * g.invL[0]), (float)(pg.K[1] * g.invL[1]), (float)(pg.K[2] * g.invL[2])};
  int c0[2], nc[2];
  for (int a = 0; a < 2; ++a) {
    double lo = (B0[a] - P / 2 - 2) * g.L[a] / pg.K[a];
    double hi = (B0[a] + Bn[a] + P / 2 + 2) * g.L[a] / pg.K[a];
    int cl = (int)floor(lo / g.colw[a]), ch = (int)floor(hi / g.colw[a]);
    c0[a] = cl;
    nc[a] = ch - cl + 1 < g.ncol[a] ? ch - cl + 1 : g.ncol[a];
  }
  const int ncand_cols = nc[0] * nc[1];
  for (int cc = 0; cc < ncand_cols; ++cc) {
    int xx = (c0[0] + cc / nc[1]) % g.ncol[0];
    if (xx < 0) xx += g.ncol[0];
    int yy = (c0[1] + cc % nc[1]) % g.ncol[1];
    if (yy < 0) yy += g.ncol[1];
    const int col = xx * g.ncol[1] + yy;
    const int cs = col_cl_start[col], ce = col_cl_start[col + 1];
    for (int c = cs + (int)threadIdx.x; c < ce; c += blockDim.x) {
      const float4 cc4 = bb_c[c], hh4 = bb_h[c];
      const float cv[3] = {cc4.x, cc4.y, cc4.z}, hv[3] = {hh4.x, hh4.y, hh4.z};
      bool ok = true;
      for (int a = 0; a < 3; ++a) {
        int lo, len;
        stencil_interval((cv[a] - hv[a]) * kl[a], (cv[a] + hv[a]) * kl[a], P, lo, len);
        ok = ok && circ_overlap(lo, len, B0[a], Bn[a], pg.K[a]);
      }
      if (ok) {
        int k = atomicAdd(&s_n, 1);
        if (k < BRICK_MAXC) s_list[k] = c;
      }
    }
  }
  __syncthreads();
  int ncand = s_n;
  if (ncand > BRICK_MAXC) {
    if (threadIdx.x == 0) atomicOr(status, TFF_FUSED_PAIR_OVERFLOW);
    ncand = BRICK_MAXC;
  }
  // two clusters per wave per round: lane >> 5 picks the cluster, lane & 31 the slot
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int gw = threadIdx.x / TFF_WAVE, nwv = blockDim.x / TFF_WAVE;
  for (int r = 2 * gw + (lane >> 5); r < ncand; r += 2 * nwv) {
    const int sl = s_list[r] * CL + (lane & (CL - 1));
    if (slot_orig[sl] < 0) continue;
    const V4<TC> p = xyzq[sl];
    TC Mv[3][P], dM[3][P];
    int base[3];
    unsigned int inmask[3];
    const double rr[3] = {(double)p.x, (double)p.y, (double)p.z};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double sa = (rr[0] * pg.inv[a] + rr[1] * pg.inv[3 + a] + rr[2] * pg.inv[6 + a]) * (double)pg.K[a];
      double t;
      anchor_axis(sa, pg.K[a], P, base[a], t);
      spline_weights<TC, P>((TC)t, Mv[a], dM[a]);
      unsigned int m = 0;
#pragma unroll
      for (int k = 0; k < P; ++k) {
        int gk = base[a] + k;
        gk = gk >= pg.K[a] ? gk - pg.K[a] : gk;
        if (gk >= B0[a] && gk < B0[a] + Bn[a]) m |= 1u << k;
      }
      inmask[a] = m;
    }
    if (!inmask[0] || !inmask[1] || !inmask[2]) continue;
    const TC qi = p.w;
#pragma unroll
    for (int ka = 0; ka < P; ++ka) {
      if (!((inmask[0] >> ka) & 1u)) continue;
      int gx = base[0] + ka;
      gx = gx >= K0 ? gx - K0 : gx;
      const TC wa = qi * Mv[0][ka];
#pragma unroll
      for (int kb = 0; kb < P; ++kb) {
        if (!((inmask[1] >> kb) & 1u)) continue;
        int gy = base[1] + kb;
        gy = gy >= K1 ? gy - K1 : gy;
        const TC wab = wa * Mv[1][kb];
        unsigned long long* row = s_q + ((gx - B0[0]) * BR + (gy - B0[1])) * BR;
#pragma unroll
        for (int kc = 0; kc < P; ++kc) {
          if (!((inmask[2] >> kc) & 1u)) continue;
          int gz = base[2] + kc;
          gz = gz >= K2 ? gz - K2 : gz;
          atomicAdd(&row[gz - B0[2]], (unsigned long long)to_fx((double)(wab * Mv[2][kc]), Q_SCALE));
        }
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < BR * BR * BR; k += blockDim.x) {
    const int lz = k % BR, ly = (k / BR) % BR, lx = k / (BR * BR);
    if (lx < Bn[0] && ly < Bn[1] && lz < Bn[2])
      grid[((long long)(B0[0] + lx) * K1 + (B0[1] + ly)) * K2 + (B0[2] + lz)] = s_q[k];
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
      *excl_part, *ncp, *seg, *seg_count, *seg_start, *slot_of_orig;
  void *xyzq, *lj;
  float4 *bb_c, *bb_h, *sb_c, *sb_h;
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
  w.sb_c = (float4*)take(16 * 4 * (size_t)g.ncl_max);
  w.sb_h = (float4*)take(16 * 4 * (size_t)g.ncl_max);
  w.slot_of_orig = (int*)take(4 * nn);
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
int launch_pme_slots(int order, const FusedWs& w, const Geo& g, const PmeGeom& pg, int nslots, const TC* phi,
                     double coul, bool spread, int* status, hipStream_t st) {
  unsigned nb = (unsigned)((nslots + FB - 1) / FB);
  const unsigned nbricks = (unsigned)(((pg.K[0] + BR - 1) / BR) * ((pg.K[1] + BR - 1) / BR) * ((pg.K[2] + BR - 1) / BR));
#define TFF_SP(P)                                                                                               \
  if (spread)                                                                                                   \
    hipLaunchKernelGGL((k_spread_bricks<TC, P>), dim3(nbricks), dim3(256), 0, st, w.ncl, (const V4<TC>*)w.xyzq, \
                       w.slot_orig, w.bb_c, w.bb_h, w.col_cl_start, g, pg, w.grid, status);                      \
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
  zero32(status, 2);
  zero64(w.facc, 3LL * nslots);
  zero64(w.eacc, 4);
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
  const unsigned nbW = (unsigned)(((long long)g.ncl_max * TFF_WAVE + FB - 1) / FB);
  hipLaunchKernelGGL(k_reorder_clusters<TC>, dim3(nbW), dim3(FB), 0, st, w.ncl, (V4<TC>*)w.xyzq, (V2<TC>*)w.lj,
                     w.slot_orig, w.slot_of_orig);
  hipLaunchKernelGGL(k_clear_pad_exclusions, dim3(nbS), dim3(FB), 0, st, w.ncl, w.slot_orig, w.excl_part);
  hipLaunchKernelGGL(k_fill_exclusions, dim3(nbA), dim3(FB), 0, st, n, w.slot_of_orig, row_ptr, col, w.excl_part,
                     status);
  hipLaunchKernelGGL(k_bbox<TC>, dim3(nbW), dim3(FB), 0, st, w.ncl, (const V4<TC>*)w.xyzq, w.slot_orig, w.bb_c,
                     w.bb_h, w.sb_c, w.sb_h);
  // 3. cluster pairs
  {
    void* tok = tff_prof_begin("fused_cluster_pairs", st);
    float rl2 = (float)(list_cutoff * list_cutoff);
    hipLaunchKernelGGL(k_cluster_pairs, dim3(nbW), dim3(FB), 0, st, w.ncl, w.cl_col, w.col_cl_start, w.bb_c,
                       w.bb_h, g, rl2, w.excl_part, w.seg, w.seg_count, status);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, w.ncl, w.seg_count, w.seg_start, w.ncp, status);
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
                       (const V4<TC>*)w.xyzq, (const V2<TC>*)w.lj, w.slot_orig, w.excl_part, w.sb_c, w.sb_h, ta,
                       (float)(list_cutoff * list_cutoff), w.facc, w.eacc);
    tff_prof_end(tok, st);
  }
  // 5. PME
  PmeGeom pg = make_geom(box, K, order, alpha);
  FftPlan pl[3];
  for (int a = 0; a < 3; ++a) make_plan(K[a], pl[a]);
  {
    void* tok = tff_prof_begin("fused_spread", st);
    int rc = launch_pme_slots<TC>(order, w, g, pg, nslots, nullptr, coul, true, status, st);
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
    int rc = launch_pme_slots<TC>(order, w, g, pg, nslots, phi, coul, false, status, st);
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
