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

#include <string.h>

#include <atomic>
#include <mutex>
#include <type_traits>
#include <vector>

#include "pair_math.h"
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
constexpr int E_SLOTS = 64;  // fixed-point energy slots per term
constexpr double F_SCALE = 4294967296.0;        // 2^32: forces and energies (kJ/mol, nm)
constexpr double Q_SCALE = 1099511627776.0;     // 2^40: PME charge grid
// padding slots sit at x = PAD_STEP (1 + slot), y = z = PAD_STEP: every padding slot is more
// than PAD_STEP / 2 from every atom and from every other padding slot under any periodic
// shift of at most one box length (boxes shorter than PAD_LIMIT)
constexpr double PAD_STEP = 1.0e6;
constexpr double PAD_LIMIT = 5.0e5;

template <typename T> struct V4 { T x, y, z, w; };
template <typename T> struct V2 { T x, y; };

struct Geo {
  double L[3], invL[3];
  int ncol[2];
  double colw[2];
  int ncl_max;
  long long pair_cap;
  int maxn;  // cluster-pair segment length per i-cluster
  int reach[2];
};

__device__ __forceinline__ long long to_fx(double v, double s) { return __double2ll_rn(v * s); }

// ---------------------------------------------------------------- 1. binning
// Atoms are binned into x-y columns, the columns are scanned, the atoms scattered into their
// column's range, and every column is sorted by (z, atom index) in LDS, so that the order
// (and everything computed from it) does not depend on the order of the atomics.
constexpr int COL_SORT_MAX = 8192;  // atoms per column sorted in LDS (64 KB of keys)

template <typename TI>
__device__ __forceinline__ void wrap_coords(const TI* __restrict__ coords, long long i, const Geo& g, double w[3]) {
  for (int a = 0; a < 3; ++a) {
    double f = (double)coords[3 * i + a] * g.invL[a];
    f -= floor(f);
    w[a] = f * g.L[a];
  }
}

// Column of every atom and its offset within the column: the lanes of a wave that share a
// column reserve one contiguous range with a single atomic (binning water boxes puts hundreds
// of atoms on each column counter, so per-atom atomics would serialise).
template <typename TI>
__global__ __launch_bounds__(FB) void k_bin_atoms(const TI* __restrict__ coords, const TI* __restrict__ q, long long n,
                                                  Geo g, int* __restrict__ col_of, int* __restrict__ off_in_col,
                                                  int* __restrict__ col_count, double* __restrict__ q2part,
                                                  const TI* __restrict__ box_dev, int* __restrict__ status) {
  const long long i = blockIdx.x * (long long)FB + threadIdx.x;
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const bool valid = i < n;
  if (box_dev && blockIdx.x == 0 && threadIdx.x == 0) {
    // the call was planned (cell grid, capacities, PME geometry) from a host copy of the box: a
    // device box that no longer holds those values marks every result of the call as invalid
    bool same = true;
    for (int k = 0; k < 9; ++k) same = same && box_dev[k] == (k % 4 == 0 ? (TI)g.L[k / 4] : (TI)0);
    if (!same) atomicOr(status, TFF_FUSED_BOX_STALE);
  }
  {
    // sum of q^2 of the block (self energy), one partial per block
    const double qi = valid ? (double)q[i] : 0.0;
    const double t = tff_block_sum(qi * qi);
    if (threadIdx.x == 0) q2part[blockIdx.x] = t;
  }
  int col = -1;
  if (valid) {
    double w[3];
    wrap_coords(coords, i, g, w);
    int cx = (int)(w[0] / g.colw[0]); cx = cx < 0 ? 0 : (cx >= g.ncol[0] ? g.ncol[0] - 1 : cx);
    int cy = (int)(w[1] / g.colw[1]); cy = cy < 0 ? 0 : (cy >= g.ncol[1] ? g.ncol[1] - 1 : cy);
    col = cx * g.ncol[1] + cy;
  }
  int off = 0;
  unsigned long long active = __ballot(valid);
  while (active) {
    const int leader = __ffsll((long long)active) - 1;
    const int c0 = __shfl(col, leader, TFF_WAVE);
    const unsigned long long same = __ballot(valid && col == c0) & active;
    int base = 0;
    if (lane == leader) base = atomicAdd(&col_count[c0], __popcll(same));
    base = __shfl(base, leader, TFF_WAVE);
    if (valid && col == c0) off = base + __popcll(same & ((1ull << lane) - 1ull));
    active &= ~same;
  }
  if (valid) {
    col_of[i] = col;
    off_in_col[i] = off;
  }
}

// exclusive scans over the columns of their atom counts and cluster counts (one workgroup)
__global__ __launch_bounds__(1024) void k_scan_columns(const int* __restrict__ col_count, int ncols,
                                                       int* __restrict__ col_start, int* __restrict__ col_cl_start,
                                                       int* __restrict__ n_clusters, int* __restrict__ status) {
  __shared__ int sa[1024], sc[1024];
  const int t = threadIdx.x;
  const int per = (ncols + 1023) / 1024;
  const int b = t * per, e = b + per < ncols ? b + per : ncols;
  int na = 0, nc = 0;
  bool big = false;
  for (int c = b; c < e; ++c) {
    const int m = col_count[c];
    na += m;
    nc += (m + CL - 1) / CL;
    big = big || m > COL_SORT_MAX;
  }
  if (big) atomicOr(status, TFF_FUSED_COLUMN_OVERFLOW);
  sa[t] = na;
  sc[t] = nc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int va = t >= off ? sa[t - off] : 0, vc = t >= off ? sc[t - off] : 0;
    __syncthreads();
    sa[t] += va;
    sc[t] += vc;
    __syncthreads();
  }
  int ra = t > 0 ? sa[t - 1] : 0, rc = t > 0 ? sc[t - 1] : 0;
  for (int c = b; c < e; ++c) {
    const int m = col_count[c];
    col_start[c] = ra;
    col_cl_start[c] = rc;
    ra += m;
    rc += (m + CL - 1) / CL;
  }
  if (t == 1023) {
    col_start[ncols] = sa[1023];
    col_cl_start[ncols] = sc[1023];
    *n_clusters = sc[1023];
  }
}

__global__ __launch_bounds__(FB) void k_scatter_atoms(long long n, const int* __restrict__ col_of,
                                                      const int* __restrict__ off_in_col,
                                                      const int* __restrict__ col_start, int* __restrict__ sorted_orig) {
  long long i = blockIdx.x * (long long)FB + threadIdx.x;
  if (i >= n) return;
  sorted_orig[col_start[col_of[i]] + off_in_col[i]] = (int)i;
}

// One workgroup per column sorts the keys (float bits of the wrapped z << 32 | atom) in LDS.  The
// keys are distinct (the atom index), so the sorted order is unique.  Bucket sort: the keys are
// counted into up to COL_BUCKETS z buckets (the bucket is monotone in the key), the counts
// scanned, the keys placed by bucket, and every key written to its bucket start plus its rank
// among the keys of its bucket; a column whose largest bucket holds more than BUCKET_RANK_MAX
// keys (atoms on shared z planes) is sorted by a bitonic network over the placed keys instead.
// Columns above COL_SORT_MAX atoms are left in scatter order (flagged by the scan).
constexpr int COL_BUCKETS = 2048;
constexpr int BUCKET_RANK_MAX = 32;
constexpr int SORT_THREADS = 256;

template <typename TI>
__device__ __forceinline__ unsigned long long column_key(const TI* __restrict__ coords, int i, const Geo& g) {
  double w[3];
  wrap_coords(coords, i, g, w);
  return ((unsigned long long)__float_as_uint((float)w[2]) << 32) | (unsigned)i;
}
// z >= 0, so the float bits order as the values and the product is monotone in the key
__device__ __forceinline__ int column_bucket(unsigned long long key, float scale, int nb) {
  const int b = (int)(__uint_as_float((unsigned)(key >> 32)) * scale);
  return b < 0 ? 0 : (b < nb - 1 ? b : nb - 1);
}

template <typename TI>
__global__ __launch_bounds__(SORT_THREADS) void k_sort_columns(const TI* __restrict__ coords, Geo g,
                                                               const int* __restrict__ col_start,
                                                               int* __restrict__ sorted_orig) {
  static_assert(COL_BUCKETS <= 8 * SORT_THREADS, "column sort: eight buckets per thread at most");
  __shared__ unsigned long long s_k[COL_SORT_MAX];
  __shared__ int s_c[COL_BUCKETS];
  __shared__ int s_tot[SORT_THREADS / TFF_WAVE], s_max[SORT_THREADS / TFF_WAVE];
  const int col = blockIdx.x;
  const int cs = col_start[col], m = col_start[col + 1] - cs;
  if (m <= 1 || m > COL_SORT_MAX) return;
  const int t = threadIdx.x, lane = t & (TFF_WAVE - 1), wv = t / TFF_WAVE;
  const int nb = m / 2 < 1 ? 1 : (m / 2 < COL_BUCKETS ? m / 2 : COL_BUCKETS);
  const float scale = (float)nb / (float)g.L[2];
  for (int b = t; b < nb; b += SORT_THREADS) s_c[b] = 0;
  __syncthreads();
  for (int k = t; k < m; k += SORT_THREADS)
    atomicAdd(&s_c[column_bucket(column_key(coords, sorted_orig[cs + k], g), scale, nb)], 1);
  __syncthreads();
  // exclusive scan of the bucket sizes (thread t owns buckets [b0, b1)) and the largest bucket
  const int per = (nb + SORT_THREADS - 1) / SORT_THREADS;
  const int b0 = t * per < nb ? t * per : nb, b1 = b0 + per < nb ? b0 + per : nb;
  int sum = 0, mx = 0;
  for (int b = b0; b < b1; ++b) {
    const int c = s_c[b];
    sum += c;
    mx = c > mx ? c : mx;
  }
  int inc = sum;
  for (int o = 1; o < TFF_WAVE; o <<= 1) {
    const int v = __shfl_up(inc, o, TFF_WAVE);
    if (lane >= o) inc += v;
  }
  for (int o = TFF_WAVE / 2; o > 0; o >>= 1) {
    const int v = __shfl_xor(mx, o, TFF_WAVE);
    mx = v > mx ? v : mx;
  }
  if (lane == TFF_WAVE - 1) {
    s_tot[wv] = inc;
    s_max[wv] = mx;
  }
  __syncthreads();
  int run = inc - sum, big = 0;
  for (int v = 0; v < SORT_THREADS / TFF_WAVE; ++v) {
    if (v < wv) run += s_tot[v];
    big = s_max[v] > big ? s_max[v] : big;
  }
  for (int b = b0; b < b1; ++b) {
    const int c = s_c[b];
    s_c[b] = run;
    run += c;
  }
  __syncthreads();
  // place the keys by bucket; afterwards s_c[b] is the end of bucket b
  for (int k = t; k < m; k += SORT_THREADS) {
    const unsigned long long key = column_key(coords, sorted_orig[cs + k], g);
    s_k[atomicAdd(&s_c[column_bucket(key, scale, nb)], 1)] = key;
  }
  __syncthreads();
  if (big <= BUCKET_RANK_MAX) {
    for (int p = t; p < m; p += SORT_THREADS) {
      const unsigned long long key = s_k[p];
      const int b = column_bucket(key, scale, nb);
      const int s = b > 0 ? s_c[b - 1] : 0, e = s_c[b];
      int r = 0;
      for (int q = s; q < e; ++q) r += s_k[q] < key ? 1 : 0;
      sorted_orig[cs + s + r] = (int)(unsigned)(key & 0xffffffffull);
    }
    return;
  }
  // bitonic network over the placed keys, padded to a power of two (pair k -> keys (lo, lo + stride))
  int np = 1;
  while (np < m) np <<= 1;
  for (int k = m + t; k < np; k += SORT_THREADS) s_k[k] = ~0ull;
  __syncthreads();
  for (int size = 2; size <= np; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = t; k < (np >> 1); k += SORT_THREADS) {
        const int lo = (k & ~(stride - 1)) * 2 + (k & (stride - 1)), hi = lo + stride;
        const bool asc = (lo & size) == 0;
        const unsigned long long x = s_k[lo], y = s_k[hi];
        if ((x > y) == asc) {
          s_k[lo] = y;
          s_k[hi] = x;
        }
      }
      __syncthreads();
    }
  }
  for (int k = t; k < m; k += SORT_THREADS) sorted_orig[cs + k] = (int)(unsigned)(s_k[k] & 0xffffffffull);
}

// zero several int32 buffers in one launch (lengths in 32-bit words)
struct ClearList {
  int* p[8];
  long long n[8];
  int count;
};
__global__ void k_clear(ClearList c) {
  const long long s = (long long)gridDim.x * blockDim.x;
  for (int b = 0; b < c.count; ++b)
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < c.n[b]; i += s) c.p[b][i] = 0;
}

// ---------------------------------------------------------------- 2. cluster arrays
// Two clusters per wave, one per half (lanes 0..31, 32..63 = slots): a cluster takes 32 consecutive atoms of its
// column's z-sorted range (the last cluster of a column is padded).  The slots are then
// ordered into four spatially compact groups of 8: sort by x and split in halves, sort each
// half by y and split, order each group by z (padding slots, keyed +inf, collect at the end;
// ties broken by slot).  Writes the cluster arrays (TC4 x, y, z, q; TC2 sigma / 2,
// 2 sqrt(eps)), slot_orig / slot_of_orig, the cluster's column, and the bounding boxes (centre,
// half extent) of the cluster and of its four groups (an empty group gets half extent -1),
// widened by a relative epsilon so that float rounding of the tests never drops a pair.
template <typename TI, typename TC>
__global__ __launch_bounds__(FB) void k_build_clusters(const TI* __restrict__ coords, const TI* __restrict__ q,
                                                       const TI* __restrict__ sig, const TI* __restrict__ eps,
                                                       Geo g, const int* __restrict__ ncl_ptr,
                                                       const int* __restrict__ col_cl_start,
                                                       const int* __restrict__ col_start,
                                                       const int* __restrict__ sorted_orig, V4<TC>* __restrict__ xyzq,
                                                       V2<TC>* __restrict__ lj, int* __restrict__ slot_orig,
                                                       int* __restrict__ slot_of_orig, int* __restrict__ cl_col,
                                                       float4* __restrict__ bb_c, float4* __restrict__ bb_h,
                                                       float4* __restrict__ sb_c, float4* __restrict__ sb_h) {
  __shared__ float4 s_p[FB / CL][CL];
  const int c = (blockIdx.x * FB + threadIdx.x) / CL;
  const int s = threadIdx.x & (CL - 1);
  const int w = threadIdx.x / CL;
  if (c >= *ncl_ptr) return;
  // column of cluster c
  int lo = 0, hi = g.ncol[0] * g.ncol[1];
  while (hi - lo > 1) {
    const int mid = (lo + hi) / 2;
    if (col_cl_start[mid] <= c) lo = mid; else hi = mid;
  }
  const int col = lo;
  if (s == 0) cl_col[c] = col;
  const int pos = col_start[col] + (c - col_cl_start[col]) * CL + s;
  const int o = pos < col_start[col + 1] ? sorted_orig[pos] : -1;
  V4<TC> p;
  V2<TC> l;
  if (o >= 0) {
    double wv[3];
    wrap_coords(coords, o, g, wv);
    p = {(TC)wv[0], (TC)wv[1], (TC)wv[2], (TC)q[o]};
    l = {(TC)(0.5 * (double)sig[o]), (TC)(2.0 * sqrt((double)eps[o] > 0 ? (double)eps[o] : 0.0))};
  } else {
    p = {(TC)0, (TC)0, (TC)0, (TC)0};
    l = {(TC)0, (TC)0};
  }
  const float inf = 3.0e38f;
  const float kx = o >= 0 ? (float)p.x : inf, ky = o >= 0 ? (float)p.y : inf, kz = o >= 0 ? (float)p.z : inf;
  // rank by x over all 32 slots
  int rx = 0;
  for (int t = 0; t < CL; ++t) {
    float vx = __shfl(kx, t, CL);
    rx += (vx < kx || (vx == kx && t < s)) ? 1 : 0;
  }
  const int half = rx >= CL / 2;
  int ry = 0;
  for (int t = 0; t < CL; ++t) {
    float vy = __shfl(ky, t, CL);
    int ht = __shfl(half, t, CL);
    ry += (ht == half && (vy < ky || (vy == ky && t < s))) ? 1 : 0;
  }
  const int grp = 2 * half + (ry >= CL / 4);
  int rz = 0;
  for (int t = 0; t < CL; ++t) {
    float vz = __shfl(kz, t, CL);
    int gt = 2 * __shfl(half, t, CL);
    int rt = __shfl(ry, t, CL);
    gt += rt >= CL / 4;
    rz += (gt == grp && (vz < kz || (vz == kz && t < s))) ? 1 : 0;
  }
  const int nl = grp * 8 + rz;
  const int ns = c * CL + nl;
  {
    // padding slots sit far away, each at its own place (see PAD_STEP)
    xyzq[ns] = o >= 0 ? p : V4<TC>{(TC)(PAD_STEP * (1.0 + ns)), (TC)PAD_STEP, (TC)PAD_STEP, (TC)0};
    lj[ns] = l;
    slot_orig[ns] = o;
    if (o >= 0) slot_of_orig[o] = ns;
    s_p[w][nl] = make_float4((float)p.x, (float)p.y, (float)p.z, o >= 0 ? 1.f : 0.f);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // bounding boxes of the ordered slots: thread s of the half wave holds slot s (the xor
  // shuffles below stay inside the half)
  const float4 qv = s_p[w][s];
  const bool v = qv.w != 0.f;
  float mn[3], mx[3], gmn[3], gmx[3];
  const float pv[3] = {qv.x, qv.y, qv.z};
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
  if (s == 0) {
    bb_c[c] = make_float4(0.5f * (mn[0] + mx[0]), 0.5f * (mn[1] + mx[1]), 0.5f * (mn[2] + mx[2]), 0.f);
    bb_h[c] = make_float4(0.5f * (mx[0] - mn[0]) * 1.0001f + 1e-6f, 0.5f * (mx[1] - mn[1]) * 1.0001f + 1e-6f,
                          0.5f * (mx[2] - mn[2]) * 1.0001f + 1e-6f, 0.f);
  }
  if ((s & 7) == 0) {
    const int gi = c * 4 + s / 8;
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

// exclusion partners of every slot as packed (cluster << 5 | slot), -1 = none; padding slots
// have none
__global__ __launch_bounds__(FB) void k_fill_exclusions(const int* __restrict__ ncl_ptr,
                                                        const int* __restrict__ slot_orig,
                                                        const int* __restrict__ slot_of_orig, long long n,
                                                        const int64_t* __restrict__ row_ptr,
                                                        const int64_t* __restrict__ col,
                                                        int* __restrict__ excl_part, int* __restrict__ status) {
  const int sl = blockIdx.x * FB + threadIdx.x;
  if (sl >= *ncl_ptr * CL) return;
  const int i = slot_orig[sl];
  int k = 0;
  if (i >= 0 && col) {
    for (long long e = row_ptr[i]; e < row_ptr[i + 1]; ++e) {
      const long long j = col[e];
      if (j == i || j < 0 || j >= n) continue;
      if (k == MAXEX) {
        atomicOr(status, TFF_FUSED_EXCL_OVERFLOW);
        break;
      }
      excl_part[sl * MAXEX + k++] = slot_of_orig[j];
    }
  }
  for (; k < MAXEX; ++k) excl_part[sl * MAXEX + k] = -1;
}

// ---------------------------------------------------------------- 3. cluster pairs
// Cluster-pair entries are (ci, cj, info) with cj >= ci; info packs
//   bits  0..15  sub-tile mask: bit gi * 4 + gj set when groups gi of ci and gj of cj have
//                bounding boxes within the list cutoff (diagonal tiles: gj >= gi only)
//   bits 16..21  periodic shift k_a + 1 (2 bits per axis): x_j + k L is the image of cluster
//                cj closest to cluster ci
//   bit  22      the tile holds excluded pairs
//   bit  23      the shift is not the minimum image of every atom pair of the tile (small
//                boxes): the pair kernel then takes the minimum image per pair
//   bit  24      always set
// The sub-tile masks are filled in by k_compact_pairs; entries whose mask stays empty are
// skipped by the pair kernel.
constexpr unsigned INFO_EXCL = 1u << 22;
constexpr unsigned INFO_MINIMG = 1u << 23;
constexpr unsigned INFO_LISTED = 1u << 24;  // set for every entry (so that info != 0)

// One wave per i-cluster; lanes sweep the clusters of the neighbouring columns in a fixed
// order and write the accepted j-clusters into the i-cluster's own segment.
__global__ __launch_bounds__(FB) void k_cluster_pairs(const int* __restrict__ ncl_ptr, const int* __restrict__ cl_col,
                                                      const int* __restrict__ col_cl_start,
                                                      const float4* __restrict__ bb_c,
                                                      const float4* __restrict__ bb_h, Geo g, float rlist2,
                                                      const int* __restrict__ excl_part,
                                                      int2* __restrict__ seg, int* __restrict__ seg_count,
                                                      int* __restrict__ status) {
  __shared__ int s_xp[FB / TFF_WAVE][CL * MAXEX];
  __shared__ int s_off[FB / TFF_WAVE][2 * TFF_WAVE], s_cs[FB / TFF_WAVE][2 * TFF_WAVE];
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
  const float4 c0v = bb_c[ci], h0 = bb_h[ci];
  const float Lv[3] = {(float)g.L[0], (float)g.L[1], (float)g.L[2]};
  const float iLv[3] = {(float)g.invL[0], (float)g.invL[1], (float)g.invL[2]};
  int nseg = 0;
  int2* my = seg + (long long)ci * g.maxn;
  const int nx = 2 * g.reach[0] + 1 >= g.ncol[0] ? g.ncol[0] : 2 * g.reach[0] + 1;
  const int ny = 2 * g.reach[1] + 1 >= g.ncol[1] ? g.ncol[1] : 2 * g.reach[1] + 1;
  const int x0 = nx == g.ncol[0] ? 0 : cx - g.reach[0];
  const int y0 = ny == g.ncol[1] ? 0 : cy - g.reach[1];
  const int ncc = nx * ny;
  // The z window of ci: clusters whose z extent meets [z_lo - rlist, z_hi + rlist] (or its
  // periodic images) can pass the bounding-box test below; every other one fails it.  The
  // clusters of a column are z ordered (their lower and upper z bounds rise along the column), so
  // per column the window is one index range, or two when it wraps round the box; 0.01 nm of slack
  // keeps rounding from dropping a cluster the test would accept.  The accepted clusters and their
  // order (columns in order, clusters ascending) are those of a scan over whole columns.
  const float rl = sqrtf(rlist2) + 0.01f;
  const double zw_lo = (double)(c0v.z - h0.z) - rl, zw_hi = (double)(c0v.z + h0.z) + rl;
  const bool zall = zw_hi - zw_lo >= g.L[2];
  const float za0 = zw_lo < 0 ? 0.f : (float)zw_lo, zb0 = zw_hi > g.L[2] ? (float)g.L[2] : (float)zw_hi;
  const bool zwrap = !zall && (zw_lo < 0 || zw_hi > g.L[2]);
  const float za1 = zw_lo < 0 ? (float)(zw_lo + g.L[2]) : 0.f;  // second interval [za1, zb1]
  const float zb1 = zw_lo < 0 ? (float)g.L[2] : (float)(zw_hi - g.L[2]);
  // Candidates: the neighbouring columns (up to 64 at a time, one per lane), each as up to two
  // cluster ranges inside the window, clipped to cj >= ci, scanned into one flat range that the
  // wave walks 64 clusters at a time (column-major order).
  for (int c0 = 0; c0 < ncc; c0 += TFF_WAVE) {
    const int cc = c0 + lane;
    int r0s = 0, r0n = 0, r1s = 0, r1n = 0;
    if (cc < ncc) {
      int xx = x0 + cc / ny, yy = y0 + cc % ny;
      xx += xx < 0 ? g.ncol[0] : (xx >= g.ncol[0] ? -g.ncol[0] : 0);
      yy += yy < 0 ? g.ncol[1] : (yy >= g.ncol[1] ? -g.ncol[1] : 0);
      const int col_id = xx * g.ncol[1] + yy;
      const int ce = col_cl_start[col_id + 1];
      const int cs = col_cl_start[col_id];
      // first cluster with z max >= a, last cluster with z min <= b (both monotone in the column)
      auto first_ge = [&](float a) {
        int lo = cs, hi = ce;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (bb_c[mid].z + bb_h[mid].z >= a) hi = mid; else lo = mid + 1;
        }
        return lo;
      };
      auto last_le = [&](float b) {
        int lo = cs - 1, hi = ce - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (bb_c[mid].z - bb_h[mid].z <= b) lo = mid; else hi = mid - 1;
        }
        return lo;
      };
      int a0 = cs, e0 = ce, a1 = 0, e1 = 0;  // [start, end) of up to two ranges
      if (!zall && cs < ce) {
        if (!zwrap) {
          a0 = first_ge(za0);
          e0 = last_le(zb0) + 1;
        } else {
          // [0, b] at the column's start, [a, L) at its end; merged when they meet
          const int l0 = last_le(zw_lo < 0 ? zb0 : zb1) + 1;
          const int f1 = first_ge(zw_lo < 0 ? za1 : za0);
          if (f1 > l0) {
            e0 = l0;
            a1 = f1;
            e1 = ce;
          }
        }
      }
      a0 = a0 > ci ? a0 : ci;
      a1 = a1 > ci ? a1 : ci;
      r0s = a0;
      r0n = e0 > a0 ? e0 - a0 : 0;
      r1s = a1;
      r1n = e1 > a1 ? e1 - a1 : 0;
    }
    const int cnt = r0n + r1n;
    int incl = cnt;
#pragma unroll
    for (int off = 1; off < TFF_WAVE; off <<= 1) {
      const int v = __shfl_up(incl, off, TFF_WAVE);
      if (lane >= off) incl += v;
    }
    const int total = __shfl(incl, TFF_WAVE - 1, TFF_WAVE);
    s_off[w][2 * lane] = incl - cnt;
    s_cs[w][2 * lane] = r0s;
    s_off[w][2 * lane + 1] = incl - cnt + r0n;
    s_cs[w][2 * lane + 1] = r1s;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int nvalid = 2 * (ncc - c0 < TFF_WAVE ? ncc - c0 : TFF_WAVE);
    for (int b = 0; b < total; b += TFF_WAVE) {
      const int k = b + lane;
      unsigned info = 0;
      int cj = -1;
      if (k < total) {
        // last range whose offset is <= k (offsets are non-decreasing; an empty range shares its
        // offset with the next one, which then holds k)
        int lo = 0, hi = nvalid - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (s_off[w][mid] <= k) lo = mid; else hi = mid - 1;
        }
        cj = s_cs[w][lo] + (k - s_off[w][lo]);
        const float4 c1 = bb_c[cj], h1 = bb_h[cj];
        const float d[3] = {c0v.x - c1.x, c0v.y - c1.y, c0v.z - c1.z};
        const float hs[3] = {h0.x + h1.x, h0.y + h1.y, h0.z + h1.z};
        float gap2 = 0.f;
        bool minimg = false;
        unsigned shift = 0;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float kk = rintf(d[a] * iLv[a]);
          const float dm = d[a] - Lv[a] * kk;
          const float gp = fmaxf(fabsf(dm) - hs[a], 0.f);
          gap2 += gp * gp;
          minimg = minimg || !(fabsf(dm) + hs[a] <= 0.49999f * Lv[a]);
          shift |= (unsigned)((int)kk + 1) << (16 + 2 * a);
        }
        if (gap2 <= rlist2) info = shift | (minimg ? INFO_MINIMG : 0u) | INFO_LISTED;
      }
      const bool ok = info != 0;
      unsigned long long bm = __ballot(ok);
      if (ok) {
        bool xf = cj == ci;
        for (int t = 0; t < nxp; ++t) xf = xf || s_xp[w][t] == cj;
        int kk = nseg + __popcll(bm & ((1ull << lane) - 1ull));
        if (kk < g.maxn) my[kk] = make_int2(cj, (int)(info | (xf ? INFO_EXCL : 0u)));
      }
      nseg += __popcll(bm);
    }
    // the next group of columns rewrites s_off / s_cs
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
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
  const int total = tff_block_scan_1024(cnt, *ncl_ptr, start, false);  // start holds ncl entries
  if (threadIdx.x == 0) {
    *ncp = total;
    status[1] = total;
  }
}

// Segments -> one compact cluster-pair list in (ci, segment order) (deterministic), with the
// sub-tile masks: bit gi * 4 + gj is set when the bounding boxes of group gi of ci and group
// gj of cj are within cut2 (minimum image of the group centres: a lower bound on the distance
// of every image of every atom pair of the two groups).  One wave per i-cluster, one lane per
// tile (64 tiles per step): a lane loads its tile's entry, then the four group boxes of cj, and
// forms all 16 bits itself (ci's group boxes are the same for every lane).
__global__ __launch_bounds__(FB) void k_compact_pairs(const int* __restrict__ ncl_ptr, const int2* __restrict__ seg,
                                                      const int* __restrict__ seg_count,
                                                      const int* __restrict__ seg_start, int maxn,
                                                      const float4* __restrict__ sb_c, const float4* __restrict__ sb_h,
                                                      Geo g, float cut2, int4* __restrict__ pairs) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int ci = (blockIdx.x * FB + threadIdx.x) / TFF_WAVE;
  if (ci >= *ncl_ptr) return;
  const int n = seg_count[ci], o = seg_start[ci];
  float4 c0[4], h0[4];
#pragma unroll
  for (int gi = 0; gi < 4; ++gi) {
    c0[gi] = sb_c[ci * 4 + gi];
    h0[gi] = sb_h[ci * 4 + gi];
  }
  const float L0 = (float)g.L[0], L1 = (float)g.L[1], L2 = (float)g.L[2];
  const float i0 = (float)g.invL[0], i1 = (float)g.invL[1], i2 = (float)g.invL[2];
  for (int k = lane; k < n; k += TFF_WAVE) {
    const int2 e = seg[(long long)ci * maxn + k];
    const int cj = e.x;
    unsigned bits = 0;
#pragma unroll
    for (int gj = 0; gj < 4; ++gj) {
      const float4 c1 = sb_c[cj * 4 + gj], h1 = sb_h[cj * 4 + gj];
#pragma unroll
      for (int gi = 0; gi < 4; ++gi) {
        if (h0[gi].x >= 0.f && h1.x >= 0.f && (ci != cj || gj >= gi)) {
          float dx = c0[gi].x - c1.x, dy = c0[gi].y - c1.y, dz = c0[gi].z - c1.z;
          dx -= L0 * rintf(dx * i0);
          dy -= L1 * rintf(dy * i1);
          dz -= L2 * rintf(dz * i2);
          const float gx = fmaxf(fabsf(dx) - h0[gi].x - h1.x, 0.f);
          const float gy = fmaxf(fabsf(dy) - h0[gi].y - h1.y, 0.f);
          const float gz = fmaxf(fabsf(dz) - h0[gi].z - h1.z, 0.f);
          if (gx * gx + gy * gy + gz * gz <= cut2) bits |= 1u << (gi * 4 + gj);
        }
      }
    }
    pairs[o + k] = make_int4(ci, cj, (int)((unsigned)e.y | bits), 0);
  }
}

// ---------------------------------------------------------------- 4. tile kernel

// sum over the 8 lanes sharing lane >> 3 (the result lands in all 8 lanes)
template <int CTRL>
__device__ __forceinline__ float dpp_f(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, false));
}
__device__ __forceinline__ float octet_sum(float v) {
  v += dpp_f<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_f<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_f<0x141>(v);  // row_half_mirror
  return v;
}
// float64: the same lane permutations on both 32-bit halves (VALU moves, no LDS crossbar round trip)
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double octet_sum(double v) {
  v += dpp_d<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp_d<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp_d<0x141>(v);  // row_half_mirror (every lane of a quad holds the quad's sum by now)
  return v;
}
// sum over the 8 lanes sharing lane & 7
template <typename T>
__device__ __forceinline__ T stride8_sum(T v) {
  v += __shfl_xor(v, 8, TFF_WAVE);
  v += __shfl_xor(v, 16, TFF_WAVE);
  v += __shfl_xor(v, 32, TFF_WAVE);
  return v;
}

// The 12 force components f[g][a] (k = 3 g + a) of one slot column are summed over the 8 lanes
// that share the column (RED) and written by those 8 lanes: lane r of the eight takes values
// r and r + 8.  The selects act on the reduced values, so the arrays stay in registers.
template <typename T, typename RED>
__device__ __forceinline__ void reduce_pick2(const T (&f)[4][3], int r, RED red, T& v0, T& v1) {
  const T k0 = red(f[0][0]), k1 = red(f[0][1]), k2 = red(f[0][2]), k3 = red(f[1][0]);
  const T k4 = red(f[1][1]), k5 = red(f[1][2]), k6 = red(f[2][0]), k7 = red(f[2][1]);
  const T k8 = red(f[2][2]), k9 = red(f[3][0]), k10 = red(f[3][1]), k11 = red(f[3][2]);
  const bool b0 = r & 1, b1 = r & 2, b2 = r & 4;
  const T s0 = b0 ? k1 : k0, s1 = b0 ? k3 : k2, s2 = b0 ? k5 : k4, s3 = b0 ? k7 : k6;
  const T t0 = b1 ? s1 : s0, t1 = b1 ? s3 : s2;
  v0 = b2 ? t1 : t0;
  const T w0 = b0 ? k9 : k8, w1 = b0 ? k11 : k10;
  v1 = b1 ? w1 : w0;
}
// offset of value k of a slot column within its cluster's 96 accumulators (slot = g * 8 + col)
__device__ __forceinline__ int acc_offset(int k, int col) { return 24 * (k / 3) + 3 * col + k % 3; }

template <typename TC>
struct TileArgs {
  TC L[3], invL[3];
  TC rc2;
  TC alpha;
  TC coul;
  TC c2alpha;  // (2 / sqrt(pi)) alpha
};

// Up to 16 sub-tiles of 8 x 8 atoms of one cluster-pair tile.  Lane (ii = lane & 7, jj = lane >> 3)
// evaluates pair (slot gi * 8 + ii of ci, slot gj * 8 + jj of cj) of sub-tile b = gi * 4 + gj.
// MINIMG: minimum image per pair and explicit slot validity (else pj is already shifted to
// the nearest image and padding slots sit far away); EXCL: the tile holds excluded pairs.
template <typename TC, bool MINIMG, bool EXCL>
__device__ __forceinline__ void tile_steps(unsigned mask, bool upper, const V4<TC> (&pi)[4], const V2<TC> (&li)[4],
                                           const V4<TC> (&pj)[4], const V2<TC> (&lJ)[4], const unsigned (&exm)[4],
                                           int jj, const TileArgs<TC>& ta, TC (&fi)[4][3], TC (&fj)[4][3], TC& e_c,
                                           TC& e_lj) {
  // j-group major, so that the first steps wait only for the first j-group's loads
#pragma unroll
  for (int b = 0; b < 16; ++b) {
    const int gj = b >> 2, gi = b & 3;
    if (!((mask >> (gi * 4 + gj)) & 1u)) continue;
    TC dx = pi[gi].x - pj[gj].x, dy = pi[gi].y - pj[gj].y, dz = pi[gi].z - pj[gj].z;
    if (MINIMG) {
      dx -= ta.L[0] * rint(dx * ta.invL[0]);
      dy -= ta.L[1] * rint(dy * ta.invL[1]);
      dz -= ta.L[2] * rint(dz * ta.invL[2]);
    }
    TC r2 = dx * dx + dy * dy + dz * dz;
    bool ok = r2 <= ta.rc2;
    if (MINIMG) ok = ok && pi[gi].x < (TC)PAD_LIMIT && pj[gj].x < (TC)PAD_LIMIT;
    if (gi == gj) ok = ok && upper;
    // about a fifth of the sub-tiles kept by the bounding-box test hold no pair within the
    // cutoff: the wave skips them (uniform branch)
    if (__builtin_amdgcn_ballot_w64(ok) == 0) continue;
    bool excl = false;
    if (EXCL) excl = (exm[gi] >> (gj * 8 + jj)) & 1u;
    // pairs that do not count are moved far away: every term below is then exactly zero
    // (exp underflows, s^6 flushes), except (erfc - 1)/r of an excluded pair beyond the
    // cutoff, which is below 1e-12 q_i q_j (and dropped with its sub-tile when no pair of the
    // sub-tile is in range)
    r2 = ok ? r2 : (TC)1e30;
    const TC rinv = rsqrt_t(r2);
    const TC rinv2 = rinv * rinv;
    const TC r = r2 * rinv;
    TC ec, em;
    erfc_exp(ta.alpha * r, ec, em);
    if (EXCL) ec = excl ? ec - (TC)1 : ec;
    const TC qr = pi[gi].w * pj[gj].w * rinv;
    TC ep4 = li[gi].y * lJ[gj].y;
    if (EXCL) ep4 = excl ? (TC)0 : ep4;
    const TC sg = li[gi].x + lJ[gj].x;
    const TC s2 = sg * sg * rinv2;
    const TC s6 = s2 * s2 * s2;
    const TC a6 = ep4 * s6;  // 4 eps s^6
    // -dE/dr / r = [k q_i q_j / r (erfc + (2/sqrt pi) a r e^{-a^2 r^2}) + 4 eps s^6 (12 s^6 - 6)] / r^2
    const TC fs = fma(qr, fma(ta.c2alpha * r, em, ec), a6 * fma((TC)12, s6, (TC)-6)) * rinv2;
    e_c = fma(qr, ec, e_c);
    e_lj += fma(a6, s6, -a6);
    // F_i += fs d, F_j -= fs d   (fs = -dE/dr / r)
    fi[gi][0] = fma(fs, dx, fi[gi][0]);
    fi[gi][1] = fma(fs, dy, fi[gi][1]);
    fi[gi][2] = fma(fs, dz, fi[gi][2]);
    fj[gj][0] = fma(-fs, dx, fj[gj][0]);
    fj[gj][1] = fma(-fs, dy, fj[gj][1]);
    fj[gj][2] = fma(-fs, dz, fj[gj][2]);
  }
}

// One wave per workgroup walks a contiguous run of the (ci-ordered) tile list (the dispatcher
// starts a new wave as soon as one finishes, which balances the uneven tiles).  F_i stays in
// registers while ci repeats; F_j of every tile is summed over the 8 lanes sharing jj and added
// to the int64 fixed-point accumulators.  The wave's energies go to its own slot of a partial
// array, summed in a fixed order afterwards, so the results are bitwise reproducible.  Charges
// are pre-multiplied by the Coulomb constant on the i side; lj holds (sigma / 2, 2 sqrt(eps))
// so that li.y * lJ.y = 4 eps_ij.
template <typename T>
__device__ __forceinline__ T wave_sum_t(T v) {
  v = octet_sum(v);
  return stride8_sum(v);
}

template <typename TC>
__device__ __forceinline__ void tiles_body(const int4* __restrict__ cpairs, const int* __restrict__ ncp_ptr,
                                           long long cap, const V4<TC>* __restrict__ xyzq,
                                           const V2<TC>* __restrict__ lj, const int* __restrict__ excl_part,
                                           const TileArgs<TC>& ta, unsigned long long* __restrict__ facc,
                                           unsigned long long* __restrict__ eacc) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int ii = lane & 7, jj = lane >> 3;
  long long ncp = *ncp_ptr;
  if (ncp > cap) ncp = cap;
  const long long nw = gridDim.x;
  const long long gw = blockIdx.x;
  const long long per = (ncp + nw - 1) / nw;
  const long long t0 = gw * per;
  const long long t1 = t0 + per < ncp ? t0 + per : ncp;
  // accumulator offsets of the two values this lane writes for a j column (jj) / an i column (ii)
  const int oj0 = acc_offset(ii, jj), oj1 = acc_offset(ii + 8 < 12 ? ii + 8 : 11, jj);
  const int oi0 = acc_offset(jj, ii), oi1 = acc_offset(jj + 8 < 12 ? jj + 8 : 11, ii);
  TC fi[4][3];
  V4<TC> pi[4];
  V2<TC> li[4];
  auto flush_i = [&](int c) {
    TC v0, v1;
    reduce_pick2(fi, jj, [](TC v) { return stride8_sum(v); }, v0, v1);
    unsigned long long* base = facc + (long long)c * (3 * CL);
    if (v0 != (TC)0) atomicAdd(base + oi0, (unsigned long long)to_fx((double)v0, F_SCALE));
    if (jj < 4 && v1 != (TC)0) atomicAdd(base + oi1, (unsigned long long)to_fx((double)v1, F_SCALE));
  };
  // energies: float within a tile (at most 16 terms per lane), then double across tiles
  double e_lj = 0.0, e_c = 0.0;
  int cur_i = -1;
  {
    for (long long t = t0; t < t1; ++t) {
      const int4 cp = cpairs[t];
      const int ci = cp.x, cj = cp.y;
      const unsigned info = (unsigned)cp.z;
      const unsigned mask = info & 0xffffu;
      if (mask == 0) continue;
      if (ci != cur_i) {
        if (cur_i >= 0) flush_i(cur_i);
        cur_i = ci;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          const int sl = ci * CL + g * 8 + ii;
          pi[g] = xyzq[sl];
          pi[g].w *= ta.coul;
          li[g] = lj[sl];
          fi[g][0] = fi[g][1] = fi[g][2] = 0;
        }
      }
      const bool upper = ci != cj || jj > ii;
      V4<TC> pj[4];
      V2<TC> lJ[4];
#pragma unroll
      for (int g = 0; g < 4; ++g) {
        const int sl = cj * CL + g * 8 + jj;
        pj[g] = xyzq[sl];
        lJ[g] = lj[sl];
      }
      unsigned exm[4] = {0u, 0u, 0u, 0u};
      if (info & INFO_EXCL) {
        // exclusion bit masks of this lane's four i-slots against the slots of cj
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
      TC fj[4][3];
#pragma unroll
      for (int g = 0; g < 4; ++g) fj[g][0] = fj[g][1] = fj[g][2] = 0;
      TC t_lj = 0, t_c = 0;
      if (info & INFO_MINIMG) {
        tile_steps<TC, true, true>(mask, upper, pi, li, pj, lJ, exm, jj, ta, fi, fj, t_c, t_lj);
      } else {
        const TC sx = ta.L[0] * (TC)((int)((info >> 16) & 3u) - 1);
        const TC sy = ta.L[1] * (TC)((int)((info >> 18) & 3u) - 1);
        const TC sz = ta.L[2] * (TC)((int)((info >> 20) & 3u) - 1);
#pragma unroll
        for (int g = 0; g < 4; ++g) {
          pj[g].x += sx;
          pj[g].y += sy;
          pj[g].z += sz;
        }
        if (info & INFO_EXCL)
          tile_steps<TC, false, true>(mask, upper, pi, li, pj, lJ, exm, jj, ta, fi, fj, t_c, t_lj);
        else
          tile_steps<TC, false, false>(mask, upper, pi, li, pj, lJ, exm, jj, ta, fi, fj, t_c, t_lj);
      }
      e_lj += (double)t_lj;
      e_c += (double)t_c;
      TC v0, v1;
      reduce_pick2(fj, ii, [](TC v) { return octet_sum(v); }, v0, v1);
      unsigned long long* base = facc + (long long)cj * (3 * CL);
      if (v0 != (TC)0) atomicAdd(base + oj0, (unsigned long long)to_fx((double)v0, F_SCALE));
      if (ii < 4 && v1 != (TC)0) atomicAdd(base + oj1, (unsigned long long)to_fx((double)v1, F_SCALE));
    }
    if (cur_i >= 0) flush_i(cur_i);
  }
  // the wave's energies go to one of E_SLOTS fixed-point slots per term (spread to keep the
  // atomics uncontended; integer sums do not depend on their order)
  const double el = wave_sum_t(e_lj), ecl = wave_sum_t(e_c);
  if (lane == 0 && t1 > t0) {
    atomicAdd(&eacc[gw % E_SLOTS], (unsigned long long)to_fx(el, F_SCALE));
    atomicAdd(&eacc[E_SLOTS + gw % E_SLOTS], (unsigned long long)to_fx(ecl, F_SCALE));
  }
}

// float: registers capped for 4 waves per SIMD; float64 keeps its full register set
__global__ __launch_bounds__(TFF_WAVE) __attribute__((amdgpu_waves_per_eu(4))) void k_tiles_f32(
    const int4* __restrict__ cpairs, const int* __restrict__ ncp_ptr, long long cap, const V4<float>* __restrict__ xyzq,
    const V2<float>* __restrict__ lj, const int* __restrict__ excl_part, TileArgs<float> ta,
    unsigned long long* __restrict__ facc, unsigned long long* __restrict__ eacc) {
  tiles_body<float>(cpairs, ncp_ptr, cap, xyzq, lj, excl_part, ta, facc, eacc);
}
__global__ __launch_bounds__(TFF_WAVE) void k_tiles_f64(const int4* __restrict__ cpairs,
                                                       const int* __restrict__ ncp_ptr, long long cap,
                                                       const V4<double>* __restrict__ xyzq,
                                                       const V2<double>* __restrict__ lj,
                                                       const int* __restrict__ excl_part, TileArgs<double> ta,
                                                       unsigned long long* __restrict__ facc,
                                                       unsigned long long* __restrict__ eacc) {
  tiles_body<double>(cpairs, ncp_ptr, cap, xyzq, lj, excl_part, ta, facc, eacc);
}

// ---------------------------------------------------------------- 4b. force-matching backward
// For upstream gradients gE (of E) and v (of F = -dE/dr), the vector-Jacobian product of (E, F):
//   dL/dtheta = gE dE/dtheta - d(D_v E)/dtheta,   dL/dr = gE dE/dr - grad(D_v E),
// with D_v E = sum_i v_i . grad_i E.  For a pair (d = r_i - r_j, dv = v_i - v_j, s = d . dv) with
// energy phi(r), f = phi'/r and h = f'/r:  D_v E_pair = f s,  grad_i(f s) = f dv + s h d.
//   Coulomb  phi = k q_i q_j c/r, c = erfc(a r) - [excluded],  g = c + (2/sqrt pi) a r e^{-a^2 r^2}:
//            f = -k q_i q_j g / r^3,  h = k q_i q_j (3 g / r^5 + (4/sqrt pi) a^3 e^{-a^2 r^2} / r^2)
//   LJ       phi = 4 eps (s12 - s6):  f = -4 eps (12 s12 - 6 s6) / r^2,
//            h = 4 eps (168 s12 - 48 s6) / r^4;  d/dsigma_i = 1/2 d/dsigma_ij,
//            d/deps_i = (.)/(2 eps_i) (0 where eps_i = 0).
// Per slot six int64 accumulators: d/dr (3), d/dq, d/dsigma, d/deps.
constexpr int GN = 6;
static_assert(4 * GN == 24, "a cluster column holds 24 backward sums per lane");

// per slot: the upstream force v (x, y, z) and 1 / (2 eps) (w; 0 where eps = 0), from lj's 2 sqrt(eps);
// with status_from set, thread 0 also starts the call's status words from those of the forward call
// whose cluster arrays and pair list are reused
template <typename TI, typename TC>
__global__ __launch_bounds__(FB) void k_fill_vslots(const int* __restrict__ ncl_ptr, const int* __restrict__ slot_orig,
                                                    const TI* __restrict__ v, const V2<TC>* __restrict__ lj,
                                                    V4<TC>* __restrict__ vsl, const int* __restrict__ status_from,
                                                    int* __restrict__ status) {
  const int sl = blockIdx.x * FB + threadIdx.x;
  if (status_from && sl == 0) {
    status[0] = status_from[0];
    status[1] = status_from[1];
  }
  if (sl >= *ncl_ptr * CL) return;
  const int i = slot_orig[sl];
  const TC e2 = lj[sl].y;
  const TC ehalf_inv = e2 > (TC)0 ? (TC)2 / (e2 * e2) : (TC)0;
  vsl[sl] = (i >= 0 && v) ? V4<TC>{(TC)v[3 * i], (TC)v[3 * i + 1], (TC)v[3 * i + 2], ehalf_inv}
                          : V4<TC>{(TC)0, (TC)0, (TC)0, ehalf_inv};
}

// the value of the lane whose index within its eight differs in bit B (the partner for B = 2 is the
// mirrored lane, whose lower bits differ too: only the opposite bit 2 matters below)
template <int B>
__device__ __forceinline__ float octet_xchg(float v) {
  return dpp_f<B == 2 ? 0x141 : B == 1 ? 0x4E : 0xB1>(v);
}
template <int B>
__device__ __forceinline__ double octet_xchg(double v) {
  return dpp_d<B == 2 ? 0x141 : B == 1 ? 0x4E : 0xB1>(v);
}
template <bool STRIDE8, int B, typename T>
__device__ __forceinline__ T xchg8(T v) {
  if constexpr (STRIDE8)
    return __shfl_xor(v, 8 << B, TFF_WAVE);
  else
    return octet_xchg<B>(v);
}
// Reduce-scatter of 24 values over eight lanes (contiguous lanes, or lanes at stride 8): lane m of
// the eight ends with the sums of values m, m + 8 and m + 16.  Each of the three exchange steps
// halves what a lane keeps: 21 exchanges where summing every value in every lane takes 72.
template <bool STRIDE8, typename T>
__device__ __forceinline__ void reduce_scatter24(const T (&a)[24], int m, T (&out)[3]) {
  const bool b2 = (m >> 2) & 1, b1 = (m >> 1) & 1, b0 = m & 1;
  T h[12];  // h[4 r + c]: value 8 r + 4 b2 + c
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const T lo = a[8 * r + c], hi = a[8 * r + 4 + c];
      h[4 * r + c] = (b2 ? hi : lo) + xchg8<STRIDE8, 2>(b2 ? lo : hi);
    }
  T q[6];  // q[2 r + c]: value 8 r + 4 b2 + 2 b1 + c
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const T lo = h[4 * r + c], hi = h[4 * r + 2 + c];
      q[2 * r + c] = (b1 ? hi : lo) + xchg8<STRIDE8, 1>(b1 ? lo : hi);
    }
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const T lo = q[2 * r], hi = q[2 * r + 1];
    out[r] = (b0 ? hi : lo) + xchg8<STRIDE8, 0>(b0 ? lo : hi);
  }
}

// One wave per workgroup over a contiguous run of the tile list (as the forward pair kernel);
// one i-group at a time (lane = (ii, jj) holds i-slot gi * 8 + ii against j-slot gj * 8 + jj),
// the tile's periodic shift (minimum image per pair on the tiles flagged for it).
// Per pair, in factored form without divisions (Q = q_i q_j, s6 = (sigma_ij / r)^6):
//   f = -k Q g / r^3 - 4 eps s6 (12 s6 - 6) / r^2
//   h =  k Q (3 g / r^3 + (4/sqrt pi) a^3 e^{-a^2 r^2}) / r^2 + 4 eps s6 (168 s6 - 48) / r^4
//   d phi / d sigma_ij = 4 eps (s6 / sigma_ij) (12 s6 - 6),  s6 / sigma_ij = s^4 sigma_ij / r^2,
//   d (f s) / d sigma_ij = -4 eps (s6 / sigma_ij) (144 s6 - 36) s / r^2,
// and 1 / (2 eps) of every slot precomputed (the w component of the per-slot upstream force).
// The i-side sums stay in registers while ci repeats (the list is ci-ordered) and are flushed once
// per run of tiles; the j-side sums once per tile.  Value q = 6 g + k of a cluster's 24 per lane
// column is component k of slot g * 8 + column.
template <typename TC>
__global__ __launch_bounds__(TFF_WAVE) void k_tiles_grad(const int4* __restrict__ cpairs,
                                                         const int* __restrict__ ncp_ptr, long long cap,
                                                         const V4<TC>* __restrict__ xyzq,
                                                         const V2<TC>* __restrict__ lj,
                                                         const V4<TC>* __restrict__ vsl,
                                                         const int* __restrict__ excl_part, TileArgs<TC> ta,
                                                         const double* __restrict__ gE_ptr,
                                                         unsigned long long* __restrict__ gacc) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int ii = lane & 7, jj = lane >> 3;
  long long ncp = *ncp_ptr;
  if (ncp > cap) ncp = cap;
  const long long nw = gridDim.x, gw = blockIdx.x;
  const long long per = (ncp + nw - 1) / nw;
  const long long t0 = gw * per;
  const long long t1 = t0 + per < ncp ? t0 + per : ncp;
  const TC gE = gE_ptr ? (TC)gE_ptr[0] : (TC)0;
  const TC c2 = (TC)1.1283791670955126;                     // 2 / sqrt(pi)
  const TC ch = (TC)2 * c2 * ta.alpha * ta.alpha * ta.alpha;  // (4 / sqrt pi) a^3
  TC ai[4][GN];
  int cur_i = -1;
  // the 24 sums of a cluster's lane column over the eight lanes sharing it, added to the
  // accumulators of slot (q / GN) * 8 + col (value q = col8 + 8 r of the lane at col8)
  auto flush = [&](const TC (&a)[4][GN], int c, int col, int col8, auto stride8) {
    TC v[3];
    reduce_scatter24<decltype(stride8)::value>(reinterpret_cast<const TC(&)[24]>(a), col8, v);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const int q = col8 + 8 * r;
      if (v[r] != (TC)0) {
        const long long sl = (long long)c * CL + (q / GN) * 8 + col;
        atomicAdd(&gacc[GN * sl + q % GN], (unsigned long long)to_fx((double)v[r], F_SCALE));
      }
    }
  };
  for (long long t = t0; t < t1; ++t) {
    const int4 cp = cpairs[t];
    const int ci = cp.x, cj = cp.y;
    const unsigned info = (unsigned)cp.z;
    const unsigned mask = info & 0xffffu;
    if (mask == 0) continue;
    if (ci != cur_i) {
      if (cur_i >= 0) flush(ai, cur_i, ii, jj, std::true_type{});
      cur_i = ci;
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int k = 0; k < GN; ++k) ai[g][k] = 0;
    }
    const bool upper = ci != cj || jj > ii;
    // tiles flagged for per-pair minimum image keep it; the others take the tile's periodic shift
    const bool minimg = (info & INFO_MINIMG) != 0u;
    const TC shx = minimg ? (TC)0 : ta.L[0] * (TC)((int)((info >> 16) & 3u) - 1);
    const TC shy = minimg ? (TC)0 : ta.L[1] * (TC)((int)((info >> 18) & 3u) - 1);
    const TC shz = minimg ? (TC)0 : ta.L[2] * (TC)((int)((info >> 20) & 3u) - 1);
    V4<TC> pj[4], vj[4];
    V2<TC> lJ[4];
    TC kqj[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      const int sl = cj * CL + g * 8 + jj;
      pj[g] = xyzq[sl];
      pj[g].x += shx;
      pj[g].y += shy;
      pj[g].z += shz;
      lJ[g] = lj[sl];
      vj[g] = vsl[sl];
      kqj[g] = ta.coul * pj[g].w;
    }
    TC aj[4][GN];
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int k = 0; k < GN; ++k) aj[g][k] = 0;
#pragma unroll
    for (int gi = 0; gi < 4; ++gi) {
      if (!((mask >> (gi * 4)) & 0xfu)) continue;
      const int sli = ci * CL + gi * 8 + ii;
      const V4<TC> pi = xyzq[sli];
      const V2<TC> li = lj[sli];
      const V4<TC> vi = vsl[sli];
      unsigned exm = 0u;
      if (info & INFO_EXCL) {
#pragma unroll
        for (int k = 0; k < MAXEX; ++k) {
          const int pv = excl_part[sli * MAXEX + k];
          if (pv >= 0 && (pv >> 5) == cj) exm |= 1u << (pv & 31);
        }
      }
      TC (&a)[GN] = ai[gi];
      const TC kqi = ta.coul * pi.w;
#pragma unroll
      for (int gj = 0; gj < 4; ++gj) {
        if (!((mask >> (gi * 4 + gj)) & 1u)) continue;
        TC dx = pi.x - pj[gj].x, dy = pi.y - pj[gj].y, dz = pi.z - pj[gj].z;
        if (minimg) {  // uniform over the wave
          dx -= ta.L[0] * rint(dx * ta.invL[0]);
          dy -= ta.L[1] * rint(dy * ta.invL[1]);
          dz -= ta.L[2] * rint(dz * ta.invL[2]);
        }
        TC r2 = dx * dx + dy * dy + dz * dz;
        bool ok = r2 <= ta.rc2 && pi.x < (TC)PAD_LIMIT && pj[gj].x < (TC)PAD_LIMIT;
        if (gi == gj) ok = ok && upper;
        if (__builtin_amdgcn_ballot_w64(ok) == 0) continue;  // no pair of the sub-tile in range
        const bool excl = (exm >> (gj * 8 + jj)) & 1u;
        r2 = ok ? r2 : (TC)1e30;
        const TC rinv = rsqrt_t(r2);
        const TC rinv2 = rinv * rinv;
        const TC u = ta.alpha * (r2 * rinv);
        TC ec, em;
        erfc_exp(u, ec, em);
        ec = excl ? ec - (TC)1 : ec;                       // c = erfc - [excluded]
        const TC grinv3 = fma(c2 * u, em, ec) * rinv * rinv2;  // g / r^3
        const TC kQ = kqi * pj[gj].w;
        const TC eps4 = excl ? (TC)0 : li.y * lJ[gj].y;
        const TC sg = li.x + lJ[gj].x;
        const TC tq = sg * rinv2;  // s^2 / sigma_ij
        const TC s2 = sg * tq;
        const TC s4 = s2 * s2;
        const TC s6 = s4 * s2;
        const TC ps = s4 * tq;     // s6 / sigma_ij
        const TC a12 = fma((TC)12, s6, (TC)-6);
        const TC e6 = eps4 * s6;
        const TC f = fma(-kQ, grinv3, -e6 * a12 * rinv2);
        const TC h = fma(kQ * rinv2, fma((TC)3, grinv3, ch * em), e6 * fma((TC)168, s6, (TC)-48) * (rinv2 * rinv2));
        const TC okf = ok ? (TC)1 : (TC)0;
        const TC dvx = vi.x - vj[gj].x, dvy = vi.y - vj[gj].y, dvz = vi.z - vj[gj].z;
        const TC sdv = dx * dvx + dy * dvy + dz * dvz;
        const TC cd = okf * fma(gE, f, -sdv * h), cv = okf * f;
        const TC gx = cd * dx - cv * dvx, gy = cd * dy - cv * dvy, gz = cd * dz - cv * dvz;
        a[0] += gx;
        a[1] += gy;
        a[2] += gz;
        aj[gj][0] -= gx;
        aj[gj][1] -= gy;
        aj[gj][2] -= gz;
        // d/dq_i = k q_j (gE c / r + (s.dv) g / r^3), d/dq_j alike with k q_i
        const TC gq = okf * fma(gE, ec * rinv, sdv * grinv3);
        a[3] = fma(kqj[gj], gq, a[3]);
        aj[gj][3] = fma(kqi, gq, aj[gj][3]);
        const TC rs = sdv * rinv2;
        const TC ts = okf * (TC)0.5 * eps4 * ps * fma(gE, a12, rs * fma((TC)144, s6, (TC)-36));
        a[4] += ts;
        aj[gj][4] += ts;
        const TC te = okf * e6 * fma(gE, s6 - (TC)1, rs * a12);
        a[5] = fma(te, vi.w, a[5]);
        aj[gj][5] = fma(te, vj[gj].w, aj[gj][5]);
      }
    }
    flush(aj, cj, jj, ii, std::false_type{});
  }
  if (cur_i >= 0) flush(ai, cur_i, ii, jj, std::true_type{});
}

// ---------------------------------------------------------------- 5. PME on cluster slots
// Owner-computes spreading: one workgroup per BR^3 block ("brick") of grid points.  The
// workgroup lists the clusters whose stencils can reach its brick (column neighbourhood,
// circular interval tests in grid units), accumulates their charges into the brick in LDS
// with int64 fixed-point atomics (order independent), then stores the brick: every grid
// point is written once, so the grid needs neither zeroing nor global atomics.
constexpr int BR = 8;
// A brick's grid points in LDS (the spread kernels): x fastest, rows of BR + 1 and planes of
// (BR + 1) BR + 1 int64 words, so that lanes adding to points at other x, y or z fall on different
// banks (x slowest without padding put every x of a row on one bank: 512 bytes apart).  Measured at
// 96k atoms, the list spread alone: 34.2 us against 41.1.
constexpr int SQ_Y = BR + 1, SQ_Z = SQ_Y * BR + 1;
constexpr int BRICK_MAXC = 1024;
constexpr int SPREAD_CHUNK = 32;  // candidate clusters compacted per round
constexpr int SPREAD_MAXCOL = 64; // candidate columns scanned per round

// Charge-grid fixed point: float mode scales by 2^28 and converts through int32 (a spline-weighted
// charge |q M M M| < 8 e, i.e. |q| < ~36 e for order >= 4, fits); float64 mode scales by 2^40.
constexpr double QF_SCALE = 268435456.0;
__device__ __forceinline__ long long grid_fx(float v) { return (long long)__float2int_rn(v * (float)QF_SCALE); }
__device__ __forceinline__ long long grid_fx(double v) { return to_fx(v, Q_SCALE); }
template <typename TC> constexpr double grid_scale() { return sizeof(TC) == 4 ? QF_SCALE : Q_SCALE; }

// x mod K: two conditional corrections each way cover -2K <= x < 3K (grid indices within a
// few stencil widths of [0, K)); anything further out (tiny grids) takes the division
__device__ __forceinline__ int mod_near(int x, int K) {
  x = x < 0 ? x + K : x;
  x = x < 0 ? x + K : x;
  x = x >= K ? x - K : x;
  x = x >= K ? x - K : x;
  if (x < 0 || x >= K) x = ((x % K) + K) % K;
  return x;
}

// do the circular index intervals [lo, lo + len) and [b0, b0 + b) of Z_K intersect?
__device__ __forceinline__ bool circ_overlap(int lo, int len, int b0, int b, int K) {
  if (len >= K) return true;
  const int d1 = mod_near(b0 - lo, K);
  const int d2 = mod_near(lo - b0, K);
  return d1 < len || d2 < b;
}

// grid-index interval [lo, lo + len) touched by the stencils of atoms with scaled coordinate
// in [slo, shi] (spline order P, anchor of the spline weights in tffpme::anchor_axis)
__device__ __forceinline__ void stencil_interval(float slo, float shi, int P, int& lo, int& len) {
  lo = (int)floorf(slo) - P / 2 - 1;
  int hi = (int)ceilf(shi) + P - P / 2;
  len = hi - lo + 1;
}

// Lists in s_list (at most BRICK_MAXC entries; *s_n gets the full count) the clusters whose
// stencils can reach the brick at B0 of extent Bn.  Called by every thread of the workgroup; *s_n
// must be zero (and visible) on entry.
template <int P>
__device__ __forceinline__ void brick_clusters(const float4* __restrict__ bb_c, const float4* __restrict__ bb_h,
                                               const int* __restrict__ col_cl_start, const Geo& g,
                                               const PmeGeom& pg, const int (&B0)[3], const int (&Bn)[3],
                                               int* s_list, int* s_coff, int* s_ccs, int* s_cln, int* s_n) {
  // candidate columns: those whose x-y extent can reach the brick
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
  // the brick's z window (stencil reach plus two grid points), as up to two intervals of [0, L)
  const double zw_lo = (B0[2] - P / 2 - 2) * g.L[2] / pg.K[2];
  const double zw_hi = (B0[2] + Bn[2] + P / 2 + 2) * g.L[2] / pg.K[2];
  const bool zall = zw_hi - zw_lo >= g.L[2];
  const float za0 = zw_lo < 0 ? 0.f : (float)zw_lo, zb0 = zw_hi > g.L[2] ? (float)g.L[2] : (float)zw_hi;
  const bool zwrap = !zall && (zw_lo < 0 || zw_hi > g.L[2]);
  const float za1 = zw_lo < 0 ? (float)(zw_lo + g.L[2]) : 0.f;          // second interval [za1, zb1]
  const float zb1 = zw_lo < 0 ? (float)g.L[2] : (float)(zw_hi - g.L[2]);
  // the candidate columns' cluster sub-ranges inside that window (clusters of a column are z
  // ordered: binary searches, one thread per column), scanned into one flat range (columns in
  // groups of SPREAD_MAXCOL), then every candidate cluster tested in one pass
  for (int g0 = 0; g0 < ncand_cols; g0 += SPREAD_MAXCOL) {
    const int ngc = ncand_cols - g0 < SPREAD_MAXCOL ? ncand_cols - g0 : SPREAD_MAXCOL;
    if (threadIdx.x < ngc) {
      const int cc = g0 + threadIdx.x;
      int xx = (c0[0] + cc / nc[1]) % g.ncol[0];
      if (xx < 0) xx += g.ncol[0];
      int yy = (c0[1] + cc % nc[1]) % g.ncol[1];
      if (yy < 0) yy += g.ncol[1];
      const int col = xx * g.ncol[1] + yy;
      const int cs = col_cl_start[col], ce = col_cl_start[col + 1];
      // first cluster with z max >= a, last cluster with z min <= b (both monotone in the column)
      auto first_ge = [&](float a) {
        int lo = cs, hi = ce;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (bb_c[mid].z + bb_h[mid].z >= a) hi = mid; else lo = mid + 1;
        }
        return lo;
      };
      auto last_le = [&](float b) {
        int lo = cs - 1, hi = ce - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (bb_c[mid].z - bb_h[mid].z <= b) lo = mid; else hi = mid - 1;
        }
        return lo;
      };
      int r0s = cs, r0e = ce, r1s = 0, r1e = 0;  // [start, end) of up to two ranges
      if (!zall && cs < ce) {
        if (!zwrap) {
          r0s = first_ge(za0);
          r0e = last_le(zb0) + 1;
        } else {
          // [0, b] at the column's start, [a, L) at its end; merged when they meet
          const int l0 = last_le(zw_lo < 0 ? zb0 : zb1) + 1;
          const int f1 = first_ge(zw_lo < 0 ? za1 : za0);
          if (f1 <= l0) {
            r0s = cs;
            r0e = ce;
          } else {
            r0s = cs;
            r0e = l0;
            r1s = f1;
            r1e = ce;
          }
        }
      }
      s_ccs[2 * threadIdx.x] = r0s;
      s_cln[2 * threadIdx.x] = r0e > r0s ? r0e - r0s : 0;
      s_ccs[2 * threadIdx.x + 1] = r1s;
      s_cln[2 * threadIdx.x + 1] = r1e > r1s ? r1e - r1s : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int run = 0;
      for (int t = 0; t < 2 * ngc; ++t) {
        s_coff[t] = run;
        run += s_cln[t];
      }
      s_coff[2 * ngc] = run;
    }
    __syncthreads();
    const int nent = 2 * ngc;
    const int total = s_coff[nent];
    for (int k = threadIdx.x; k < total; k += blockDim.x) {
      int lo = 0, hi = nent - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_coff[mid] <= k) lo = mid; else hi = mid - 1;
      }
      const int c = s_ccs[lo] + (k - s_coff[lo]);
      const float4 cc4 = bb_c[c], hh4 = bb_h[c];
      const float cv[3] = {cc4.x, cc4.y, cc4.z}, hv[3] = {hh4.x, hh4.y, hh4.z};
      bool ok = true;
      for (int a = 0; a < 3; ++a) {
        int lo2, len;
        stencil_interval((cv[a] - hv[a]) * kl[a], (cv[a] + hv[a]) * kl[a], P, lo2, len);
        ok = ok && circ_overlap(lo2, len, B0[a], Bn[a], pg.K[a]);
      }
      if (ok) {
        const int q = atomicAdd(s_n, 1);
        if (q < BRICK_MAXC) s_list[q] = c;
      }
    }
    __syncthreads();
  }
}

// DIPOLE: instead of the charges, spread the dipoles q_i v_i . grad_r theta_i(m) (the grid of the
// force-matching backward), from per-slot vectors vsl and the Cartesian-to-grid scale jsc = K / L.
template <typename TC, int P, bool DIPOLE>
__global__ __launch_bounds__(256) void k_spread_bricks(const int* __restrict__ ncl_ptr, const V4<TC>* __restrict__ xyzq,
                                                       const int* __restrict__ slot_orig, const float4* __restrict__ bb_c,
                                                       const float4* __restrict__ bb_h,
                                                       const int* __restrict__ col_cl_start, Geo g, PmeGeom pg,
                                                       unsigned long long* __restrict__ grid, int* __restrict__ status,
                                                       const V4<TC>* __restrict__ vsl, V4<TC> jsc) {
  __shared__ unsigned long long s_q[SQ_Z * BR];
  __shared__ int s_list[BRICK_MAXC];
  __shared__ V4<TC> s_pos[SPREAD_CHUNK * CL];
  __shared__ int s_slot[DIPOLE ? SPREAD_CHUNK * CL : 1];
  __shared__ int s_coff[2 * SPREAD_MAXCOL + 1], s_ccs[2 * SPREAD_MAXCOL], s_cln[2 * SPREAD_MAXCOL];
  __shared__ int s_n, s_na;
  const int K0 = pg.K[0], K1 = pg.K[1], K2 = pg.K[2];
  const int nbx = (K0 + BR - 1) / BR, nby = (K1 + BR - 1) / BR, nbz = (K2 + BR - 1) / BR;
  const int bid = blockIdx.x;
  if (bid >= nbx * nby * nbz) return;
  const int bz = bid % nbz, by = (bid / nbz) % nby, bx = bid / (nbz * nby);
  const int B0[3] = {bx * BR, by * BR, bz * BR};
  const int Bn[3] = {K0 - B0[0] < BR ? K0 - B0[0] : BR, K1 - B0[1] < BR ? K1 - B0[1] : BR,
                     K2 - B0[2] < BR ? K2 - B0[2] : BR};
  for (int k = threadIdx.x; k < SQ_Z * BR; k += blockDim.x) s_q[k] = 0ull;
  if (threadIdx.x == 0) s_n = 0;
  __syncthreads();
  brick_clusters<P>(bb_c, bb_h, col_cl_start, g, pg, B0, Bn, s_list, s_coff, s_ccs, s_cln, &s_n);
  int ncand = s_n;
  if (ncand > BRICK_MAXC) {
    if (threadIdx.x == 0) atomicOr(status, TFF_FUSED_BRICK_OVERFLOW);
    ncand = BRICK_MAXC;
  }
  // Candidate clusters in chunks of SPREAD_CHUNK: the slots whose stencil reaches the brick are
  // compacted into s_pos (a cheap anchor test), then one lane per compacted slot adds its P^3
  // contributions to the brick (LDS int64 atomics: the sums do not depend on the list order).
  for (int r0 = 0; r0 < ncand; r0 += SPREAD_CHUNK) {
    if (threadIdx.x == 0) s_na = 0;
    __syncthreads();
    const int nslot = (ncand - r0 < SPREAD_CHUNK ? ncand - r0 : SPREAD_CHUNK) * CL;
    for (int k = threadIdx.x; k < nslot; k += blockDim.x) {
      const int sl = s_list[r0 + k / CL] * CL + (k % CL);
      if (slot_orig[sl] < 0) continue;
      const V4<TC> p = xyzq[sl];
      const double rr[3] = {(double)p.x, (double)p.y, (double)p.z};
      bool ov = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double sa = (rr[0] * pg.inv[a] + rr[1] * pg.inv[3 + a] + rr[2] * pg.inv[6 + a]) * (double)pg.K[a];
        int base;
        double t;
        anchor_axis(sa, pg.K[a], P, base, t);
        ov = ov && circ_overlap(base, P, B0[a], Bn[a], pg.K[a]);
      }
      if (ov) {
        const int idx = atomicAdd(&s_na, 1);
        s_pos[idx] = p;
        if (DIPOLE) s_slot[idx] = sl;
      }
    }
    __syncthreads();
    const int na = s_na;
    for (int k = threadIdx.x; k < na; k += blockDim.x) {
      const V4<TC> p = s_pos[k];
      TC Mv[3][P], dM[3][P];
      int base[3];
      unsigned int inmask[3];
      const double rr[3] = {(double)p.x, (double)p.y, (double)p.z};
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const double sa = (rr[0] * pg.inv[a] + rr[1] * pg.inv[3 + a] + rr[2] * pg.inv[6 + a]) * (double)pg.K[a];
        double t;
        anchor_axis(sa, pg.K[a], P, base[a], t);
        spline_weights<TC, P>((TC)t, Mv[a], dM[a]);
        unsigned int m = 0;
#pragma unroll
        for (int q = 0; q < P; ++q) {
          int gk = base[a] + q;
          gk = gk >= pg.K[a] ? gk - pg.K[a] : gk;
          if (gk >= B0[a] && gk < B0[a] + Bn[a]) m |= 1u << q;
        }
        inmask[a] = m;
      }
      const TC qi = p.w;
      TC wv[3] = {(TC)0, (TC)0, (TC)0};
      if (DIPOLE) {
        const V4<TC> vv = vsl[s_slot[k]];
        // d theta / d r_a = -K_a / L_a M'_a ... (the weights fall as the scaled coordinate grows)
        wv[0] = -qi * vv.x * jsc.x;
        wv[1] = -qi * vv.y * jsc.y;
        wv[2] = -qi * vv.z * jsc.z;
      }
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
          // dipole: value(kc) = cz * M_z(kc) + dz * M_z'(kc)
          const TC cz = DIPOLE ? wv[0] * dM[0][ka] * Mv[1][kb] + wv[1] * Mv[0][ka] * dM[1][kb] : (TC)0;
          const TC dz = DIPOLE ? wv[2] * Mv[0][ka] * Mv[1][kb] : (TC)0;
          unsigned long long* row = s_q + (gy - B0[1]) * SQ_Y + (gx - B0[0]);
#pragma unroll
          for (int kc = 0; kc < P; ++kc) {
            if (!((inmask[2] >> kc) & 1u)) continue;
            int gz = base[2] + kc;
            gz = gz >= K2 ? gz - K2 : gz;
            if (DIPOLE)
              atomicAdd(&row[(gz - B0[2]) * SQ_Z], (unsigned long long)to_fx((double)(cz * Mv[2][kc] + dz * dM[2][kc]),
                                                                            Q_SCALE));
            else
              atomicAdd(&row[(gz - B0[2]) * SQ_Z], (unsigned long long)grid_fx(wab * Mv[2][kc]));
          }
        }
      }
    }
    __syncthreads();
  }
  __syncthreads();
  for (int k = threadIdx.x; k < BR * BR * BR; k += blockDim.x) {
    const int lz = k % BR, ly = (k / BR) % BR, lx = k / (BR * BR);
    if (lx < Bn[0] && ly < Bn[1] && lz < Bn[2])
      grid[((long long)(B0[0] + lx) * K1 + (B0[1] + ly)) * K2 + (B0[2] + lz)] = s_q[lz * SQ_Z + ly * SQ_Y + lx];
  }
}

// ---- charge spread of the float path: per-slot spline tables + home-brick lists
// A slot's home brick is the brick holding its stencil base (the grid point of its first spline
// weight on every axis).  k_slot_splines evaluates each slot's spline weights once (the
// arithmetic of k_spread_bricks and k_gather_slots) into wts[sl * WST ...] = M(t_x + k),
// M(t_y + k), M(t_z + k) for k < P, then q, and appends (slot, packed base) to its home brick's
// list (capacity cap per brick; the lanes of a wave that share a home brick reserve one range
// with one atomic).  k_spread_lists then spreads each brick from the lists of the home bricks whose
// stencils can reach it: no search over clusters.  Needs P <= 6 and BR + P <= K_a <= 1023 (every
// stencil meets a brick in one run of grid points; bases fit 10 bits).
template <int P> constexpr int spline_stride() { return (3 * P + 1 + 3) & ~3; }
constexpr int LIST_WST = 20;  // spline_stride<6>(): the table stride of the largest order served

struct BrickGrid {
  int nb[3];  // bricks per axis
  int cap;    // list entries per brick
};

template <int P>
__global__ __launch_bounds__(FB) void k_slot_splines(const int* __restrict__ ncl_ptr, const V4<float>* __restrict__ xyzq,
                                                     const int* __restrict__ slot_orig, PmeGeom pg, BrickGrid bg,
                                                     float* __restrict__ wts, int* __restrict__ bcount,
                                                     int2* __restrict__ blist, int* __restrict__ status) {
  const int sl = blockIdx.x * FB + threadIdx.x;
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const bool valid = sl < *ncl_ptr * CL && slot_orig[sl] >= 0;
  int hb = -1, packed = 0;
  if (valid) {
    const V4<float> p = xyzq[sl];
    const double rr[3] = {(double)p.x, (double)p.y, (double)p.z};
    int base[3];
    float* w = wts + (size_t)sl * LIST_WST;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double sa = (rr[0] * pg.inv[a] + rr[1] * pg.inv[3 + a] + rr[2] * pg.inv[6 + a]) * (double)pg.K[a];
      double t;
      float Mv[P], dM[P];
      anchor_axis(sa, pg.K[a], P, base[a], t);
      spline_weights<float, P>((float)t, Mv, dM);
#pragma unroll
      for (int k = 0; k < P; ++k) w[a * P + k] = Mv[k];
    }
    w[3 * P] = p.w;
    hb = ((base[0] / BR) * bg.nb[1] + base[1] / BR) * bg.nb[2] + base[2] / BR;
    packed = base[0] | (base[1] << 10) | (base[2] << 20);
  }
  // wave-aggregated append: group the lanes by home brick (ballots only), then every group's first
  // lane reserves the group's range, all groups at once
  unsigned long long todo = __ballot(valid), grp = 0;
  int leader = 0;
  while (todo) {
    const int l = __ffsll((long long)todo) - 1;
    const int h = __shfl(hb, l, TFF_WAVE);
    const unsigned long long m = __ballot(valid && hb == h);
    todo &= ~m;
    if (valid && hb == h) {
      grp = m;
      leader = l;
    }
  }
  int start = 0;
  if (valid && lane == leader) start = atomicAdd(&bcount[hb], __popcll(grp));
  start = __shfl(start, leader, TFF_WAVE);
  if (valid) {
    const int r = start + __popcll(grp & ((1ull << lane) - 1ull));
    if (r < bg.cap)
      blist[(long long)hb * bg.cap + r] = make_int2(sl, packed);
    else
      atomicOr(status, TFF_FUSED_BRICK_OVERFLOW);
  }
}

// Charge spread from the home-brick lists: one workgroup per BR^3 brick, in rounds of up to 256
// listed slots.  Each thread tests one slot's stencil against the brick (one run of grid points per
// axis) and compacts into LDS the slot's weights over its in-brick runs (q M_x, M_y, M_z, zero past
// the run) and the runs' starts and lengths.  Then lane k takes x point k % P of compacted slot
// k / P and adds its y-z box of points ((q M_x M_y) M_z, the products k_spread_bricks forms) to the
// brick in LDS with int64 fixed-point atomics.  Integer sums do not depend on the order, so the grid
// equals k_spread_bricks' bit for bit; the brick is stored once, so the grid needs neither clearing
// nor global atomics.
constexpr int LIST_ROUND = 256;
template <int P>
__global__ __launch_bounds__(LIST_ROUND) void k_spread_lists(const float* __restrict__ wts,
                                                             const int* __restrict__ bcount,
                                                             const int2* __restrict__ blist, BrickGrid bg,
                                                             PmeGeom pg, unsigned long long* __restrict__ grid) {
  constexpr int NT = LIST_ROUND;
  constexpr int REC = spline_stride<P>();  // per compacted slot: the weights of its runs, then the runs
  __shared__ unsigned long long s_q[SQ_Z * BR];
  __shared__ __align__(16) float s_w[NT][REC];
  __shared__ int s_hoff[28], s_hid[27], s_cnt[27];
  __shared__ int s_na, s_nh;
  const int K0 = pg.K[0], K1 = pg.K[1], K2 = pg.K[2];
  const int bid = blockIdx.x;
  if (bid >= bg.nb[0] * bg.nb[1] * bg.nb[2]) return;
  const int bi[3] = {bid / (bg.nb[2] * bg.nb[1]), (bid / bg.nb[2]) % bg.nb[1], bid % bg.nb[2]};
  const int Ks[3] = {K0, K1, K2};
  int B0[3], Bn[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    B0[a] = bi[a] * BR;
    Bn[a] = Ks[a] - B0[a] < BR ? Ks[a] - B0[a] : BR;
  }
  for (int k = threadIdx.x; k < SQ_Z * BR; k += NT) s_q[k] = 0ull;
  // home bricks whose slots' stencils can reach this brick: per axis the distinct h among b, b - 1,
  // b - 2 (mod the brick count) whose base range plus stencil meets the brick; lane t < 27 of the
  // first wave takes combination t and loads its list length (the other waves skip this scalar
  // work: the CU's scalar unit is shared)
  const int t = threadIdx.x;
  if (t < TFF_WAVE) {
    int hl[3][3], hn[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a) {
      if (Ks[a] % BR == 0 && bg.nb[a] >= 2) {
        // full bricks: the brick itself and the one below (P <= BR + 1)
        hl[a][0] = bi[a];
        hl[a][1] = bi[a] > 0 ? bi[a] - 1 : bg.nb[a] - 1;
        hn[a] = 2;
        continue;
      }
      for (int d = 0; d < 3; ++d) {
        int h = bi[a] - d;
        while (h < 0) h += bg.nb[a];
        bool dup = false;
        for (int e = 0; e < hn[a]; ++e) dup = dup || hl[a][e] == h;
        const int hB = Ks[a] - h * BR < BR ? Ks[a] - h * BR : BR;
        if (!dup && circ_overlap(h * BR, hB + P - 1, B0[a], Bn[a], Ks[a])) hl[a][hn[a]++] = h;
      }
    }
    const int nh = hn[0] * hn[1] * hn[2];
    if (t < nh) {
      const int x = t / (hn[1] * hn[2]), y = (t / hn[2]) % hn[1], z = t % hn[2];
      const int h = (hl[0][x] * bg.nb[1] + hl[1][y]) * bg.nb[2] + hl[2][z];
      const int c = bcount[h];
      s_hid[t] = h;
      s_cnt[t] = c < bg.cap ? c : bg.cap;
    }
    if (t == 0) s_nh = nh;
  }
  {
    __syncthreads();
    const int nh = s_nh;
    if (t == 0) {
      int run = 0;
      for (int k = 0; k < nh; ++k) {
        s_hoff[k] = run;
        run += s_cnt[k];
      }
      s_hoff[nh] = run;
    }
    __syncthreads();
  }
  const int nh = s_nh, total = s_hoff[nh];
  // candidates in passes of 2 NT: each thread loads and tests its two (k and k + NT) and loads the
  // spline tables of those that hit before the first is spread, so the global loads of both rounds
  // overlap
  for (int p0 = 0; p0 < total; p0 += 2 * NT) {
    float rec2[2][REC];
    bool hit2[2] = {false, false};
    int2 e2[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int k = p0 + c * NT + threadIdx.x;
      e2[c] = make_int2(0, 0);
      if (k < total) {
        int hi = 0;
        for (int j = 1; j < nh; ++j) hi += s_hoff[j] <= k ? 1 : 0;
        e2[c] = blist[(long long)s_hid[hi] * bg.cap + (k - s_hoff[hi])];
      }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int k = p0 + c * NT + threadIdx.x;
      if (k >= total) continue;
      const int2 e = e2[c];
      const int bs[3] = {e.y & 1023, (e.y >> 10) & 1023, (e.y >> 20) & 1023};
      bool hit = true;
      int q0[3], l0[3], nr[3];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        // the stencil [bs, bs + P) and the brick [B0, B0 + Bn) as arcs of Z_K: one common run
        int d = B0[a] - bs[a];
        d = d < 0 ? d + Ks[a] : d;  // brick start relative to the stencil start
        if (d < P) {
          q0[a] = d;
          l0[a] = 0;
          nr[a] = P - d < Bn[a] ? P - d : Bn[a];
        } else {
          int f = bs[a] - B0[a];
          f = f < 0 ? f + Ks[a] : f;  // stencil start relative to the brick start
          q0[a] = 0;
          l0[a] = f;
          nr[a] = f < Bn[a] ? (P < Bn[a] - f ? P : Bn[a] - f) : 0;
        }
        hit = hit && nr[a] > 0;
      }
      if (!hit) continue;
      hit2[c] = true;
      const float* src = wts + (size_t)e.x * LIST_WST;
      float tb[3 * P + 1];
#pragma unroll
      for (int i = 0; i <= 3 * P; ++i) tb[i] = src[i];
#pragma unroll
      for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int i = 0; i < P; ++i) {
          float v = 0.f;
#pragma unroll
          for (int kk = 0; kk < P; ++kk)
            if (i < nr[a] && q0[a] + i == kk) v = tb[a * P + kk];
          // z weights carry the fixed-point scale (a power of two: (a b) 2^28 = a (b 2^28) exactly)
          rec2[c][a * P + i] = a == 0 ? tb[3 * P] * v : (a == 2 ? v * (float)QF_SCALE : v);
        }
      rec2[c][3 * P] = __int_as_float(l0[0] | (l0[1] << 3) | (l0[2] << 6) | (nr[0] << 9) | (nr[1] << 13) |
                                      (nr[2] << 17));
#pragma unroll
      for (int i = 3 * P + 1; i < REC; ++i) rec2[c][i] = 0.f;
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (p0 + c * NT >= total) break;
      if (threadIdx.x == 0) s_na = 0;
      __syncthreads();
      // 1. compact
      if (hit2[c]) {
        const int idx = atomicAdd(&s_na, 1);
        float4* dst = reinterpret_cast<float4*>(s_w[idx]);
#pragma unroll
        for (int i = 0; i < REC / 4; ++i)
          dst[i] = make_float4(rec2[c][4 * i], rec2[c][4 * i + 1], rec2[c][4 * i + 2], rec2[c][4 * i + 3]);
      }
      __syncthreads();
      // 2. lane k: x point k % P of compacted slot k / P, its y-z box
      const int nlanes = s_na * P;
      for (int k = threadIdx.x; k < nlanes; k += NT) {
        const int t = k / P, ix = k - t * P;
        const float* w = s_w[t];
        const int meta = __float_as_int(w[3 * P]);
        const int nx = (meta >> 9) & 15, ny = (meta >> 13) & 15, nz = (meta >> 17) & 15;
        if (ix >= nx) continue;
        const float wx = w[ix];
        float wy[P], wz[P];
#pragma unroll
        for (int i = 0; i < P; ++i) {
          wy[i] = w[P + i];
          wz[i] = w[2 * P + i];
        }
        const int lx = (meta & 7) + ix, ly0 = (meta >> 3) & 7, lz0 = (meta >> 6) & 7;
#pragma unroll
        for (int iy = 0; iy < P; ++iy) {
          if (iy >= ny) break;
          const float wab = wx * wy[iy];
          long long v[P];
#pragma unroll
          for (int iz = 0; iz < P; ++iz) v[iz] = (long long)__float2int_rn(wab * wz[iz]);
          unsigned long long* row = s_q + lz0 * SQ_Z + (ly0 + iy) * SQ_Y + lx;
#pragma unroll
          for (int iz = 0; iz < P; ++iz)
            if (iz < nz) atomicAdd(&row[iz * SQ_Z], (unsigned long long)v[iz]);
        }
      }
      __syncthreads();
    }
  }
  for (int k = threadIdx.x; k < BR * BR * BR; k += NT) {
    const int lz = k % BR, ly = (k / BR) % BR, lx = k / (BR * BR);
    if (lx < Bn[0] && ly < Bn[1] && lz < Bn[2])
      grid[((long long)(B0[0] + lx) * K1 + (B0[1] + ly)) * K2 + (B0[2] + lz)] = s_q[lz * SQ_Z + ly * SQ_Y + lx];
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
  // force = -dE/dr into the slot's reciprocal-space accumulator (one thread per slot; the pair
  // kernel, which may run concurrently, adds into the other accumulator)
  facc[3 * sl + 0] = (unsigned long long)to_fx(-kq * gxr, F_SCALE);
  facc[3 * sl + 1] = (unsigned long long)to_fx(-kq * gyr, F_SCALE);
  facc[3 * sl + 2] = (unsigned long long)to_fx(-kq * gzr, F_SCALE);
}

// Backward gather, one thread per slot: from the potential grid Phi (of the charges) and Psi (of
// the dipoles q v . grad theta), the reciprocal-space parts
//   dL/dq_i = k [gE phi(r_i) - v_i . grad phi(r_i) - psi(r_i)],
//   dL/dr_i = k q_i [gE grad phi(r_i) - H_phi(r_i) v_i - grad psi(r_i)]
// (orthorhombic box: d s_a / d r_a = K_a / L_a), as int64 fixed point into gpme (4 per slot).
template <typename TC, int P>
__global__ __launch_bounds__(FB) void k_gather_grad(const int* __restrict__ ncl_ptr, const TC* __restrict__ phi_grid,
                                                    const TC* __restrict__ psi_grid, const V4<TC>* __restrict__ xyzq,
                                                    const V4<TC>* __restrict__ vsl, const int* __restrict__ slot_orig,
                                                    PmeGeom g, double coul, const double* __restrict__ gE_ptr,
                                                    unsigned long long* __restrict__ gpme) {
  const int sl = blockIdx.x * FB + threadIdx.x;
  if (sl >= *ncl_ptr * CL || slot_orig[sl] < 0) return;
  const V4<TC> p = xyzq[sl];
  const V4<TC> v = vsl[sl];
  TC M[3][P], D1[3][P], D2[3][P];
  int base[3];
  const double r[3] = {(double)p.x, (double)p.y, (double)p.z};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double sa = (r[0] * g.inv[a] + r[1] * g.inv[3 + a] + r[2] * g.inv[6 + a]) * (double)g.K[a];
    double t;
    anchor_axis(sa, g.K[a], P, base[a], t);
    spline_weights2<TC, P>((TC)t, M[a], D1[a], D2[a]);
  }
  const int K1 = g.K[1], K2 = g.K[2];
  // separable sums over z, then y, then x (derivatives in t; d/ds = -d/dt)
  double phi = 0, px = 0, py = 0, pz = 0, hxx = 0, hyy = 0, hzz = 0, hxy = 0, hxz = 0, hyz = 0;
  double psi = 0, sx = 0, sy = 0, sz = 0;
  int gx = base[0];
#pragma unroll 1
  for (int a = 0; a < P; ++a) {
    double ty0 = 0, ty1 = 0, ty2 = 0, tz1 = 0, tz2 = 0, tyz = 0, uy0 = 0, uy1 = 0, uz1 = 0;
    int gy = base[1];
#pragma unroll
    for (int b = 0; b < P; ++b) {
      const long long row = ((long long)gx * K1 + gy) * K2;
      double z0 = 0, z1 = 0, z2 = 0, w0 = 0, w1 = 0;
      int gz = base[2];
#pragma unroll
      for (int c = 0; c < P; ++c) {
        const double f = (double)phi_grid[row + gz], q = (double)psi_grid[row + gz];
        z0 += (double)M[2][c] * f;
        z1 += (double)D1[2][c] * f;
        z2 += (double)D2[2][c] * f;
        w0 += (double)M[2][c] * q;
        w1 += (double)D1[2][c] * q;
        gz = gz + 1 == K2 ? 0 : gz + 1;
      }
      const double my = M[1][b], dy = D1[1][b], ddy = D2[1][b];
      ty0 += my * z0;
      ty1 += dy * z0;
      ty2 += ddy * z0;
      tz1 += my * z1;
      tz2 += my * z2;
      tyz += dy * z1;
      uy0 += my * w0;
      uy1 += dy * w0;
      uz1 += my * w1;
      gy = gy + 1 == K1 ? 0 : gy + 1;
    }
    const double mx = M[0][a], dx = D1[0][a], ddx = D2[0][a];
    phi += mx * ty0;
    px += dx * ty0;
    py += mx * ty1;
    pz += mx * tz1;
    hxx += ddx * ty0;
    hyy += mx * ty2;
    hzz += mx * tz2;
    hxy += dx * ty1;
    hxz += dx * tz1;
    hyz += mx * tyz;
    psi += mx * uy0;
    sx += dx * uy0;
    sy += mx * uy1;
    sz += mx * uz1;
    gx = gx + 1 == g.K[0] ? 0 : gx + 1;
  }
  const double j0 = g.K[0] * g.inv[0], j1 = g.K[1] * g.inv[4], j2 = g.K[2] * g.inv[8];
  // Cartesian gradient (first derivatives change sign from t to s) and Hessian
  const double gp[3] = {-px * j0, -py * j1, -pz * j2};
  const double gs[3] = {-sx * j0, -sy * j1, -sz * j2};
  const double H[3][3] = {{hxx * j0 * j0, hxy * j0 * j1, hxz * j0 * j2},
                          {hxy * j0 * j1, hyy * j1 * j1, hyz * j1 * j2},
                          {hxz * j0 * j2, hyz * j1 * j2, hzz * j2 * j2}};
  const double gE = gE_ptr ? gE_ptr[0] : 0.0;
  const double vv[3] = {(double)v.x, (double)v.y, (double)v.z};
  const double q = (double)p.w;
  const double dq = coul * (gE * phi - (vv[0] * gp[0] + vv[1] * gp[1] + vv[2] * gp[2]) - psi);
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double hv = H[a][0] * vv[0] + H[a][1] * vv[1] + H[a][2] * vv[2];
    gpme[4 * (long long)sl + a] = (unsigned long long)to_fx(coul * q * (gE * gp[a] - hv - gs[a]), F_SCALE);
  }
  gpme[4 * (long long)sl + 3] = (unsigned long long)to_fx(dq, F_SCALE);
}

// gradients back in the caller's atom order (pair + reciprocal accumulators; self energy
// -k alpha / sqrt(pi) sum q^2 contributes gE * (-2 k alpha / sqrt(pi)) q_i to dL/dq_i)
template <typename TI>
__global__ __launch_bounds__(FB) void k_finalize_grad(const int* __restrict__ ncl_ptr, const int* __restrict__ slot_orig,
                                                      const unsigned long long* __restrict__ gacc,
                                                      const unsigned long long* __restrict__ gpme,
                                                      const TI* __restrict__ q, double coul, double alpha,
                                                      const double* __restrict__ gE_ptr, TI* __restrict__ dr,
                                                      TI* __restrict__ dq, TI* __restrict__ dsig, TI* __restrict__ deps,
                                                      const int* __restrict__ status, int* __restrict__ status_host) {
  const int sl = blockIdx.x * FB + threadIdx.x;
  if (status_host && sl == 0) {
    // every kernel that sets a flag ran before this one on the stream; word 0 last, after a
    // system-scope fence: a host that sees it sees word 1
    status_host[1] = status[1];
    __threadfence_system();
    status_host[0] = status[0];
  }
  if (sl >= *ncl_ptr * CL) return;
  const int i = slot_orig[sl];
  if (i < 0) return;
  const long long b6 = (long long)GN * sl, b4 = 4 * (long long)sl;
  if (dr)
    for (int a = 0; a < 3; ++a) dr[3 * i + a] = (TI)((double)(long long)(gacc[b6 + a] + gpme[b4 + a]) / F_SCALE);
  if (dq) {
    const double gE = gE_ptr ? gE_ptr[0] : 0.0;
    dq[i] = (TI)((double)(long long)(gacc[b6 + 3] + gpme[b4 + 3]) / F_SCALE -
                 gE * 2.0 * coul * alpha * 0.5641895835477563 * (double)q[i]);
  }
  if (dsig) dsig[i] = (TI)((double)(long long)gacc[b6 + 4] / F_SCALE);
  if (deps) deps[i] = (TI)((double)(long long)gacc[b6 + 5] / F_SCALE);
}

// ---------------------------------------------------------------- 6. finalize
// Forces (pair + reciprocal accumulators) back in the caller's atom order; block 0 also sums
// the energy terms:
// out = [Lennard-Jones, real-space Coulomb, reciprocal, self, total].
template <typename TI>
__global__ __launch_bounds__(FB) void k_finalize(const int* __restrict__ ncl_ptr, const int* __restrict__ slot_orig,
                                                 const unsigned long long* __restrict__ facc,
                                                 const unsigned long long* __restrict__ facc_pme, TI* __restrict__ forces,
                                                 const unsigned long long* __restrict__ eacc,
                                                 const double* __restrict__ kparts, int nk,
                                                 const double* __restrict__ qparts, int nq, double coul,
                                                 double alpha, double* __restrict__ out, const int* __restrict__ status,
                                                 int* __restrict__ status_out, int* __restrict__ status_host) {
  const int sl = blockIdx.x * FB + threadIdx.x;
  if (sl == 0) {
    // every kernel that sets a flag ran before this one on the stream (after the join)
    const int s0 = status[0], s1 = status[1];
    status_out[0] = s0;
    status_out[1] = s1;
    if (status_host) {
      // word 0 last, after a system-scope fence: a host that sees it sees word 1
      status_host[1] = s1;
      __threadfence_system();
      status_host[0] = s0;
    }
  }
  if (sl < *ncl_ptr * CL) {
    const int i = slot_orig[sl];
    if (i >= 0)
      for (int a = 0; a < 3; ++a)
        forces[3 * i + a] = (TI)((double)(long long)(facc[3 * sl + a] + facc_pme[3 * sl + a]) / F_SCALE);
  }
  if (blockIdx.x != 0) return;
  const double ek = tff_sum_partials(kparts, nk);
  const double q2 = tff_sum_partials(qparts, nq);
  if (threadIdx.x != 0) return;
  long long sl_lj = 0, sl_c = 0;
  for (int k = 0; k < E_SLOTS; ++k) {
    sl_lj += (long long)eacc[k];
    sl_c += (long long)eacc[E_SLOTS + k];
  }
  out[0] = (double)sl_lj / F_SCALE;
  out[1] = (double)sl_c / F_SCALE;
  out[2] = coul * ek;
  out[3] = -coul * alpha * 0.5641895835477563 * q2;
  out[4] = out[0] + out[1] + out[2] + out[3];
}



// ---------------------------------------------------------------- host side

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

// waves of the pair kernel: about 16 tile-list slots each (some 6 listed tiles at water density)
int tile_waves(const Geo& g) {
  long long w = (g.pair_cap + 15) / 16;
  if (w < 256) w = 256;
  if (w > (1 << 20)) w = 1 << 20;
  return (int)w;
}

// z extent of the reciprocal-space array: even K2 uses the real-input half spectrum
int spectrum_nz(const int K[3]) { return K[2] % 2 == 0 ? K[2] / 2 + 1 : K[2]; }

struct FusedWs {
  int *col_of, *off_in_col, *col_count, *col_start, *sorted_orig, *col_cl_start, *ncl, *cl_col, *slot_orig,
      *excl_part, *ncp, *seg_count, *seg_start, *slot_of_orig;
  int* st;  // the forward call's status words while it runs (copied out by k_finalize)
  int2* seg;
  void *xyzq, *lj;
  float4 *bb_c, *bb_h, *sb_c, *sb_h;
  int4* cpairs;
  unsigned long long *facc, *facc_pme, *grid;
  void* work;
  double *kparts, *qparts;
  unsigned long long* eacc;
  // float compute only: per-slot spline tables and home-brick lists of the charge spread
  float* wts;
  int* bcount;
  int2* blist;
  BrickGrid bg;
  // backward only
  void* vsl;
  unsigned long long *gacc, *gpme, *grid_d;
  size_t bytes;
  int nk, nw_tiles, nq;
};

FusedWs carve(void* base, long long n, const Geo& g, const int K[3], size_t tc, bool grad = false) {
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
  w.col_of = (int*)take(4 * nn);
  w.off_in_col = (int*)take(4 * nn);
  w.col_count = (int*)take(4 * (size_t)(ncols + 1));
  w.col_start = (int*)take(4 * (size_t)(ncols + 1));
  w.sorted_orig = (int*)take(4 * nn);
  w.col_cl_start = (int*)take(4 * (size_t)(ncols + 1));
  w.ncl = (int*)take(16);
  w.ncp = (int*)take(16);
  w.st = (int*)take(16);
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
  w.cpairs = (int4*)take(16 * (size_t)g.pair_cap);
  w.seg = (int2*)take(8 * (size_t)g.pair_cap);
  w.seg_count = (int*)take(4 * (size_t)g.ncl_max);
  w.seg_start = (int*)take(4 * (size_t)g.ncl_max);
  w.facc = (unsigned long long*)take(8 * 3 * slots);
  w.facc_pme = (unsigned long long*)take(8 * 3 * slots);
  w.nw_tiles = tile_waves(g);
  w.eacc = (unsigned long long*)take(8 * 2 * E_SLOTS);
  w.grid = (unsigned long long*)take(8 * ng);
  w.work = take(2 * tc * ng);
  PassCfg px = pass_cfg(1, K[0], K[1] * spectrum_nz(K), 2 * tc);
  w.nk = px.tiles;
  w.kparts = (double*)take(8 * (size_t)w.nk);
  w.nq = (int)((nn + FB - 1) / FB);
  w.qparts = (double*)take(8 * (size_t)w.nq);
  if (tc == 4) {
    for (int a = 0; a < 3; ++a) w.bg.nb[a] = (K[a] + BR - 1) / BR;
    const long long nbr = (long long)w.bg.nb[0] * w.bg.nb[1] * w.bg.nb[2];
    // four times the mean occupancy (plus a margin) per brick; more sets TFF_FUSED_BRICK_OVERFLOW
    w.bg.cap = (int)(4 * ((long long)nn + nbr - 1) / nbr + 64);
    w.wts = (float*)take(4 * (size_t)LIST_WST * slots);
    w.bcount = (int*)take(4 * (size_t)nbr);
    w.blist = (int2*)take(8 * (size_t)nbr * w.bg.cap);
  }
  if (grad) {
    w.vsl = take(4 * tc * slots);
    w.gacc = (unsigned long long*)take(8 * GN * slots);
    w.gpme = (unsigned long long*)take(8 * 4 * slots);
    w.grid_d = (unsigned long long*)take(8 * ng);
  }
  w.bytes = off;
  return w;
}

// the list spread serves the float path for orders up to 6 on grids with BR + P <= K_a <= 1023
bool spread_from_lists(int P, const int K[3]) {
  if (P > 6) return false;
  for (int a = 0; a < 3; ++a)
    if (K[a] < BR + P || K[a] > 1023) return false;
  return true;
}

template <typename TC>
int launch_pme_slots(int order, const FusedWs& w, const Geo& g, const PmeGeom& pg, int nslots, const TC* phi,
                     double coul, bool spread, int* status, hipStream_t st) {
  unsigned nb = (unsigned)((nslots + FB - 1) / FB);
  const unsigned nbricks = (unsigned)(((pg.K[0] + BR - 1) / BR) * ((pg.K[1] + BR - 1) / BR) * ((pg.K[2] + BR - 1) / BR));
#define TFF_SP(P)                                                                                               \
  if (spread && sizeof(TC) == 4 && spread_from_lists(P, pg.K)) {                                                \
    hipLaunchKernelGGL((k_slot_splines<P < 7 ? P : 6>), dim3(nb), dim3(FB), 0, st, w.ncl,                      \
                       (const V4<float>*)w.xyzq, w.slot_orig, pg, w.bg, w.wts, w.bcount, w.blist, status);      \
    hipLaunchKernelGGL((k_spread_lists<P < 7 ? P : 6>), dim3(nbricks), dim3(LIST_ROUND), 0, st,              \
                       (const float*)w.wts,                                                                      \
                       (const int*)w.bcount, (const int2*)w.blist, w.bg, pg, w.grid);                         \
  } else if (spread)                                                                                            \
    hipLaunchKernelGGL((k_spread_bricks<TC, P, false>), dim3(nbricks), dim3(256), 0, st, w.ncl,               \
                       (const V4<TC>*)w.xyzq, w.slot_orig, w.bb_c, w.bb_h, w.col_cl_start, g, pg, w.grid, status, \
                       (const V4<TC>*)nullptr, V4<TC>{});                                                        \
  else                                                                                                          \
    hipLaunchKernelGGL((k_gather_slots<TC, P>), dim3(nb), dim3(FB), 0, st, w.ncl, phi, (const V4<TC>*)w.xyzq,   \
                       w.slot_orig, pg, coul, w.facc_pme)
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

// Per-device auxiliary stream (default priority) and fork / join events.  The fork makes the
// auxiliary stream wait for the work enqueued so far on the caller's stream; the join makes
// the caller's stream wait for the auxiliary work, so the call stays ordered on the caller's
// stream (and can be captured into a graph).  The mutex covers each record / wait pair.
struct AuxStream {
  hipStream_t s = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
};
std::mutex g_aux_mu;
AuxStream g_aux[64];

// 0: the reciprocal-space chain runs on the caller's stream, in order (a timing aid); 2: the charge
// spread runs on the caller's stream before the fork, the FFT and the gather on the auxiliary
// stream; 3: the whole chain runs on the auxiliary stream; 1: 3 below SPREAD_FIRST_ATOMS, 2 from
// there on.  Measured per evaluation (mixed): 96k atoms 0.309 ms as 3 against 0.312 as 2; 999k
// atoms 2.616 ms as 3 against 2.533 as 2 (the pair kernel starves a concurrent spread, which
// then leaves the FFT and the gather to run after it).
constexpr long long SPREAD_FIRST_ATOMS = 300000;
std::atomic<int> g_overlap{1};

// *sp: the auxiliary stream, waiting for the work enqueued so far on st (st itself when the
// overlap is off); false when the stream or its events could not be set up
// creates the auxiliary stream and its events of device dev once (caller holds g_aux_mu)
bool aux_stream_setup_locked(int dev) {
  AuxStream& a = g_aux[dev];
  if (!a.s) {
    if (hipStreamCreateWithFlags(&a.s, hipStreamNonBlocking) != hipSuccess) return false;
    if (hipEventCreateWithFlags(&a.fork, hipEventDisableTiming) != hipSuccess) return false;
    if (hipEventCreateWithFlags(&a.join, hipEventDisableTiming) != hipSuccess) return false;
  }
  return true;
}

bool aux_stream_fork(hipStream_t st, hipStream_t* sp) {
  *sp = st;
  if (!g_overlap.load()) return true;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return false;
  std::lock_guard<std::mutex> lk(g_aux_mu);
  AuxStream& a = g_aux[dev];
  if (!aux_stream_setup_locked(dev)) return false;
  if (hipEventRecord(a.fork, st) != hipSuccess) return false;
  if (hipStreamWaitEvent(a.s, a.fork, 0) != hipSuccess) return false;
  *sp = a.s;
  return true;
}

void aux_stream_join(hipStream_t st, hipStream_t sp) {
  if (sp == st) return;
  int dev = 0;
  hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_aux_mu);
  AuxStream& a = g_aux[dev];
  hipEventRecord(a.join, a.s);
  hipStreamWaitEvent(st, a.join, 0);
}

// Reciprocal-space convolution of a fixed-point grid: forward transforms, influence function with
// the k-space energy partials, inverse transforms; the real potential grid phi_out may alias
// grid_in (it is written by the last pass only).
template <typename TC>
void run_recip_fft(const FusedWs& w, const int K[3], const TffBox& box, double alpha, const FftPlan* pl,
                   const TC* mx, const TC* my, const TC* mz, const unsigned long long* grid_in,
                   double grid_inv_scale, TC* phi_out, double* kparts, const char* prof_name, hipStream_t st) {
  ConvArgs cv{};
  for (int a = 0; a < 3; ++a) cv.K[a] = K[a];
  for (int k = 0; k < 9; ++k) cv.inv[k] = box.inv[k];
  cv.vol = box.det;
  cv.alpha = alpha;
  cv.grid_fixed = 1;
  cv.grid_inv_scale = grid_inv_scale;
  const size_t ce = 2 * sizeof(TC);
  // real-input transforms along z when K2 is even (half spectrum, K2 / 2 + 1 planes), else complex
  const bool rz = K[2] % 2 == 0;
  const int nzc = spectrum_nz(K);
  cv.nz = nzc;
  cv.half = rz ? 1 : 0;
  PassCfg cz = pass_cfg(K[0] * K[1], K[2], 1, ce);
  PassCfg cy = pass_cfg(K[0], K[1], nzc, ce);
  PassCfg cxp = pass_cfg(1, K[0], K[1] * nzc, ce);
  FftPlan pzh;
  make_plan(rz ? K[2] / 2 : 1, pzh);
  const ZPlanCfg zc = zcfg(K[0] * K[1], K[2], ce);
  {
    void* tok = tff_prof_begin(prof_name, st);
    if (rz)
      hipLaunchKernelGGL(k_fft_z_r2c<TC>, dim3(zc.blocks), dim3(FFT_THREADS), zc.lds, st, grid_in, (cx<TC>*)w.work,
                         K[0] * K[1], K[2], zc.AB, zc.ab_shift, pzh, 1, cv.grid_inv_scale);
    else
      hipLaunchKernelGGL((k_fft_pass<TC, PASS_Z_FWD>), dim3(cz.tiles), dim3(FFT_THREADS), cz.lds, st, grid_in, w.work,
                         cz.pa, pl[2], cv, mx, my, mz, 1, kparts);
    hipLaunchKernelGGL((k_fft_pass<TC, PASS_Y_FWD>), dim3(cy.tiles), dim3(FFT_THREADS), cy.lds, st, w.work, w.work,
                       cy.pa, pl[1], cv, mx, my, mz, 1, kparts);
    hipLaunchKernelGGL((k_fft_pass<TC, PASS_X_CONV>), dim3(cxp.tiles), dim3(FFT_THREADS), cxp.lds, st, w.work,
                       w.work, cxp.pa, pl[0], cv, mx, my, mz, 1, kparts);
    hipLaunchKernelGGL((k_fft_pass<TC, PASS_Y_INV>), dim3(cy.tiles), dim3(FFT_THREADS), cy.lds, st, w.work, w.work,
                       cy.pa, pl[1], cv, mx, my, mz, 1, kparts);
    if (rz)
      hipLaunchKernelGGL(k_fft_z_c2r<TC>, dim3(zc.blocks), dim3(FFT_THREADS), zc.lds, st, (const cx<TC>*)w.work, phi_out,
                         K[0] * K[1], K[2], zc.AB, zc.ab_shift, pzh);
    else
      hipLaunchKernelGGL((k_fft_pass<TC, PASS_Z_INV>), dim3(cz.tiles), dim3(FFT_THREADS), cz.lds, st, w.work, phi_out,
                         cz.pa, pl[2], cv, mx, my, mz, 1, kparts);
    tff_prof_end(tok, st);
  }
}

// Everything of one forward call up to the join of the reciprocal-space chain: reads the inputs,
// writes only the workspace (status words in w.st), so the same launch sequence can be captured
// once into a graph and replayed for every call with the same inputs and workspace.
template <typename TI, typename TC>
int run_fused_core(const TI* coords, const TI* q, const TI* sig, const TI* eps, long long n, const TffBox& box,
                   const Geo& g, const int64_t* row_ptr, const int64_t* col, double cutoff, double list_cutoff,
                   double alpha, double coul, const int K[3], int order, const TC* mx, const TC* my, const TC* mz,
                   const TI* box_dev, const FusedWs& w, hipStream_t st) {
  if (order < 3 || order > 8) {  // checked before any launch: the stream fork below must be joined
    tff_set_error("fused: unsupported spline order %d (3..8)", order);
    return TFF_ERR_UNSUPPORTED;
  }
  int* status = w.st;
  const unsigned nbA = (unsigned)((n + FB - 1) / FB);
  const int nslots = g.ncl_max * CL;
  const unsigned nbS = (unsigned)((nslots + FB - 1) / FB);
  const int ncols = g.ncol[0] * g.ncol[1];
  {
    ClearList cl{};
    auto add = [&](void* p, long long words) {
      cl.p[cl.count] = (int*)p;
      cl.n[cl.count] = words;
      ++cl.count;
    };
    add(w.col_count, ncols + 1);
    add(w.ncp, 4);
    add(status, 2);
    add(w.facc, 6LL * nslots);
    add(w.slot_orig, nslots);
    add(w.eacc, 4 * E_SLOTS);
    if (w.bcount) add(w.bcount, (long long)w.bg.nb[0] * w.bg.nb[1] * w.bg.nb[2]);
    hipLaunchKernelGGL(k_clear, dim3(1024), dim3(256), 0, st, cl);
  }
  // 1. sort
  hipLaunchKernelGGL(k_bin_atoms<TI>, dim3(nbA), dim3(FB), 0, st, coords, q, n, g, w.col_of, w.off_in_col,
                     w.col_count, w.qparts, box_dev, status);
  hipLaunchKernelGGL(k_scan_columns, dim3(1), dim3(1024), 0, st, w.col_count, ncols, w.col_start, w.col_cl_start,
                     w.ncl, status);
  hipLaunchKernelGGL(k_scatter_atoms, dim3(nbA), dim3(FB), 0, st, n, w.col_of, w.off_in_col, w.col_start,
                     w.sorted_orig);
  hipLaunchKernelGGL(k_sort_columns<TI>, dim3(ncols), dim3(SORT_THREADS), 0, st, coords, g, w.col_start, w.sorted_orig);
  // 2. cluster arrays
  const unsigned nbW = (unsigned)(((long long)g.ncl_max * TFF_WAVE + FB - 1) / FB);
  const unsigned nbB = (unsigned)(((long long)g.ncl_max * CL + FB - 1) / FB);  // two clusters per wave
  hipLaunchKernelGGL((k_build_clusters<TI, TC>), dim3(nbB), dim3(FB), 0, st, coords, q, sig, eps, g, w.ncl,
                     w.col_cl_start, w.col_start, w.sorted_orig, (V4<TC>*)w.xyzq, (V2<TC>*)w.lj, w.slot_orig,
                     w.slot_of_orig, w.cl_col, w.bb_c, w.bb_h, w.sb_c, w.sb_h);
  // 3. PME on the auxiliary stream (needs only the cluster arrays), concurrent with 4-5; the
  // exclusion table (needed from the pair search on) beside the first reciprocal-space kernels
  PmeGeom pg = make_geom(box, K, order, alpha);
  FftPlan pl[3];
  for (int a = 0; a < 3; ++a) make_plan(K[a], pl[a]);
  const int mode = g_overlap.load();
  const bool spread_first = mode == 2 || (mode == 1 && n >= SPREAD_FIRST_ATOMS);
  auto fill_exclusions = [&](hipStream_t s) {
    hipLaunchKernelGGL(k_fill_exclusions, dim3(nbS), dim3(FB), 0, s, w.ncl, w.slot_orig, w.slot_of_orig, n, row_ptr,
                       col, w.excl_part, status);
  };
  auto spread = [&](hipStream_t s) {
    void* tok = tff_prof_begin("fused_spread", s);
    const int rc = launch_pme_slots<TC>(order, w, g, pg, nslots, nullptr, coul, true, status, s);
    tff_prof_end(tok, s);
    return rc;
  };
  hipStream_t sp = st;
  if (!aux_stream_fork(st, &sp)) {
    tff_set_error("fused: could not create the auxiliary stream");
    return TFF_ERR_LAUNCH;
  }
  if (spread_first) {
    // the spread alone on the caller's stream, the exclusion table beside it; the FFT and the
    // gather fork after the spread
    fill_exclusions(sp);
    const int rc = spread(st);
    if (rc) return rc;
    aux_stream_join(st, sp);
    if (!aux_stream_fork(st, &sp)) {
      tff_set_error("fused: could not create the auxiliary stream");
      return TFF_ERR_LAUNCH;
    }
  } else {
    const int rc = spread(sp);
    if (rc) return rc;
  }
  TC* phi = (TC*)w.grid;  // the real potential grid reuses the charge grid's memory after the z pass
  run_recip_fft<TC>(w, K, box, alpha, pl, mx, my, mz, w.grid, 1.0 / grid_scale<TC>(), phi, w.kparts, "fused_fft", sp);
  {
    void* tok = tff_prof_begin("fused_gather", sp);
    int rc = launch_pme_slots<TC>(order, w, g, pg, nslots, phi, coul, false, status, sp);
    tff_prof_end(tok, sp);
    if (rc) return rc;
  }
  if (!spread_first) fill_exclusions(st);
  // 4. cluster pairs
  {
    void* tok = tff_prof_begin("fused_cluster_pairs", st);
    float rl2 = (float)(list_cutoff * list_cutoff);
    hipLaunchKernelGGL(k_cluster_pairs, dim3(nbW), dim3(FB), 0, st, w.ncl, w.cl_col, w.col_cl_start, w.bb_c,
                       w.bb_h, g, rl2, w.excl_part, w.seg, w.seg_count, status);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, w.ncl, w.seg_count, w.seg_start, w.ncp, status);
    // sub-tiles are pruned at the interaction cutoff itself: pairs between rc and the list
    // cutoff contribute nothing to this evaluation
    hipLaunchKernelGGL(k_compact_pairs, dim3(nbW), dim3(FB), 0, st, w.ncl, w.seg, w.seg_count, w.seg_start, g.maxn,
                       w.sb_c, w.sb_h, g, (float)(cutoff * cutoff), w.cpairs);
    tff_prof_end(tok, st);
  }
  // 5. tiles
  {
    TileArgs<TC> ta;
    for (int a = 0; a < 3; ++a) {
      ta.L[a] = (TC)g.L[a];
      ta.invL[a] = (TC)g.invL[a];
    }
    ta.rc2 = (TC)(cutoff * cutoff);
    ta.alpha = (TC)alpha;
    ta.coul = (TC)coul;
    ta.c2alpha = (TC)(1.1283791670955126 * alpha);
    void* tok = tff_prof_begin("fused_tiles", st);
    if constexpr (sizeof(TC) == 4)
      hipLaunchKernelGGL(k_tiles_f32, dim3((unsigned)w.nw_tiles), dim3(TFF_WAVE), 0, st, w.cpairs, w.ncp,
                         g.pair_cap, (const V4<float>*)w.xyzq, (const V2<float>*)w.lj, w.excl_part, ta, w.facc,
                         w.eacc);
    else
      hipLaunchKernelGGL(k_tiles_f64, dim3((unsigned)w.nw_tiles), dim3(TFF_WAVE), 0, st, w.cpairs, w.ncp,
                         g.pair_cap, (const V4<double>*)w.xyzq, (const V2<double>*)w.lj, w.excl_part, ta, w.facc,
                         w.eacc);
    tff_prof_end(tok, st);
  }
  aux_stream_join(st, sp);
  TFF_CHECK_LAUNCH("fused energy+forces");
  return TFF_OK;
}

// 6. outputs, after the reciprocal-space work: forces in atom order, energy terms, status words
template <typename TI>
int launch_fused_finalize(const FusedWs& w, const Geo& g, double coul, double alpha, double* energies, TI* forces,
                          int* status_out, int* status_host, hipStream_t st) {
  const unsigned nbS = (unsigned)(((long long)g.ncl_max * CL + FB - 1) / FB);
  hipLaunchKernelGGL(k_finalize<TI>, dim3(nbS), dim3(FB), 0, st, w.ncl, w.slot_orig, w.facc, w.facc_pme, forces, w.eacc,
                     w.kparts, w.nk, w.qparts, w.nq, coul, alpha, energies, w.st, status_out, status_host);
  TFF_CHECK_LAUNCH("fused energy+forces");
  return TFF_OK;
}

// Force-matching backward: rebuilds the cluster pair list (not stored from the forward call),
// then the pair backward, the charge and dipole reciprocal convolutions and the gather.
template <typename TI, typename TC>
int run_fused_grad(const TI* coords, const TI* q, const TI* sig, const TI* eps, long long n, const TffBox& box,
                   const Geo& g, const int64_t* row_ptr, const int64_t* col, double cutoff, double list_cutoff,
                   double alpha, double coul, const int K[3], int order, const TC* mx, const TC* my, const TC* mz,
                   const double* gE, const TI* gF, TI* dr, TI* dq, TI* dsig, TI* deps, int* status, const TI* box_dev,
                   int* status_host, const FusedWs& w, bool reuse, hipStream_t st) {
  if (order < 3 || order > 8) {
    tff_set_error("fused: unsupported spline order %d (3..8)", order);
    return TFF_ERR_UNSUPPORTED;
  }
  const unsigned nbA = (unsigned)((n + FB - 1) / FB);
  const int nslots = g.ncl_max * CL;
  const unsigned nbS = (unsigned)((nslots + FB - 1) / FB);
  const int ncols = g.ncol[0] * g.ncol[1];
  {
    ClearList cl{};
    auto add = [&](void* p, long long words) {
      cl.p[cl.count] = (int*)p;
      cl.n[cl.count] = words;
      ++cl.count;
    };
    if (!reuse) {
      add(w.col_count, ncols + 1);
      add(w.ncp, 4);
      add(status, 2);
      add(w.slot_orig, nslots);
      // the brick counters of the charge spread (float: the forward's list spread runs again)
      if (w.bcount) add(w.bcount, (long long)w.bg.nb[0] * w.bg.nb[1] * w.bg.nb[2]);
    }
    add(w.gacc, 2LL * GN * nslots);
    hipLaunchKernelGGL(k_clear, dim3(1024), dim3(256), 0, st, cl);
  }
  const unsigned nbW = (unsigned)(((long long)g.ncl_max * TFF_WAVE + FB - 1) / FB);
  // reuse: the cluster arrays, the exclusion table and the pair list are those the forward call
  // left in the workspace (for the same inputs); otherwise they are rebuilt here
  if (!reuse) {
    hipLaunchKernelGGL(k_bin_atoms<TI>, dim3(nbA), dim3(FB), 0, st, coords, q, n, g, w.col_of, w.off_in_col,
                       w.col_count, w.qparts, box_dev, status);
    hipLaunchKernelGGL(k_scan_columns, dim3(1), dim3(1024), 0, st, w.col_count, ncols, w.col_start,
                       w.col_cl_start, w.ncl, status);
    hipLaunchKernelGGL(k_scatter_atoms, dim3(nbA), dim3(FB), 0, st, n, w.col_of, w.off_in_col, w.col_start,
                       w.sorted_orig);
    hipLaunchKernelGGL(k_sort_columns<TI>, dim3(ncols), dim3(SORT_THREADS), 0, st, coords, g, w.col_start,
                       w.sorted_orig);
    const unsigned nbB = (unsigned)(((long long)g.ncl_max * CL + FB - 1) / FB);  // two clusters per wave
    hipLaunchKernelGGL((k_build_clusters<TI, TC>), dim3(nbB), dim3(FB), 0, st, coords, q, sig, eps, g, w.ncl,
                       w.col_cl_start, w.col_start, w.sorted_orig, (V4<TC>*)w.xyzq, (V2<TC>*)w.lj, w.slot_orig,
                       w.slot_of_orig, w.cl_col, w.bb_c, w.bb_h, w.sb_c, w.sb_h);
  }
  hipLaunchKernelGGL((k_fill_vslots<TI, TC>), dim3(nbS), dim3(FB), 0, st, w.ncl, w.slot_orig, gF, (const V2<TC>*)w.lj,
                     (V4<TC>*)w.vsl, reuse ? (const int*)w.st : nullptr, status);
  // the reciprocal-space part (charge and dipole grids, both transforms, the gather) needs only
  // the cluster arrays and the per-slot upstream forces: it runs on the auxiliary stream beside the
  // pair search and the pair kernel, joined before the final kernel
  hipStream_t sp = st;
  if (!aux_stream_fork(st, &sp)) {
    tff_set_error("fused backward: could not create the auxiliary stream");
    return TFF_ERR_LAUNCH;
  }
  // the pair search and the pair kernel are enqueued first, so that the pair kernel starts as soon
  // as the caller's stream reaches it rather than after the host has enqueued the reciprocal chain
  if (!reuse) {
    hipLaunchKernelGGL(k_fill_exclusions, dim3(nbS), dim3(FB), 0, st, w.ncl, w.slot_orig, w.slot_of_orig, n,
                       row_ptr, col, w.excl_part, status);
    hipLaunchKernelGGL(k_cluster_pairs, dim3(nbW), dim3(FB), 0, st, w.ncl, w.cl_col, w.col_cl_start, w.bb_c,
                       w.bb_h, g, (float)(list_cutoff * list_cutoff), w.excl_part, w.seg, w.seg_count, status);
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, w.ncl, w.seg_count, w.seg_start, w.ncp, status);
    hipLaunchKernelGGL(k_compact_pairs, dim3(nbW), dim3(FB), 0, st, w.ncl, w.seg, w.seg_count, w.seg_start, g.maxn,
                       w.sb_c, w.sb_h, g, (float)(cutoff * cutoff), w.cpairs);
  }
  {
    TileArgs<TC> ta;
    for (int a = 0; a < 3; ++a) {
      ta.L[a] = (TC)g.L[a];
      ta.invL[a] = (TC)g.invL[a];
    }
    ta.rc2 = (TC)(cutoff * cutoff);
    ta.alpha = (TC)alpha;
    ta.coul = (TC)coul;
    ta.c2alpha = (TC)(1.1283791670955126 * alpha);
    void* tok = tff_prof_begin("fused_grad_tiles", st);
    hipLaunchKernelGGL(k_tiles_grad<TC>, dim3((unsigned)w.nw_tiles), dim3(TFF_WAVE), 0, st, w.cpairs, w.ncp,
                       g.pair_cap, (const V4<TC>*)w.xyzq, (const V2<TC>*)w.lj, (const V4<TC>*)w.vsl, w.excl_part, ta,
                       gE, w.gacc);
    tff_prof_end(tok, st);
  }
  PmeGeom pg = make_geom(box, K, order, alpha);
  FftPlan pl[3];
  for (int a = 0; a < 3; ++a) make_plan(K[a], pl[a]);
  const unsigned nbricks = (unsigned)(((K[0] + BR - 1) / BR) * ((K[1] + BR - 1) / BR) * ((K[2] + BR - 1) / BR));
  V4<TC> jsc{(TC)(K[0] * box.inv[0]), (TC)(K[1] * box.inv[4]), (TC)(K[2] * box.inv[8]), (TC)0};
  TC* phi = (TC*)w.grid;
  TC* psi = (TC*)w.grid_d;
  const unsigned nb = (unsigned)((nslots + FB - 1) / FB);
#define TFF_GRAD_PME(P)                                                                                          \
  {                                                                                                              \
    hipLaunchKernelGGL((k_spread_bricks<TC, P, true>), dim3(nbricks), dim3(256), 0, sp, w.ncl,                   \
                       (const V4<TC>*)w.xyzq, w.slot_orig, w.bb_c, w.bb_h, w.col_cl_start, g, pg, w.grid_d,       \
                       status, (const V4<TC>*)w.vsl, jsc);                                                       \
    run_recip_fft<TC>(w, K, box, alpha, pl, mx, my, mz, w.grid_d, 1.0 / Q_SCALE, psi, w.kparts,                 \
                      "fused_grad_fft", sp);                                                                     \
    hipLaunchKernelGGL((k_gather_grad<TC, P>), dim3(nb), dim3(FB), 0, sp, w.ncl, phi, psi, (const V4<TC>*)w.xyzq, \
                       (const V4<TC>*)w.vsl, w.slot_orig, pg, coul, gE, w.gpme);                                 \
  }
  // the potential grid phi of the charges: the forward call's, left in place (reuse), or made
  // again by the forward's own spread and transforms, so that both give the same bits
  if (!reuse) {
    const int rc = launch_pme_slots<TC>(order, w, g, pg, nslots, (const TC*)nullptr, coul, true, status, sp);
    if (rc) {
      aux_stream_join(st, sp);
      return rc;
    }
    run_recip_fft<TC>(w, K, box, alpha, pl, mx, my, mz, w.grid, 1.0 / grid_scale<TC>(), phi, w.kparts,
                      "fused_grad_fft", sp);
  }
  switch (order) {
    case 3: TFF_GRAD_PME(3); break;
    case 4: TFF_GRAD_PME(4); break;
    case 5: TFF_GRAD_PME(5); break;
    case 6: TFF_GRAD_PME(6); break;
    case 7: TFF_GRAD_PME(7); break;
    default: TFF_GRAD_PME(8); break;
  }
#undef TFF_GRAD_PME
  aux_stream_join(st, sp);
  hipLaunchKernelGGL(k_finalize_grad<TI>, dim3(nbS), dim3(FB), 0, st, w.ncl, w.slot_orig, w.gacc, w.gpme, q, coul,
                     alpha, gE, dr, dq, dsig, deps, status, status_host);
  TFF_CHECK_LAUNCH("fused energy+forces backward");
  return TFF_OK;
}

}  // namespace

extern "C" void tff_fused_set_overlap(int mode) { g_overlap.store(mode >= 0 && mode <= 3 ? mode : 1); }

extern "C" size_t tff_fused_workspace_bytes(int compute_dtype, long long natoms, const double* box_host,
                                            double list_cutoff, int K0, int K1, int K2) {
  TffBox box = tff_make_box(box_host);
  Geo g;
  if (!make_geo(box, natoms, list_cutoff, g)) return 0;
  int K[3] = {K0, K1, K2};
  return carve(nullptr, natoms, g, K, compute_dtype == TFF_F64 ? 8 : 4).bytes;
}

namespace {
// checks shared by the forward and backward entry points
int fused_check(const TffBox& box, long long natoms, double cutoff, double list_cutoff, double ewald_alpha,
                const int K[3], int order, const int64_t* excl_row_ptr, const int64_t* excl_col, Geo& g) {
  if (!make_geo(box, natoms, list_cutoff, g)) {
    tff_set_error("fused: needs an orthorhombic periodic box (diagonal box matrix)");
    return TFF_ERR_UNSUPPORTED;
  }
  if (natoms <= 0 || !(cutoff > 0) || list_cutoff < cutoff || ewald_alpha < 0) {
    tff_set_error("fused: invalid arguments");
    return TFF_ERR_ARG;
  }
  if ((excl_row_ptr == nullptr) != (excl_col == nullptr)) {
    tff_set_error("fused: exclusion row_ptr and col must be given together");
    return TFF_ERR_ARG;
  }
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
  return TFF_OK;
}
}  // namespace

extern "C" size_t tff_fused_backward_workspace_bytes(int compute_dtype, long long natoms, const double* box_host,
                                                     double list_cutoff, int K0, int K1, int K2) {
  TffBox box = tff_make_box(box_host);
  Geo g;
  if (!make_geo(box, natoms, list_cutoff, g)) return 0;
  int K[3] = {K0, K1, K2};
  return carve(nullptr, natoms, g, K, compute_dtype == TFF_F64 ? 8 : 4, true).bytes;
}

namespace {
int fused_backward(int io_dtype, int compute_dtype, const void* coords, const void* charges, const void* sigma,
                   const void* epsilon, long long natoms, const double* box_host, const int64_t* excl_row_ptr,
                   const int64_t* excl_col, double cutoff, double list_cutoff, double ewald_alpha,
                   double coul_constant, int K0, int K1, int K2, int order, const void* xmoduli,
                   const void* ymoduli, const void* zmoduli, const double* grad_energy, const void* grad_forces,
                   void* grad_coords, void* grad_charges, void* grad_sigma, void* grad_epsilon, int32_t* status_out,
                   const void* box_device, int32_t* status_host, void* workspace, size_t workspace_bytes,
                   bool reuse, hipStream_t stream) {
  TffBox box = tff_make_box(box_host);
  Geo g;
  int K[3] = {K0, K1, K2};
  int rc = fused_check(box, natoms, cutoff, list_cutoff, ewald_alpha, K, order, excl_row_ptr, excl_col, g);
  if (rc) return rc;
  if (!coords || !charges || !sigma || !epsilon || !status_out || !xmoduli || !ymoduli || !zmoduli) {
    tff_set_error("fused backward: invalid arguments");
    return TFF_ERR_ARG;
  }
  size_t tc = compute_dtype == TFF_F64 ? 8 : 4;
  size_t need = carve(nullptr, natoms, g, K, tc, true).bytes;
  if (!workspace || workspace_bytes < need) {
    tff_set_error("fused backward: workspace too small (%zu < %zu)", workspace_bytes, need);
    return TFF_ERR_WORKSPACE;
  }
  FusedWs w = carve(workspace, natoms, g, K, tc, true);
#define TFF_RUN(TI, TC)                                                                                            \
  return run_fused_grad<TI, TC>((const TI*)coords, (const TI*)charges, (const TI*)sigma, (const TI*)epsilon,       \
                                natoms, box, g, excl_row_ptr, excl_col, cutoff, list_cutoff, ewald_alpha,          \
                                coul_constant, K, order, (const TC*)xmoduli, (const TC*)ymoduli,                  \
                                (const TC*)zmoduli, grad_energy, (const TI*)grad_forces, (TI*)grad_coords,         \
                                (TI*)grad_charges, (TI*)grad_sigma, (TI*)grad_epsilon, (int*)status_out,        \
                                (const TI*)box_device, (int*)status_host, w, reuse, stream)
  if (io_dtype == TFF_F64 && compute_dtype == TFF_F32) TFF_RUN(double, float);
  if (io_dtype == TFF_F32 && compute_dtype == TFF_F32) TFF_RUN(float, float);
  if (io_dtype == TFF_F64 && compute_dtype == TFF_F64) TFF_RUN(double, double);
  if (io_dtype == TFF_F32 && compute_dtype == TFF_F64) TFF_RUN(float, double);
#undef TFF_RUN
  tff_set_error("fused backward: unsupported dtype combination");
  return TFF_ERR_ARG;
}
}  // namespace

extern "C" int tff_fused_energy_forces_backward(
    int io_dtype, int compute_dtype, const void* coords, const void* charges, const void* sigma, const void* epsilon,
    long long natoms, const double* box_host, const int64_t* excl_row_ptr, const int64_t* excl_col, double cutoff,
    double list_cutoff, double ewald_alpha, double coul_constant, int K0, int K1, int K2, int order,
    const void* xmoduli, const void* ymoduli, const void* zmoduli, const double* grad_energy,
    const void* grad_forces, void* grad_coords, void* grad_charges, void* grad_sigma, void* grad_epsilon,
    int32_t* status_out, const void* box_device, int32_t* status_host, void* workspace, size_t workspace_bytes,
    hipStream_t stream) {
  return fused_backward(io_dtype, compute_dtype, coords, charges, sigma, epsilon, natoms, box_host, excl_row_ptr,
                        excl_col, cutoff, list_cutoff, ewald_alpha, coul_constant, K0, K1, K2, order, xmoduli,
                        ymoduli, zmoduli, grad_energy, grad_forces, grad_coords, grad_charges, grad_sigma,
                        grad_epsilon, status_out, box_device, status_host, workspace, workspace_bytes, false, stream);
}

extern "C" int tff_fused_energy_forces_backward_reuse(
    int io_dtype, int compute_dtype, const void* coords, const void* charges, const void* sigma, const void* epsilon,
    long long natoms, const double* box_host, const int64_t* excl_row_ptr, const int64_t* excl_col, double cutoff,
    double list_cutoff, double ewald_alpha, double coul_constant, int K0, int K1, int K2, int order,
    const void* xmoduli, const void* ymoduli, const void* zmoduli, const double* grad_energy,
    const void* grad_forces, void* grad_coords, void* grad_charges, void* grad_sigma, void* grad_epsilon,
    int32_t* status_out, const void* box_device, int32_t* status_host, void* workspace, size_t workspace_bytes,
    hipStream_t stream) {
  return fused_backward(io_dtype, compute_dtype, coords, charges, sigma, epsilon, natoms, box_host, excl_row_ptr,
                        excl_col, cutoff, list_cutoff, ewald_alpha, coul_constant, K0, K1, K2, order, xmoduli,
                        ymoduli, zmoduli, grad_energy, grad_forces, grad_coords, grad_charges, grad_sigma,
                        grad_epsilon, status_out, box_device, status_host, workspace, workspace_bytes, true, stream);
}

namespace {
// ---------------------------------------------------------------- graph replay of the forward core
// The forward core (19 kernels on two streams) is captured once per key into a HIP graph and
// replayed; k_finalize, which writes the caller's outputs, is launched after the replay, so every
// call still returns its own output buffers.  A key holds everything the captured launches read:
// the device, the input and workspace addresses, the box, every scalar parameter and the stream
// schedule.  Replays read the inputs at their addresses when they run, so coordinates updated in
// place are picked up.  Launched one by one, the pair search waits behind the reciprocal-space
// kernels of the other stream; the same kernels replayed give bitwise equal results.
constexpr int GRAPH_CACHE = 8;  // captured forward cores kept per device (least recently used evicted)

struct FusedKey {
  int dev, io, cd, order, K[3], mode;
  long long natoms;
  const void* p[11];  // coords, charges, sigma, epsilon, excl row_ptr, excl col, x/y/z moduli, box_device, workspace
  double d[13];       // box, cutoff, list cutoff, alpha, Coulomb constant
};

struct GraphEntry {
  FusedKey key;
  hipGraphExec_t ex;
  unsigned long long used;
};

std::atomic<int> g_graphs_mode{1};
std::mutex g_graph_mu;
std::vector<GraphEntry> g_graphs[64];
unsigned long long g_graph_clock = 0;
hipStream_t g_capture_stream[64];
FusedKey g_seen[64];  // per device: the key of the last call that found no graph
bool g_seen_valid[64];

// a graph is captured and launched for a call on st: by the mode (0: never; 1: float32 arithmetic
// below SPREAD_FIRST_ATOMS atoms; 2: always), no event timing requested, st not itself being
// captured (an enclosing capture records the kernels one by one instead).  Mode 1 follows the
// measurements per evaluation (graph against one by one): 96k atoms mixed 0.294 against 0.304 ms,
// fp64 0.773 against 0.761, 999k atoms mixed 2.350 against 2.351.
bool graphs_usable(hipStream_t st, bool f32, long long n) {
  const int mode = g_graphs_mode.load();
  if (mode == 0 || (mode == 1 && !(f32 && n < SPREAD_FIRST_ATOMS)) || tff_prof_active()) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return cs == hipStreamCaptureStatusNone;
}

// Captures key's forward core into a new cache entry (caller holds g_graph_mu) when the previous
// uncached key seen on this device was the same key: a key seen once (inputs at new addresses on
// every call, one-off calls) is not captured.  *ex stays null when the call should run without a
// graph; a failure of launch_core itself is returned.
template <typename F>
int graph_capture(const FusedKey& key, F&& launch_core, hipGraphExec_t* ex) {
  *ex = nullptr;
  if (!g_seen_valid[key.dev] || memcmp(&g_seen[key.dev], &key, sizeof key) != 0) {
    g_seen[key.dev] = key;
    g_seen_valid[key.dev] = true;
    return TFF_OK;
  }
  std::vector<GraphEntry>& cache = g_graphs[key.dev];
  hipStream_t& cs = g_capture_stream[key.dev];
  if (!cs && hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) {
    cs = nullptr;
    (void)hipGetLastError();
    return TFF_OK;
  }
  {
    // the auxiliary stream is created outside the capture
    std::lock_guard<std::mutex> lk2(g_aux_mu);
    if (!aux_stream_setup_locked(key.dev)) return TFF_OK;
  }
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    return TFF_OK;
  }
  const int rc = launch_core(cs);
  hipGraph_t graph = nullptr;
  const hipError_t e = hipStreamEndCapture(cs, &graph);
  if (rc != TFF_OK || e != hipSuccess || !graph) {
    if (graph) hipGraphDestroy(graph);
    (void)hipGetLastError();
    return rc;
  }
  hipGraphExec_t g = nullptr;
  const hipError_t ei = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ei != hipSuccess || !g) {
    (void)hipGetLastError();
    return TFF_OK;
  }
  if ((int)cache.size() >= GRAPH_CACHE) {
    size_t v = 0;
    for (size_t k = 1; k < cache.size(); ++k)
      if (cache[k].used < cache[v].used) v = k;
    // a launch of the evicted graph may still be queued: let it finish first
    (void)hipDeviceSynchronize();
    hipGraphExecDestroy(cache[v].ex);
    cache.erase(cache.begin() + (long)v);
  }
  cache.push_back({key, g, ++g_graph_clock});
  g_seen_valid[key.dev] = false;
  *ex = g;
  return TFF_OK;
}

// Launches the executable graph of key on st, from the cache or newly captured (graph_capture).
// The launch happens under the cache's lock, so no other thread evicts the graph meanwhile.
// *launched stays false when this call should run without a graph.
template <typename F>
int graph_launch(const FusedKey& key, F&& launch_core, hipStream_t st, bool* launched) {
  *launched = false;
  std::lock_guard<std::mutex> lk(g_graph_mu);
  hipGraphExec_t ex = nullptr;
  for (GraphEntry& e : g_graphs[key.dev])
    if (memcmp(&e.key, &key, sizeof key) == 0) {
      e.used = ++g_graph_clock;
      ex = e.ex;
      break;
    }
  if (!ex) {
    const int rc = graph_capture(key, launch_core, &ex);
    if (rc || !ex) return rc;
  }
  if (hipGraphLaunch(ex, st) != hipSuccess) {
    tff_set_error("fused: graph launch failed: %s", hipGetErrorString(hipGetLastError()));
    return TFF_ERR_LAUNCH;
  }
  *launched = true;
  return TFF_OK;
}

template <typename TI, typename TC>
int fused_forward_t(int io_dtype, int compute_dtype, const TI* coords, const TI* q, const TI* sig, const TI* eps,
                    long long n, const double* box_host, const TffBox& box, const Geo& g, const int64_t* row_ptr,
                    const int64_t* col, double cutoff, double list_cutoff, double alpha, double coul,
                    const int K[3], int order, const TC* mx, const TC* my, const TC* mz, double* energies,
                    TI* forces, int* status_out, const TI* box_dev, int* status_host, void* workspace,
                    const FusedWs& w, bool graphs, hipStream_t st) {
  auto core = [&](hipStream_t s) {
    return run_fused_core<TI, TC>(coords, q, sig, eps, n, box, g, row_ptr, col, cutoff, list_cutoff, alpha, coul, K,
                                  order, mx, my, mz, box_dev, w, s);
  };
  if (graphs && graphs_usable(st, sizeof(TC) == 4, n)) {
    FusedKey key;
    memset(&key, 0, sizeof key);  // padding included, the key is compared bytewise
    if (hipGetDevice(&key.dev) != hipSuccess || key.dev < 0 || key.dev >= 64) {
      tff_set_error("fused: no current device");
      return TFF_ERR_LAUNCH;
    }
    key.io = io_dtype;
    key.cd = compute_dtype;
    key.order = order;
    for (int a = 0; a < 3; ++a) key.K[a] = K[a];
    key.mode = g_overlap.load();
    key.natoms = n;
    const void* ptrs[11] = {coords, q, sig, eps, row_ptr, col, mx, my, mz, box_dev, workspace};
    for (int k = 0; k < 11; ++k) key.p[k] = ptrs[k];
    for (int k = 0; k < 9; ++k) key.d[k] = box_host[k];
    key.d[9] = cutoff;
    key.d[10] = list_cutoff;
    key.d[11] = alpha;
    key.d[12] = coul;
    bool launched = false;
    const int rc = graph_launch(key, core, st, &launched);
    if (rc) return rc;
    if (launched) return launch_fused_finalize<TI>(w, g, coul, alpha, energies, forces, status_out, status_host, st);
  }
  const int rc = core(st);
  if (rc) return rc;
  return launch_fused_finalize<TI>(w, g, coul, alpha, energies, forces, status_out, status_host, st);
}

int fused_forward(int io_dtype, int compute_dtype, const void* coords, const void* charges, const void* sigma,
                  const void* epsilon, long long natoms, const double* box_host, const int64_t* excl_row_ptr,
                  const int64_t* excl_col, double cutoff, double list_cutoff, double ewald_alpha,
                  double coul_constant, int K0, int K1, int K2, int order, const void* xmoduli, const void* ymoduli,
                  const void* zmoduli, double* energies_out, void* forces_out, int32_t* status_out,
                  const void* box_device, int32_t* status_host, void* workspace, size_t workspace_bytes, bool graphs,
                  hipStream_t stream) {
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
  return fused_forward_t<TI, TC>(io_dtype, compute_dtype, (const TI*)coords, (const TI*)charges, (const TI*)sigma, \
                                 (const TI*)epsilon, natoms, box_host, box, g, excl_row_ptr, excl_col, cutoff,     \
                                 list_cutoff, ewald_alpha, coul_constant, K, order, (const TC*)xmoduli,            \
                                 (const TC*)ymoduli, (const TC*)zmoduli, energies_out, (TI*)forces_out,            \
                                 (int*)status_out, (const TI*)box_device, (int*)status_host, workspace, w, graphs, \
                                 stream)
  if (io_dtype == TFF_F64 && compute_dtype == TFF_F32) TFF_RUN(double, float);
  if (io_dtype == TFF_F32 && compute_dtype == TFF_F32) TFF_RUN(float, float);
  if (io_dtype == TFF_F64 && compute_dtype == TFF_F64) TFF_RUN(double, double);
  if (io_dtype == TFF_F32 && compute_dtype == TFF_F64) TFF_RUN(float, double);
#undef TFF_RUN
  tff_set_error("fused: unsupported dtype combination");
  return TFF_ERR_ARG;
}
}  // namespace

extern "C" void tff_fused_set_graphs(int mode) { g_graphs_mode.store(mode >= 0 && mode <= 2 ? mode : 1); }

extern "C" int tff_fused_energy_forces(int io_dtype, int compute_dtype, const void* coords, const void* charges,
                                       const void* sigma, const void* epsilon, long long natoms,
                                       const double* box_host, const int64_t* excl_row_ptr,
                                       const int64_t* excl_col, double cutoff, double list_cutoff,
                                       double ewald_alpha, double coul_constant, int K0, int K1, int K2, int order,
                                       const void* xmoduli, const void* ymoduli, const void* zmoduli,
                                       double* energies_out, void* forces_out, int32_t* status_out,
                                       const void* box_device, int32_t* status_host, void* workspace,
                                       size_t workspace_bytes, hipStream_t stream) {
  return fused_forward(io_dtype, compute_dtype, coords, charges, sigma, epsilon, natoms, box_host, excl_row_ptr,
                       excl_col, cutoff, list_cutoff, ewald_alpha, coul_constant, K0, K1, K2, order, xmoduli,
                       ymoduli, zmoduli, energies_out, forces_out, status_out, box_device, status_host, workspace,
                       workspace_bytes, true, stream);
}

// ---------------------------------------------------------------- allocation-free variant
// The workspace of tff_fused_energy_forces_noalloc is this array: each device holds one copy for
// the life of the process.  Host code never passes the array's name to a kernel (on the host it
// is only a stand-in for the symbol): the entry point looks up the device address of its copy
// with hipGetSymbolAddress once per device, at its first call there, and uses that address.
// Every call on one device shares the copy (and keeps the previous call's contents; every kernel
// of the call clears or overwrites what it reads first), so calls that overlap on one device race.
static __device__ __align__(256) unsigned char scratch[TFF_FUSED_SCRATCH_BYTES];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[64];

// device address of this device's scratch copy and the auxiliary stream, set up once per device
int scratch_setup(void** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) {
    tff_set_error("fused noalloc: no current device");
    return TFF_ERR_LAUNCH;
  }
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        tff_set_error("fused noalloc: hipGetSymbolAddress of the scratch array failed (%s)", hipGetErrorString(e));
        return TFF_ERR_LAUNCH;
      }
      g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
  }
  std::lock_guard<std::mutex> lk(g_aux_mu);
  if (!aux_stream_setup_locked(dev)) {
    tff_set_error("fused noalloc: could not create the auxiliary stream");
    return TFF_ERR_LAUNCH;
  }
  return TFF_OK;
}
}  // namespace

extern "C" int tff_fused_energy_forces_noalloc(int io_dtype, int compute_dtype, const void* coords,
                                               const void* charges, const void* sigma, const void* epsilon,
                                               long long natoms, const double* box_host,
                                               const int64_t* excl_row_ptr, const int64_t* excl_col, double cutoff,
                                               double list_cutoff, double ewald_alpha, double coul_constant, int K0,
                                               int K1, int K2, int order, const void* xmoduli, const void* ymoduli,
                                               const void* zmoduli, double* energies_out, void* forces_out,
                                               int32_t* status_out, const void* box_device, int32_t* status_host,
                                               hipStream_t stream) {
  const size_t need = tff_fused_workspace_bytes(compute_dtype, natoms, box_host, list_cutoff, K0, K1, K2);
  if (need > TFF_FUSED_SCRATCH_BYTES) {
    tff_set_error("fused noalloc: this system needs %zu bytes of scratch, above the %zu of the static array "
                  "(use tff_fused_energy_forces with a workspace)", need, (size_t)TFF_FUSED_SCRATCH_BYTES);
    return TFF_ERR_WORKSPACE;
  }
  void* ws = nullptr;
  const int rc = scratch_setup(&ws);
  if (rc) return rc;
  return fused_forward(io_dtype, compute_dtype, coords, charges, sigma, epsilon, natoms, box_host, excl_row_ptr,
                       excl_col, cutoff, list_cutoff, ewald_alpha, coul_constant, K0, K1, K2, order, xmoduli,
                       ymoduli, zmoduli, energies_out, forces_out, status_out, box_device, status_host, ws,
                       (size_t)TFF_FUSED_SCRATCH_BYTES, false, stream);
}
