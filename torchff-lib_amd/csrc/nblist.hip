// Periodic neighbor list for gfx950: all-pairs tiles and a cell-list search.
//
// Replaces the reference's build_neighbor_list_nsquared / _cell_list ops
// (csrc/nblist/nblist_nsquared_cuda.cu:29-198, nblist_clist_cuda.cu:44-525).
// Output contract (nblist_interface.cpp:5-8): pairs (max_npairs, 2) int64,
// rows past the found count set to -1, each row (a, b) with a > b (a == b for
// self pairs), npairs int32 holding the number found (it may exceed
// max_npairs; then the extra pairs were not stored).
//
// Predicate: |minimg(r_a - r_b)|^2 <= cutoff^2 in float64 with a fixed
// operation order and no contraction, on the input coordinates promoted to
// float64 (so float32 inputs are judged exactly as the float64 CPU oracle
// judges the same values); self pairs only with include_self; CSR exclusions.
//
// Cell search: atoms are binned in fractional space into nc^3 cells of
// width >= cutoff/2 along every reciprocal direction (or >= cutoff when the
// box is too small for 5 cells), counting-sorted by cell, and one workgroup
// of four waves per home cell walks the half shell of 62 (or 13) neighbour
// cells plus the home cell, the home atoms and their exclusion partners staged
// in LDS once and each neighbour cell's atoms staged by the wave that takes
// it.  The ni x nj candidate block of a cell pair is spread over the 64 lanes,
// so lane utilisation does not depend on cell occupancy.  Two passes: the
// first counts each wave's accepted pairs, a scan turns the counts into
// offsets, the second repeats the walk and writes each wave's pairs from its
// offset in ballot order (no shared counter: one global atomic per few
// hundred pairs on a single address cost more than the search itself), so the
// output order is the same on every call.  The rows past the found count are
// set to -1 after the search (each output row is written once).

#include <stdarg.h>
#include <stdio.h>

#include "tff_common.h"
#include "torchff_hip.h"

namespace {

constexpr int NB_BLOCK = 256;
constexpr int NB_WAVES = NB_BLOCK / TFF_WAVE;
constexpr int NB_BUF = 512;  // pairs per wave buffer

struct PairBuf {
  long long* s;  // 2*NB_BUF entries (a, b)
  int n;
};

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void flush(PairBuf& pb, int lane, long long* pairs, long long max_npairs,
                                      int* npairs) {
  if (pb.n == 0) return;
  wave_sync_lds();
  int base = 0;
  if (lane == 0) base = atomicAdd(npairs, pb.n);
  base = __shfl(base, 0, TFF_WAVE);
  for (int k = lane; k < pb.n; k += TFF_WAVE) {
    long long pos = (long long)base + k;
    if (pos < max_npairs) {
      longlong2 v;
      v.x = pb.s[2 * k];
      v.y = pb.s[2 * k + 1];
      reinterpret_cast<longlong2*>(pairs)[pos] = v;
    }
  }
  wave_sync_lds();
  pb.n = 0;
}

__device__ __forceinline__ void push(PairBuf& pb, bool ok, long long a, long long b, int lane,
                                     long long* pairs, long long max_npairs, int* npairs) {
  unsigned long long m = __ballot(ok);
  int cnt = __popcll(m);
  if (cnt == 0) return;
  if (ok) {
    int pos = pb.n + __popcll(m & ((1ull << lane) - 1ull));
    long long hi = a > b ? a : b, lo = a > b ? b : a;
    pb.s[2 * pos] = hi;
    pb.s[2 * pos + 1] = lo;
  }
  pb.n += cnt;
  if (pb.n > NB_BUF - TFF_WAVE) flush(pb, lane, pairs, max_npairs, npairs);
}

// rows from the found count on (the found rows are already written) -> -1
__global__ void k_fill_tail(long long* pairs, long long max_npairs, const int* npairs) {
  const long long n0 = (long long)*npairs < max_npairs ? (long long)*npairs : max_npairs;
  long long i = 2 * n0 + blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < 2 * max_npairs; i += stride) pairs[i] = -1;
}

__global__ void k_fill_pairs(long long* pairs, long long n2, int* npairs) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n2; i += stride) pairs[i] = -1;
  if (blockIdx.x == 0 && threadIdx.x == 0) *npairs = 0;
}

template <typename T>
__global__ void k_promote(const T* __restrict__ coords, long long n, double4* __restrict__ out) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = make_double4((double)coords[3 * i], (double)coords[3 * i + 1], (double)coords[3 * i + 2], 0.0);
}

// ---------------------------------------------------------------- all pairs
// One wavefront per 64x64 tile (I <= J) of the upper triangle.
__global__ __launch_bounds__(NB_BLOCK) void k_nsquared(const double4* __restrict__ xyz, long long n,
                                                       TffBox box, double cutoff2, int include_self,
                                                       const int64_t* __restrict__ row_ptr,
                                                       const int64_t* __restrict__ col,
                                                       long long* __restrict__ pairs, long long max_npairs,
                                                       int* __restrict__ npairs, long long ntiles,
                                                       long long nb) {
  __shared__ long long sbuf[NB_WAVES][2 * NB_BUF];
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int w = threadIdx.x / TFF_WAVE;
  PairBuf pb{sbuf[w], 0};
  long long gw = (long long)blockIdx.x * NB_WAVES + w;
  long long nwaves = (long long)gridDim.x * NB_WAVES;
  for (long long t = gw; t < ntiles; t += nwaves) {
    // tile t -> (I, J), I <= J, row-major over I
    double fI = floor(((2.0 * nb + 1.0) - sqrt((2.0 * nb + 1.0) * (2.0 * nb + 1.0) - 8.0 * (double)t)) * 0.5);
    long long I = (long long)fI;
    if (I < 0) I = 0;
    while (I > 0 && I * nb - I * (I - 1) / 2 > t) --I;
    while ((I + 1) * nb - (I + 1) * I / 2 <= t) ++I;
    long long J = I + (t - (I * nb - I * (I - 1) / 2));
    long long i = I * TFF_WAVE + lane;
    double4 ri = i < n ? xyz[i] : make_double4(0, 0, 0, 0);
    long long j0 = J * TFF_WAVE;
    int jn = (int)((n - j0) < TFF_WAVE ? (n - j0) : TFF_WAVE);
    for (int jj = 0; jj < jn; ++jj) {
      long long j = j0 + jj;
      double4 rj = xyz[j];
      bool ok = false;
      if (i < n && j >= i) {
        if (j == i) {
          ok = include_self != 0;
        } else {
          double d2 = tff_dist2_d(box, ri.x - rj.x, ri.y - rj.y, ri.z - rj.z);
          ok = d2 <= cutoff2;
        }
        if (ok) ok = !tff_is_excluded(i, j, row_ptr, col);
      }
      push(pb, ok, i, j, lane, pairs, max_npairs, npairs);
    }
  }
  flush(pb, lane, pairs, max_npairs, npairs);
}

// ---------------------------------------------------------------- cell list
struct CellGrid {
  int nc[3];
  int reach;  // 1 or 2 cells
  int ncell;
  // float32 prefilter of the cutoff test (orthorhombic box, every axis at least 2 reach + 3 cells,
  // so a cell pair's periodic shift gives the minimum image): candidates whose float32 squared
  // distance is below lo are in range, above hi out of range, the rest take the float64 predicate
  int pf;
  float lo, hi;
  float L[3];
};

__global__ void k_bin(const double4* __restrict__ xyz, long long n, TffBox box, CellGrid g,
                      int* __restrict__ cell_of, int* __restrict__ count) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double4 r = xyz[i];
  double f[3];
  f[0] = r.x * box.inv[0] + r.y * box.inv[3] + r.z * box.inv[6];
  f[1] = r.x * box.inv[1] + r.y * box.inv[4] + r.z * box.inv[7];
  f[2] = r.x * box.inv[2] + r.y * box.inv[5] + r.z * box.inv[8];
  int c[3];
  for (int k = 0; k < 3; ++k) {
    double fk = f[k] - floor(f[k]);
    int ck = (int)(fk * g.nc[k]);
    c[k] = ck < 0 ? 0 : (ck >= g.nc[k] ? g.nc[k] - 1 : ck);
  }
  int cell = (c[0] * g.nc[1] + c[1]) * g.nc[2] + c[2];
  cell_of[i] = cell;
  atomicAdd(&count[cell], 1);
}

// Exclusive scan of count[0..m) into start[0..m], single workgroup of 1024 threads.
__global__ __launch_bounds__(1024) void k_scan(const int* __restrict__ count, int m, int* __restrict__ start) {
  tff_block_scan_1024(count, m, start, true);
}

// also the prefilter's float32 copy (orthorhombic boxes): coordinates wrapped into the box, the atom
// index in w (bit pattern)
__global__ void k_scatter(const double4* __restrict__ xyz, long long n, TffBox box, const int* __restrict__ cell_of,
                          const int* __restrict__ start, int* __restrict__ cursor,
                          int* __restrict__ sorted_idx, double4* __restrict__ sorted_xyz,
                          float4* __restrict__ sorted_f) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = cell_of[i];
  int pos = start[c] + atomicAdd(&cursor[c], 1);
  sorted_idx[pos] = (int)i;
  const double4 r = xyz[i];
  sorted_xyz[pos] = r;
  if (box.ortho) {
    const double f[3] = {r.x * box.inv[0], r.y * box.inv[4], r.z * box.inv[8]};
    sorted_f[pos] = make_float4((float)(r.x - box.b[0] * floor(f[0])), (float)(r.y - box.b[4] * floor(f[1])),
                                (float)(r.z - box.b[8] * floor(f[2])), __int_as_float((int)i));
  }
}

// One workgroup of CP_WAVES waves per home cell.  The home cell's atoms (in chunks of 64) and their
// exclusion partners (up to CP_EX each; more are looked up in the CSR arrays) are staged in LDS
// once; each wave takes every CP_WAVES-th cell of the half shell (the cells' starts and sizes loaded
// at once, one lane per cell), stages that cell's atoms (first 64 prefetched into registers while
// the previous cell is walked) in its own LDS area and walks the ni x nj candidate block 64 at a
// time, the lane's (a, b) advanced incrementally (one division per cell).  WRITE = false: each
// wave's accepted-pair count into wcount[home * CP_WAVES + wave]; WRITE = true: the pairs from
// woff[home * CP_WAVES + wave] on, in ballot order (rows at or past max_npairs are not stored).
constexpr int CP_WAVES = 4;
constexpr int CP_HOME = 64;  // home atoms per chunk
constexpr int CP_EX = 4;     // exclusion partners per home atom kept in LDS

template <bool WRITE>
__global__ __launch_bounds__(CP_WAVES * TFF_WAVE) void k_cell_pairs(
    const int* __restrict__ start, const int* __restrict__ sorted_idx, const double4* __restrict__ sxyz,
    const float4* __restrict__ sf, CellGrid g, TffBox box, double cutoff2, int include_self,
    const int64_t* __restrict__ row_ptr, const int64_t* __restrict__ col, int* __restrict__ wcount,
    const int* __restrict__ woff, long long* __restrict__ pairs, long long max_npairs, int* __restrict__ npairs) {
  // staged atoms: x, y, z and the atom index (exact in a double) in w
  __shared__ double4 s_h[CP_HOME];
  // the home atoms' exclusion partners, -1 padded; all -2: more than CP_EX (CSR lookup)
  __shared__ int4 s_hx[CP_HOME];
  __shared__ double4 s_n[CP_WAVES][TFF_WAVE];
  // prefilter copies (wrapped float32 coordinates, atom index bits in w); the neighbour copies carry
  // their cell pair's periodic shift
  __shared__ float4 s_hf[CP_HOME];
  __shared__ float4 s_nf[CP_WAVES][TFF_WAVE];
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int wv = threadIdx.x / TFF_WAVE;
  const int home = blockIdx.x;
  if (home >= g.ncell) return;
  // WRITE: the next row of this wave; else its count
  long long pos = WRITE ? (long long)woff[home * CP_WAVES + wv] : 0;
  if (WRITE && home == 0 && threadIdx.x == 0) *npairs = woff[g.ncell * CP_WAVES];
  const int hs = start[home];
  const int ni_all = start[home + 1] - hs;
  const int hz = home % g.nc[2];
  const int hy = (home / g.nc[2]) % g.nc[1];
  const int hx = home / (g.nc[2] * g.nc[1]);
  const int R = g.reach;
  const int side = 2 * R + 1;
  const int noff = side * side * side;
  for (int h0 = 0; h0 < ni_all; h0 += CP_HOME) {
    const int ni = ni_all - h0 < CP_HOME ? ni_all - h0 : CP_HOME;
    __syncthreads();  // the previous chunk's readers are done
    if (threadIdx.x < ni) {
      const int pa = hs + h0 + threadIdx.x;
      const int ai = sorted_idx[pa];
      double4 r = sxyz[pa];
      r.w = (double)ai;
      s_h[threadIdx.x] = r;
      if (g.pf) s_hf[threadIdx.x] = sf[pa];
      int ex[CP_EX] = {-1, -1, -1, -1};
      if (col) {
        const int64_t e0 = row_ptr[ai], e1 = row_ptr[ai + 1];
        if (e1 - e0 <= CP_EX) {
          for (int k = 0; k < CP_EX; ++k)
            if (k < e1 - e0) ex[k] = (int)col[e0 + k];
        } else {
          for (int k = 0; k < CP_EX; ++k) ex[k] = -2;
        }
      }
      s_hx[threadIdx.x] = make_int4(ex[0], ex[1], ex[2], ex[3]);
    }
    __syncthreads();
    // this wave's cells of the half shell (offsets with linear index >= centre, the centre first
    // being the home cell): o = noff / 2 + wv + CP_WAVES c; lane c holds cell c's start and size
    const int ncw = (noff - noff / 2 - wv + CP_WAVES - 1) / CP_WAVES;
    int my_js = 0, my_nj = 0;
    float my_sx = 0.f, my_sy = 0.f, my_sz = 0.f;  // the cell's periodic shift (prefilter)
    if (lane < ncw) {
      const int o = noff / 2 + wv + CP_WAVES * lane;
      const int ox = o / (side * side) - R, oy = (o / side) % side - R, oz = o % side - R;
      int nx = hx + ox, ny = hy + oy, nz = hz + oz;
      my_sx = nx < 0 ? -g.L[0] : (nx >= g.nc[0] ? g.L[0] : 0.f);
      my_sy = ny < 0 ? -g.L[1] : (ny >= g.nc[1] ? g.L[1] : 0.f);
      my_sz = nz < 0 ? -g.L[2] : (nz >= g.nc[2] ? g.L[2] : 0.f);
      nx += nx < 0 ? g.nc[0] : (nx >= g.nc[0] ? -g.nc[0] : 0);
      ny += ny < 0 ? g.nc[1] : (ny >= g.nc[1] ? -g.nc[1] : 0);
      nz += nz < 0 ? g.nc[2] : (nz >= g.nc[2] ? -g.nc[2] : 0);
      const int nb = (nx * g.nc[1] + ny) * g.nc[2] + nz;
      my_js = start[nb];
      my_nj = start[nb + 1] - my_js;
    }
    // the first 64 atoms of the next cell are loaded into registers while the current one is walked
    double4 pre_x = make_double4(0, 0, 0, 0);
    float4 pre_f = make_float4(0, 0, 0, 0);
    int pre_i = 0;
    if (ncw > 0) {
      const int js = __shfl(my_js, 0, TFF_WAVE), nj = __shfl(my_nj, 0, TFF_WAVE);
      if (lane < nj) {
        pre_x = sxyz[js + lane];
        pre_i = sorted_idx[js + lane];
        if (g.pf) pre_f = sf[js + lane];
      }
    }
    for (int c = 0; c < ncw; ++c) {
      const int js = __shfl(my_js, c, TFF_WAVE);
      const int nj_all = __shfl(my_nj, c, TFF_WAVE);
      const bool self_cell = wv == 0 && c == 0;
      const float shx = __shfl(my_sx, c, TFF_WAVE), shy = __shfl(my_sy, c, TFF_WAVE),
                  shz = __shfl(my_sz, c, TFF_WAVE);
      // this cell's first 64 atoms from the registers, then the next cell's load is issued (also
      // when this cell is empty)
      if (lane < nj_all) {
        s_n[wv][lane] = make_double4(pre_x.x, pre_x.y, pre_x.z, (double)pre_i);
        s_nf[wv][lane] = make_float4(pre_f.x + shx, pre_f.y + shy, pre_f.z + shz, pre_f.w);
      }
      tff_wave_sync_lds();
      if (c + 1 < ncw) {
        const int jsn = __shfl(my_js, c + 1, TFF_WAVE), njn = __shfl(my_nj, c + 1, TFF_WAVE);
        if (lane < njn) {
          pre_x = sxyz[jsn + lane];
          pre_i = sorted_idx[jsn + lane];
          if (g.pf) pre_f = sf[jsn + lane];
        }
      }
      for (int j0 = 0; j0 < nj_all; j0 += TFF_WAVE) {
        const int nj = nj_all - j0 < TFF_WAVE ? nj_all - j0 : TFF_WAVE;
        if (j0 > 0) {
          if (lane < nj) {
            const double4 r = sxyz[js + j0 + lane];
            s_n[wv][lane] = make_double4(r.x, r.y, r.z, (double)sorted_idx[js + j0 + lane]);
            if (g.pf) {
              const float4 f = sf[js + j0 + lane];
              s_nf[wv][lane] = make_float4(f.x + shx, f.y + shy, f.z + shz, f.w);
            }
          }
          tff_wave_sync_lds();
        }
        const int total = ni * nj;
        int a = lane / nj, b = lane - a * nj;
        const int da = TFF_WAVE / nj, db = TFF_WAVE - da * nj;
        for (int t0 = 0; t0 < total; t0 += TFF_WAVE) {
          bool ok = false;
          int ai = 0, bj = 0;
          if (a < ni) {
            const int pa = h0 + a, pb = j0 + b;  // positions within their cells
            if (!self_cell || pa <= pb) {
              const int4 hx = s_hx[a];
              bool exact = true;
              if (g.pf) {
                const float4 fa = s_hf[a], fb = s_nf[wv][b];
                ai = __float_as_int(fa.w);
                bj = __float_as_int(fb.w);
                const float ex = fa.x - fb.x, ey = fa.y - fb.y, ez = fa.z - fb.z;
                const float d2f = ex * ex + ey * ey + ez * ez;
                ok = d2f < g.lo;
                exact = !(d2f < g.lo) && !(d2f > g.hi);
              }
              if (self_cell && pa == pb) {
                const double4 ra = s_h[a];
                ai = bj = (int)ra.w;
                ok = include_self != 0;
              } else if (exact) {
                const double4 ra = s_h[a], rb = s_n[wv][b];
                ai = (int)ra.w;
                bj = (int)rb.w;
                const double d2 = tff_dist2_d(box, ra.x - rb.x, ra.y - rb.y, ra.z - rb.z);
                ok = d2 <= cutoff2;
              }
              if (ok && col) {
                bool ex = (hx.x == bj) | (hx.y == bj) | (hx.z == bj) | (hx.w == bj);
                if (hx.x == -2) ex = tff_is_excluded(ai, bj, row_ptr, col);
                ok = !ex;
              }
            }
          }
          const unsigned long long m = __ballot(ok);
          if (WRITE && ok) {
            const long long row = pos + __popcll(m & ((1ull << lane) - 1ull));
            if (row < max_npairs) {
              longlong2 v;
              v.x = ai > bj ? ai : bj;
              v.y = ai > bj ? bj : ai;
              reinterpret_cast<longlong2*>(pairs)[row] = v;
            }
          }
          pos += __popcll(m);
          a += da;
          b += db;
          if (b >= nj) {
            b -= nj;
            ++a;
          }
        }
        tff_wave_sync_lds();  // before the next chunk restages s_n
      }
    }
  }
  if (!WRITE && lane == 0) wcount[home * CP_WAVES + wv] = (int)pos;
}
bool choose_grid(const TffBox& box, double cutoff, long long n, CellGrid& g) {
  if (!box.periodic || !(cutoff > 0.0)) return false;
  double w[3];
  for (int k = 0; k < 3; ++k) {
    double c0 = box.inv[k], c1 = box.inv[3 + k], c2 = box.inv[6 + k];
    w[k] = 1.0 / sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  }
  const double rc = cutoff * (1.0 + 1e-9);
  const double cap = 8.0 * (double)(n > 512 ? n : 512);
  for (int reach = 2; reach >= 1; --reach) {
    int nc[3];
    bool okk = true;
    double tot = 1.0;
    for (int k = 0; k < 3; ++k) {
      double v = floor(w[k] * reach / rc);
      if (v > 1024) v = 1024;
      nc[k] = (int)v;
      if (nc[k] < 2 * reach + 1) okk = false;
      tot *= v;
    }
    if (!okk) continue;
    if (tot > cap) {
      if (reach == 2) continue;  // too sparse for half cells: try full cells
      // shrink the grid uniformly (cells only grow, so the stencil stays valid)
      double s = cbrt(cap / tot);
      for (int k = 0; k < 3; ++k) {
        int v = (int)floor(nc[k] * s);
        nc[k] = v < 3 ? 3 : v;
      }
    }
    g.nc[0] = nc[0]; g.nc[1] = nc[1]; g.nc[2] = nc[2];
    g.reach = reach;
    g.ncell = nc[0] * nc[1] * nc[2];
    g.pf = box.ortho && nc[0] >= 2 * reach + 3 && nc[1] >= 2 * reach + 3 && nc[2] >= 2 * reach + 3;
    {
      // float32 error of the squared distance: each displacement component within 6 u L (the two
      // roundings of the wrapped coordinates, the shift, two subtractions), so d^2 within
      // 21 u L d + 3 u d^2; the band is eight times that at d = cutoff
      const double u = 5.9604644775390625e-8, Lm = fmax(box.b[0], fmax(box.b[4], box.b[8]));
      const double band = u * (168.0 * Lm * cutoff + 24.0 * cutoff * cutoff);
      g.lo = (float)(cutoff * cutoff - band);
      g.hi = (float)(cutoff * cutoff + band);
      for (int k = 0; k < 3; ++k) g.L[k] = (float)box.b[4 * k];
    }
    return true;
  }
  return false;
}

int launch_nsquared(const double4* xyz, long long n, const TffBox& box, double cutoff2, int include_self,
                    const int64_t* row_ptr, const int64_t* col, long long* pairs, long long max_npairs,
                    int* npairs, hipStream_t stream) {
  long long nb = (n + TFF_WAVE - 1) / TFF_WAVE;
  long long ntiles = nb * (nb + 1) / 2;
  if (ntiles == 0) return TFF_OK;
  long long blocks = (ntiles + NB_WAVES - 1) / NB_WAVES;
  if (blocks > 8192) blocks = 8192;
  void* tok = tff_prof_begin("nblist_nsquared", stream);
  hipLaunchKernelGGL(k_nsquared, dim3((unsigned)blocks), dim3(NB_BLOCK), 0, stream, xyz, n, box, cutoff2,
                     include_self, row_ptr, col, pairs, max_npairs, npairs, ntiles, nb);
  tff_prof_end(tok, stream);
  TFF_CHECK_LAUNCH("k_nsquared");
  return TFF_OK;
}

struct NbWs {
  double4* xyz;
  double4* sxyz;
  int* cell_of;
  int* count;
  int* start;
  int* cursor;
  int* sidx;
  int *wcount, *woff;  // per (home cell, wave) pair counts and their exclusive scan
  float4* sf;          // prefilter copies of the sorted atoms
  size_t bytes;
};

NbWs carve(void* base, long long n, int ncell) {
  NbWs w{};
  size_t off = 0;
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* q = p ? p + off : nullptr;
    off += tff_align(b);
    return q;
  };
  w.xyz = (double4*)take(sizeof(double4) * (size_t)(n > 0 ? n : 1));
  w.sxyz = (double4*)take(sizeof(double4) * (size_t)(n > 0 ? n : 1));
  w.cell_of = (int*)take(sizeof(int) * (size_t)(n > 0 ? n : 1));
  w.sidx = (int*)take(sizeof(int) * (size_t)(n > 0 ? n : 1));
  w.count = (int*)take(sizeof(int) * (size_t)(ncell + 1));
  w.start = (int*)take(sizeof(int) * (size_t)(ncell + 1));
  w.cursor = (int*)take(sizeof(int) * (size_t)(ncell + 1));
  w.wcount = (int*)take(sizeof(int) * ((size_t)ncell * CP_WAVES + 1));
  w.woff = (int*)take(sizeof(int) * ((size_t)ncell * CP_WAVES + 1));
  w.sf = (float4*)take(sizeof(float4) * (size_t)(n > 0 ? n : 1));
  w.bytes = off;
  return w;
}

__global__ void k_zero_int(int* p, int m) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m) p[i] = 0;
}

}  // namespace

extern "C" size_t tff_nblist_workspace_bytes(long long natoms, const double* box_host, double cutoff,
                                             int algorithm) {
  TffBox box = tff_make_box(box_host);
  CellGrid g{{1, 1, 1}, 1, 0};
  int ncell = 0;
  if (algorithm == TFF_NBLIST_CELL_LIST && choose_grid(box, cutoff, natoms, g)) ncell = g.ncell;
  return carve(nullptr, natoms, ncell).bytes;
}

extern "C" int tff_build_neighbor_list(int dtype, const void* coords, long long natoms, const double* box_host,
                                       double cutoff, int algorithm, int include_self,
                                       const int64_t* excl_row_ptr, const int64_t* excl_col,
                                       int64_t* pairs, long long max_npairs, int32_t* npairs, void* workspace,
                                       size_t workspace_bytes, hipStream_t stream) {
  if (natoms < 0 || max_npairs < 0 || !npairs || (max_npairs > 0 && !pairs)) {
    tff_set_error("tff_build_neighbor_list: invalid arguments");
    return TFF_ERR_ARG;
  }
  if (natoms >= (1LL << 31) - 1) {
    tff_set_error("tff_build_neighbor_list: natoms exceeds int32 indexing");
    return TFF_ERR_ARG;
  }
  if (algorithm == TFF_NBLIST_CELL_LIST && !box_host) {
    tff_set_error("Cell-list neighbor list requires a periodic box");
    return TFF_ERR_ARG;
  }
  if ((excl_row_ptr == nullptr) != (excl_col == nullptr)) {
    tff_set_error("tff_build_neighbor_list: exclusion row_ptr and col must be given together");
    return TFF_ERR_ARG;
  }
  TffBox box = tff_make_box(box_host);
  size_t need = tff_nblist_workspace_bytes(natoms, box_host, cutoff, algorithm);
  if (workspace_bytes < need || (need > 0 && !workspace)) {
    tff_set_error("tff_build_neighbor_list: workspace too small (%zu < %zu)", workspace_bytes, need);
    return TFF_ERR_WORKSPACE;
  }
  long long* P = reinterpret_cast<long long*>(pairs);
  int* NP = reinterpret_cast<int*>(npairs);
  // the count starts at zero; rows past the found count are set to -1 after the search (the found
  // rows are written once)
  hipLaunchKernelGGL(k_fill_pairs, dim3(1), dim3(64), 0, stream, P, 0LL, NP);
  TFF_CHECK_LAUNCH("k_fill_pairs");
  auto fill_tail = [&]() {
    long long blocks = (2 * max_npairs + 255) / 256;
    if (blocks > 65535) blocks = 65535;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_fill_tail, dim3((unsigned)blocks), dim3(256), 0, stream, P, max_npairs, (const int*)NP);
    TFF_CHECK_LAUNCH("k_fill_tail");
    return TFF_OK;
  };
  if (natoms == 0) return fill_tail();
  const double cutoff2 = cutoff * cutoff;
  CellGrid g{{1, 1, 1}, 1, 0};
  bool use_cells = algorithm == TFF_NBLIST_CELL_LIST && choose_grid(box, cutoff, natoms, g);
  NbWs ws = carve(workspace, natoms, use_cells ? g.ncell : 0);
  unsigned nbk = (unsigned)((natoms + 255) / 256);
  if (dtype == TFF_F64)
    hipLaunchKernelGGL(k_promote<double>, dim3(nbk), dim3(256), 0, stream, (const double*)coords, natoms, ws.xyz);
  else if (dtype == TFF_F32)
    hipLaunchKernelGGL(k_promote<float>, dim3(nbk), dim3(256), 0, stream, (const float*)coords, natoms, ws.xyz);
  else {
    tff_set_error("tff_build_neighbor_list: unsupported dtype %d", dtype);
    return TFF_ERR_ARG;
  }
  TFF_CHECK_LAUNCH("k_promote");
  if (!use_cells) {
    const int rc = launch_nsquared(ws.xyz, natoms, box, cutoff2, include_self, excl_row_ptr, excl_col, P,
                                   max_npairs, NP, stream);
    return rc ? rc : fill_tail();
  }
  const int m = g.ncell;
  hipLaunchKernelGGL(k_zero_int, dim3((m + 1 + 255) / 256), dim3(256), 0, stream, ws.count, m + 1);
  hipLaunchKernelGGL(k_zero_int, dim3((m + 1 + 255) / 256), dim3(256), 0, stream, ws.cursor, m + 1);
  hipLaunchKernelGGL(k_bin, dim3(nbk), dim3(256), 0, stream, ws.xyz, natoms, box, g, ws.cell_of, ws.count);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, stream, ws.count, m, ws.start);
  hipLaunchKernelGGL(k_scatter, dim3(nbk), dim3(256), 0, stream, ws.xyz, natoms, box, ws.cell_of, ws.start,
                     ws.cursor, ws.sidx, ws.sxyz, ws.sf);
  void* tok = tff_prof_begin("nblist_cell_pairs", stream);
  hipLaunchKernelGGL(k_cell_pairs<false>, dim3(m), dim3(CP_WAVES * TFF_WAVE), 0, stream, ws.start, ws.sidx,
                     ws.sxyz, (const float4*)ws.sf, g, box, cutoff2, include_self, excl_row_ptr, excl_col,
                     ws.wcount, (const int*)ws.woff, P, max_npairs, NP);
  hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, stream, ws.wcount, m * CP_WAVES, ws.woff);
  hipLaunchKernelGGL(k_cell_pairs<true>, dim3(m), dim3(CP_WAVES * TFF_WAVE), 0, stream, ws.start, ws.sidx,
                     ws.sxyz, (const float4*)ws.sf, g, box, cutoff2, include_self, excl_row_ptr, excl_col,
                     ws.wcount, (const int*)ws.woff, P, max_npairs, NP);
  tff_prof_end(tok, stream);
  TFF_CHECK_LAUNCH("k_cell_pairs");
  return fill_tail();
}
