// Cluster-pair neighbour list in the reference's five-tensor format, its decoding into atom
// pairs, and the nonbonded energy / forces evaluated directly on it (no (P, 2) pair list).
//
// Format (torchff/nblist.py:260-344; csrc/nblist/nblist_cluster_pairs_cuda.cu:18-525, which the
// reference's setup.py does not build, and the commented schemas of
// csrc/nonbonded/nonbonded_interface.cpp:5-8):
//   sorted_atom_indices (N, int32)     slot -> atom; cluster c owns slots 32c .. 32c + 31
//   cluster_exclusions (2, E, int32)   cluster pairs (x, y), x <= y, whose 32 x 32 block holds an
//                                      excluded atom pair, plus every diagonal (x, x)
//   bitmask_exclusions (E, 32, uint32) word [e][l], bit k: skip (slot 32x + l, slot 32y + k)
//                                      (excluded, an empty slot, or k <= l on the diagonal)
//   interacting_clusters (M, int32)    row -> cluster x
//   interacting_atoms (M, 32, int32)   atoms of clusters y > x (not in x's exclusion set) whose
//                                      minimum-image distance to x's bounding sphere is within the
//                                      cutoff, 32 per row in cluster-then-slot order, -1 padded
// Every atom pair within the cutoff and not excluded is visited exactly once: clusters listed in
// x's exclusion set through their bitmask block, all others through x's rows.
//
// Build (deterministic: counting passes, scans, then writes at fixed offsets):
//   k_cp_keys / k_cp_scan / k_cp_scatter / k_cp_rank  cell sort (cells cs x cs x cs/10 as in the
//        reference, z fastest; refined when they would hold more than 8 atoms on average), the
//        atoms of a cell ordered by index;
//   k_cp_bounds      half a wave per cluster: minimum-image offsets from the first atom, bounding
//                    box centre and enclosing radius;
//   k_cp_excl        half a wave per cluster: distinct partner clusters y >= x of its exclusions;
//   k_cp_rows<WRITE> half a wave per cluster: sphere tests against the clusters above it, then
//                    atom tests; the count pass sizes the rows, the write pass fills them.
// The sphere tests are exact for the fractional-rounding minimum image while the radii plus the
// cutoff stay below half the smallest box height; beyond that the candidates are kept untested
// (the traversal applies the exact cutoff test to every pair).
//
// Traversal (decode and nonbonded): half a wave per entry (exclusion blocks first, then rows);
// lane l holds atom i of slot 32x + l and atom j of column l, and walks the 32 columns by
// shuffles.  Pairs come out in entry, lane, column order at offsets from a count pass and a scan.

#include "tff_common.h"
#include "torchff_hip.h"

namespace {

constexpr int CPW = 32;        // atoms per cluster / row width of the format
constexpr int CPB = 256;       // threads per workgroup (4 waves, 8 half-waves)
constexpr int CP_EXMAX = 16;   // exclusion partners per atom
constexpr int CP_XCL = 32;     // distinct exclusion partner clusters per cluster
constexpr int CP_MAXBLK = 4096;

enum { CP_FLAG_EXCL_ATOM = 1, CP_FLAG_EXCL_CLUSTER = 2 };

__device__ __forceinline__ int half_lane() { return threadIdx.x & 31; }
__device__ __forceinline__ int half_index() { return (blockIdx.x * CPB + threadIdx.x) >> 5; }
__device__ __forceinline__ int half_count() { return (gridDim.x * CPB) >> 5; }
// bits of this half-wave in a wave-wide ballot
__device__ __forceinline__ unsigned half_ballot(bool v) {
  const unsigned long long b = __builtin_amdgcn_ballot_w64(v);
  return (unsigned)(b >> (threadIdx.x & 32));
}
__device__ __forceinline__ int half_min(int v) {
  for (int o = 16; o > 0; o >>= 1) {
    const int w = __shfl_xor(v, o, 32);
    v = w < v ? w : v;
  }
  return v;
}
__device__ __forceinline__ double half_max(double v) {
  for (int o = 16; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, 32);
    v = w > v ? w : v;
  }
  return v;
}
__device__ __forceinline__ double half_minf(double v) {
  for (int o = 16; o > 0; o >>= 1) {
    const double w = __shfl_xor(v, o, 32);
    v = w < v ? w : v;
  }
  return v;
}

struct CpGeom {
  TffBox box;
  double rc2;
  double rc;
  double half_h;  // half the smallest box height
  int nc[3];
  int natoms;
  int nclusters;
};

template <typename T>
__device__ __forceinline__ void load_xyz(const T* c, int a, double& x, double& y, double& z) {
  x = (double)c[3 * a];
  y = (double)c[3 * a + 1];
  z = (double)c[3 * a + 2];
}

// ---------------------------------------------------------------- cell sort
template <typename T>
__global__ void k_cp_keys(const T* __restrict__ coords, CpGeom g, int* __restrict__ cell_of,
                          int* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.natoms) return;
  double x, y, z;
  load_xyz(coords, i, x, y, z);
  const double f[3] = {x * g.box.inv[0] + y * g.box.inv[3] + z * g.box.inv[6],
                       x * g.box.inv[1] + y * g.box.inv[4] + z * g.box.inv[7],
                       x * g.box.inv[2] + y * g.box.inv[5] + z * g.box.inv[8]};
  int c[3];
  for (int k = 0; k < 3; ++k) {
    const double fk = f[k] - floor(f[k]);
    const int ck = (int)(fk * g.nc[k]);
    c[k] = ck < 0 ? 0 : (ck >= g.nc[k] ? g.nc[k] - 1 : ck);
  }
  const int cell = (c[0] * g.nc[1] + c[1]) * g.nc[2] + c[2];
  cell_of[i] = cell;
  atomicAdd(&count[cell], 1);
}

// exclusive scan of cnt[0..m) into off[0..m] (off[m] = total), one workgroup of 1024 threads
template <typename C, typename O>
__global__ __launch_bounds__(1024) void k_cp_scan(const C* __restrict__ cnt, long long m, O* __restrict__ off) {
  tff_block_scan_1024<C, O>(cnt, m, off, true);
}

__global__ void k_cp_scatter(int n, const int* __restrict__ cell_of, const int* __restrict__ start,
                             int* __restrict__ cursor, int* __restrict__ seg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell_of[i];
  seg[start[c] + atomicAdd(&cursor[c], 1)] = i;
}

// place each atom at its rank by index within its cell
__global__ void k_cp_rank(int n, const int* __restrict__ cell_of, const int* __restrict__ start,
                          const int* __restrict__ seg, int* __restrict__ sorted, int* __restrict__ inv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = cell_of[i];
  const int b = start[c], e = start[c + 1];
  int r = 0;
  for (int k = b; k < e; ++k) r += seg[k] < i;
  sorted[b + r] = i;
  inv[i] = b + r;
}

// ---------------------------------------------------------------- cluster bounds
template <typename T>
__global__ __launch_bounds__(CPB) void k_cp_bounds(const T* __restrict__ coords, CpGeom g,
                                                   const int* __restrict__ sorted, double4* __restrict__ sphere) {
  const int l = half_lane();
  for (int c = half_index(); c < g.nclusters; c += half_count()) {
    const int slot = c * CPW + l;
    const bool ok = slot < g.natoms;
    double x = 0, y = 0, z = 0;
    if (ok) load_xyz(coords, sorted[slot], x, y, z);
    const double ax = __shfl(x, 0, 32), ay = __shfl(y, 0, 32), az = __shfl(z, 0, 32);
    double dx = x - ax, dy = y - ay, dz = z - az;
    tff_min_image_d(g.box, dx, dy, dz);
    const double lo[3] = {half_minf(ok ? dx : 1e300), half_minf(ok ? dy : 1e300), half_minf(ok ? dz : 1e300)};
    const double hi[3] = {half_max(ok ? dx : -1e300), half_max(ok ? dy : -1e300), half_max(ok ? dz : -1e300)};
    const double mx = 0.5 * (lo[0] + hi[0]), my = 0.5 * (lo[1] + hi[1]), mz = 0.5 * (lo[2] + hi[2]);
    const double ex = dx - mx, ey = dy - my, ez = dz - mz;
    const double r = half_max(ok ? sqrt(ex * ex + ey * ey + ez * ez) : 0.0);
    if (l == 0) sphere[c] = make_double4(ax + mx, ay + my, az + mz, r * (1.0 + 1e-12) + 1e-12);
  }
}

// ---------------------------------------------------------------- exclusion blocks
// the smallest partner cluster above `last` among atom i's exclusions (INT_MAX if none)
__device__ __forceinline__ int next_partner(int i, int x, int last, const long long* __restrict__ rp,
                                            const long long* __restrict__ col, const int* __restrict__ inv,
                                            int natoms, int* flags) {
  int best = 0x7fffffff;
  if (i < 0 || !rp) return best;  // no exclusion list: only the diagonal block
  const long long b = rp[i], e = rp[i + 1];
  if (e - b > CP_EXMAX && flags) atomicOr(flags, CP_FLAG_EXCL_ATOM);
  for (long long k = b; k < e && k < b + CP_EXMAX; ++k) {
    const long long j = col[k];
    if (j < 0 || j >= natoms) continue;
    const int cy = inv[j] / CPW;
    if (cy >= x && cy > last && cy < best) best = cy;
  }
  return best;
}

__global__ __launch_bounds__(CPB) void k_cp_excl(CpGeom g, const int* __restrict__ sorted,
                                                 const int* __restrict__ inv, const long long* __restrict__ rp,
                                                 const long long* __restrict__ col, int* __restrict__ xlist,
                                                 int* __restrict__ xcount, int* __restrict__ flags) {
  const int l = half_lane();
  for (int x = half_index(); x < g.nclusters; x += half_count()) {
    const int slot = x * CPW + l;
    const int i = slot < g.natoms ? sorted[slot] : -1;
    int n = 0;
    if (l == 0) xlist[x * CP_XCL] = x;  // the diagonal block is always present
    n = 1;
    int last = x;
    while (true) {
      const int nx = half_min(next_partner(i, x, last, rp, col, inv, g.natoms, flags));
      if (nx == 0x7fffffff) break;
      if (n >= CP_XCL) {
        if (l == 0) atomicOr(flags, CP_FLAG_EXCL_CLUSTER);
        break;
      }
      if (l == 0) xlist[x * CP_XCL + n] = nx;
      ++n;
      last = nx;
    }
    if (l == 0) xcount[x] = n;
  }
}

__device__ __forceinline__ bool excluded_pair(int i, int j, const long long* __restrict__ rp,
                                              const long long* __restrict__ col) {
  if (!rp) return false;
  const long long b = rp[i], e = rp[i + 1];
  for (long long k = b; k < e; ++k)
    if (col[k] == j) return true;
  return false;
}

__global__ __launch_bounds__(CPB) void k_cp_excl_emit(CpGeom g, const int* __restrict__ sorted,
                                                      const long long* __restrict__ rp,
                                                      const long long* __restrict__ col,
                                                      const int* __restrict__ xlist, const int* __restrict__ xcount,
                                                      const int* __restrict__ xoff, int ecap,
                                                      int* __restrict__ coo, unsigned* __restrict__ mask) {
  const int l = half_lane();
  for (int x = half_index(); x < g.nclusters; x += half_count()) {
    const int slot = x * CPW + l;
    const int i = slot < g.natoms ? sorted[slot] : -1;
    const int n = xcount[x];
    for (int k = 0; k < n; ++k) {
      const int o = xoff[x] + k;
      if (o >= ecap) break;
      const int y = xlist[x * CP_XCL + k];
      const int jslot = y * CPW + l;
      const int jl = jslot < g.natoms ? sorted[jslot] : -1;
      unsigned m = 0;
      for (int c = 0; c < CPW; ++c) {
        const int j = __shfl(jl, c, 32);
        if (i < 0 || j < 0 || (x == y && c <= l) || excluded_pair(i, j, rp, col)) m |= 1u << c;
      }
      mask[(size_t)o * CPW + l] = m;
      if (l == 0) {
        coo[o] = x;
        coo[ecap + o] = y;
      }
    }
  }
}

// ---------------------------------------------------------------- interacting rows
template <typename T, bool WRITE>
__global__ __launch_bounds__(CPB) void k_cp_rows(const T* __restrict__ coords, CpGeom g,
                                                 const int* __restrict__ sorted,
                                                 const double4* __restrict__ sphere,
                                                 const int* __restrict__ xlist, const int* __restrict__ xcount,
                                                 int* __restrict__ natoms_of, const long long* __restrict__ row_off,
                                                 long long mcap, int* __restrict__ ic, int* __restrict__ ia) {
  const int l = half_lane();
  for (int x = half_index(); x < g.nclusters; x += half_count()) {
    const double4 sx = sphere[x];
    const int nx = xcount[x];
    const bool all_atoms = sx.w + g.rc > g.half_h;
    long long total = 0;
    const long long base = WRITE ? row_off[x] : 0;
    for (int y0 = x + 1; y0 < g.nclusters; y0 += CPW) {
      const int y = y0 + l;
      bool inc = false;
      if (y < g.nclusters) {
        bool ex = false;
        for (int k = 1; k < nx; ++k) ex |= xlist[x * CP_XCL + k] == y;
        if (!ex) {
          const double4 sy = sphere[y];
          const double bound = sx.w + sy.w + g.rc;
          if (bound > g.half_h) {
            inc = true;
          } else {
            double dx = sx.x - sy.x, dy = sx.y - sy.y, dz = sx.z - sy.z;
            tff_min_image_d(g.box, dx, dy, dz);
            inc = dx * dx + dy * dy + dz * dz <= bound * bound;
          }
        }
      }
      unsigned cand = half_ballot(inc);
      while (cand) {
        const int k = __builtin_ctz(cand);
        cand &= cand - 1;
        const int yy = y0 + k;
        const int slot = yy * CPW + l;
        bool hit = false;
        int a = -1;
        if (slot < g.natoms) {
          a = sorted[slot];
          if (all_atoms) {
            hit = true;
          } else {
            double px, py, pz;
            load_xyz(coords, a, px, py, pz);
            double dx = px - sx.x, dy = py - sx.y, dz = pz - sx.z;
            tff_min_image_d(g.box, dx, dy, dz);
            const double rr = sx.w + g.rc;
            hit = dx * dx + dy * dy + dz * dz <= rr * rr;
          }
        }
        const unsigned hb = half_ballot(hit);
        if (WRITE && hit) {
          const long long ord = total + __builtin_popcount(hb & ((1u << l) - 1u));
          const long long row = base + ord / CPW;
          if (row < mcap) ia[row * CPW + ord % CPW] = a;
        }
        total += __builtin_popcount(hb);
      }
    }
    if (!WRITE) {
      if (l == 0) natoms_of[x] = (int)((total + CPW - 1) / CPW);
    } else {
      const long long rows = (total + CPW - 1) / CPW;
      for (long long r = l; r < rows; r += CPW)
        if (base + r < mcap) ic[base + r] = x;
      // pad the last row
      const long long last = total % CPW;
      if (last && l >= last && base + rows - 1 < mcap) ia[(base + rows - 1) * CPW + l] = -1;
    }
  }
}

__global__ void k_cp_fill(int* p, long long from, long long to, int v) {
  for (long long k = from + blockIdx.x * (long long)blockDim.x + threadIdx.x; k < to;
       k += (long long)gridDim.x * blockDim.x)
    p[k] = v;
}

__global__ void k_cp_totals(const int* __restrict__ xoff, const long long* __restrict__ row_off, int nclusters,
                            const int* __restrict__ flags, long long* __restrict__ out) {
  out[0] = xoff[nclusters];
  out[1] = row_off[nclusters];
  out[2] = flags[0];
}

// ---------------------------------------------------------------- traversal
struct Entries {
  const int* sorted;
  const int* coo;      // (2, E)
  long long E;
  const unsigned* mask;
  const int* ic;       // (M)
  const int* ia;       // (M, 32)
  long long M;
};

// lane's atom i, column atom j and skip mask for entry e; false if the entry is padding
// the atom of a slot, -1 for a slot past the end or an index out of range
__device__ __forceinline__ int slot_atom(const int* __restrict__ sorted, long long slot, int natoms) {
  if (slot >= natoms) return -1;
  const int a = sorted[slot];
  return a >= 0 && a < natoms ? a : -1;
}

__device__ __forceinline__ bool entry_atoms(const Entries& en, long long e, int natoms, int nclusters, int l, int& i,
                                            int& j, unsigned& skip) {
  if (e < en.E) {
    const int x = en.coo[e], y = en.coo[en.E + e];
    if (x < 0 || y < 0 || x > y || y >= nclusters) return false;
    i = slot_atom(en.sorted, (long long)x * CPW + l, natoms);
    j = slot_atom(en.sorted, (long long)y * CPW + l, natoms);
    skip = en.mask[e * CPW + l];
    return true;
  }
  const long long r = e - en.E;
  const int x = en.ic[r];
  if (x < 0 || x >= nclusters) return false;
  i = slot_atom(en.sorted, (long long)x * CPW + l, natoms);
  j = en.ia[r * CPW + l];
  if (j >= natoms) j = -1;
  skip = 0;
  return true;
}

enum { TR_COUNT = 0, TR_WRITE = 1 };

template <typename T, int MODE>
__global__ __launch_bounds__(CPB) void k_cp_decode(const T* __restrict__ coords, CpGeom g, Entries en,
                                                   int* __restrict__ cnt, const long long* __restrict__ off,
                                                   long long cap, int* __restrict__ pairs) {
  const int l = half_lane();
  const long long total = en.E + en.M;
  for (long long e = half_index(); e < total; e += half_count()) {
    int i = -1, j = -1;
    unsigned skip = 0xffffffffu;
    const bool live = entry_atoms(en, e, g.natoms, g.nclusters, l, i, j, skip);
    double xi = 0, yi = 0, zi = 0, xj = 0, yj = 0, zj = 0;
    if (live && i >= 0) load_xyz(coords, i, xi, yi, zi);
    if (live && j >= 0) load_xyz(coords, j, xj, yj, zj);
    int n = 0;
    unsigned hits = 0;
    for (int c = 0; c < CPW; ++c) {
      const int jc = __shfl(j, c, 32);
      const double dx0 = xi - __shfl(xj, c, 32), dy0 = yi - __shfl(yj, c, 32), dz0 = zi - __shfl(zj, c, 32);
      double dx = dx0, dy = dy0, dz = dz0;
      tff_min_image_d(g.box, dx, dy, dz);
      const bool ok = live && i >= 0 && jc >= 0 && !((skip >> c) & 1u) && dx * dx + dy * dy + dz * dz <= g.rc2;
      hits |= (unsigned)ok << c;
      n += ok;
    }
    if (MODE == TR_COUNT) {
      int s = n;
      for (int o = 16; o > 0; o >>= 1) s += __shfl_xor(s, o, 32);
      if (l == 0) cnt[e] = s;
    } else {
      // lanes below this one write first
      int pre = n;
      for (int o = 1; o < 32; o <<= 1) {
        const int v = __shfl_up(pre, o, 32);
        if (l >= o) pre += v;
      }
      long long p = off[e] + pre - n;
      for (int c = 0; c < CPW; ++c) {
        const int jc = __shfl(j, c, 32);
        if ((hits >> c) & 1u) {
          if (p < cap) {
            pairs[2 * p] = i;
            pairs[2 * p + 1] = jc;
          }
          ++p;
        }
      }
    }
  }
}

__global__ void k_cp_npairs(const long long* __restrict__ off, long long total, long long cap, int* __restrict__ np,
                            int* __restrict__ pairs) {
  const long long n = off[total];
  if (blockIdx.x == 0 && threadIdx.x == 0) np[0] = (int)n;
  // -1 padding after the last pair
  for (long long k = n + blockIdx.x * (long long)blockDim.x + threadIdx.x; k < cap;
       k += (long long)gridDim.x * blockDim.x) {
    pairs[2 * k] = -1;
    pairs[2 * k + 1] = -1;
  }
}

// nonbonded energy and gradients on the cluster pairs: k q_i q_j (1/r - [1/rc]) + 4 eps s^6 (s^6 - 1),
// sigma and eps by Lorentz-Berthelot, the same pair term as compute_nonbonded_energy_from_atom_pairs
template <typename T>
__global__ __launch_bounds__(CPB) void k_cp_nonbonded(const T* __restrict__ coords, CpGeom g, Entries en,
                                                      const T* __restrict__ q, const T* __restrict__ sig,
                                                      const T* __restrict__ eps, double coul, int do_shift,
                                                      double* __restrict__ e_part, double* __restrict__ g_xyz,
                                                      double* __restrict__ g_q, double* __restrict__ g_sig,
                                                      double* __restrict__ g_eps) {
  const int l = half_lane();
  const long long total = en.E + en.M;
  const T inv_rc = (T)(1.0 / g.rc);
  double e_acc = 0.0;
  for (long long e = half_index(); e < total; e += half_count()) {
    int i = -1, j = -1;
    unsigned skip = 0xffffffffu;
    const bool live = entry_atoms(en, e, g.natoms, g.nclusters, l, i, j, skip);
    T xi = 0, yi = 0, zi = 0, qi = 0, si = 0, ei = 0;
    T xj = 0, yj = 0, zj = 0, qj = 0, sj = 0, ej = 0;
    if (live && i >= 0) {
      xi = coords[3 * i]; yi = coords[3 * i + 1]; zi = coords[3 * i + 2];
      qi = q[i]; si = sig[i]; ei = eps[i];
    }
    if (live && j >= 0) {
      xj = coords[3 * j]; yj = coords[3 * j + 1]; zj = coords[3 * j + 2];
      qj = q[j]; sj = sig[j]; ej = eps[j];
    }
    double fi[3] = {0, 0, 0}, gqi = 0, gsi = 0, gei = 0;
    for (int c = 0; c < CPW; ++c) {
      const int jc = __shfl(j, c, 32);
      const T xc = __shfl(xj, c, 32), yc = __shfl(yj, c, 32), zc = __shfl(zj, c, 32);
      const T qc = __shfl(qj, c, 32), sc = __shfl(sj, c, 32), ec = __shfl(ej, c, 32);
      // the pair predicate in float64, as the decoded list
      double ddx = (double)xi - (double)xc, ddy = (double)yi - (double)yc, ddz = (double)zi - (double)zc;
      tff_min_image_d(g.box, ddx, ddy, ddz);
      const bool ok = live && i >= 0 && jc >= 0 && !((skip >> c) & 1u) && ddx * ddx + ddy * ddy + ddz * ddz <= g.rc2;
      T dedr_over_r = 0, gqc = 0, gsc = 0, gec = 0;
      T dx = 0, dy = 0, dz = 0;
      if (ok) {
        dx = xi - xc;
        dy = yi - yc;
        dz = zi - zc;
        tff_min_image<T>(g.box, dx, dy, dz);
        const T r2 = dx * dx + dy * dy + dz * dz;
        const T rinv = (T)1 / sqrt(r2);
        const T rinv2 = rinv * rinv;
        const T qq = qi * qc * (T)coul;
        const T shift = do_shift ? inv_rc : (T)0;
        T en_ = qq * (rinv - shift);
        dedr_over_r = -qq * rinv * rinv2;
        const T tmp = (rinv - shift) * (T)coul;
        gqi += (double)(qc * tmp);
        gqc = qi * tmp;
        const T sij = (si + sc) * (T)0.5;
        const T eij = sqrt(ei * ec);
        const T sr2 = sij * sij * rinv2;
        const T s6 = sr2 * sr2 * sr2;
        en_ += (T)4 * eij * s6 * (s6 - (T)1);
        dedr_over_r += -(T)24 * eij * s6 * ((T)2 * s6 - (T)1) * rinv2;
        const T dsig = (T)24 * eij * s6 * ((T)2 * s6 - (T)1) / sij;
        const T deps = (T)4 * s6 * (s6 - (T)1);
        gsi += (double)(dsig * (T)0.5);
        gsc = dsig * (T)0.5;
        if (eij > (T)0) {
          const T de = deps * (T)0.5 / eij;
          gei += (double)(de * ec);
          gec = de * ei;
        }
        e_acc += (double)en_;
        fi[0] += (double)(dedr_over_r * dx);
        fi[1] += (double)(dedr_over_r * dy);
        fi[2] += (double)(dedr_over_r * dz);
      }
      // the column atom's share: summed over the half-wave in the input precision (only the
      // requested gradients), one atomic per column
      const unsigned any = half_ballot(ok);
      if (any) {
        if (g_xyz) {
          T cx = ok ? -(dedr_over_r * dx) : (T)0, cy = ok ? -(dedr_over_r * dy) : (T)0,
            cz = ok ? -(dedr_over_r * dz) : (T)0;
          for (int o = 16; o > 0; o >>= 1) {
            cx += __shfl_xor(cx, o, 32);
            cy += __shfl_xor(cy, o, 32);
            cz += __shfl_xor(cz, o, 32);
          }
          if (l == c) {
            atomicAdd(&g_xyz[3 * jc], (double)cx);
            atomicAdd(&g_xyz[3 * jc + 1], (double)cy);
            atomicAdd(&g_xyz[3 * jc + 2], (double)cz);
          }
        }
        if (g_q || g_sig || g_eps) {
          T cq = ok ? gqc : (T)0, cs = ok ? gsc : (T)0, ce = ok ? gec : (T)0;
          for (int o = 16; o > 0; o >>= 1) {
            cq += __shfl_xor(cq, o, 32);
            cs += __shfl_xor(cs, o, 32);
            ce += __shfl_xor(ce, o, 32);
          }
          if (l == c) {
            if (g_q) atomicAdd(&g_q[jc], (double)cq);
            if (g_sig) atomicAdd(&g_sig[jc], (double)cs);
            if (g_eps) atomicAdd(&g_eps[jc], (double)ce);
          }
        }
      }
    }
    if (live && i >= 0) {
      if (g_xyz && (fi[0] != 0.0 || fi[1] != 0.0 || fi[2] != 0.0)) {
        atomicAdd(&g_xyz[3 * i], fi[0]);
        atomicAdd(&g_xyz[3 * i + 1], fi[1]);
        atomicAdd(&g_xyz[3 * i + 2], fi[2]);
      }
      if (g_q && gqi != 0.0) atomicAdd(&g_q[i], gqi);
      if (g_sig && gsi != 0.0) atomicAdd(&g_sig[i], gsi);
      if (g_eps && gei != 0.0) atomicAdd(&g_eps[i], gei);
    }
  }
  const double s = tff_block_sum(e_acc);
  if (threadIdx.x == 0) e_part[blockIdx.x] = s;
}

template <typename T>
__global__ void k_cp_energy_final(const double* __restrict__ part, int n, double scale, T* __restrict__ out) {
  const double s = tff_sum_partials(part, n);
  if (threadIdx.x == 0) out[0] = (T)(s * scale);
}

template <typename T>
__global__ void k_cp_cast(const double* __restrict__ s, long long n, double scale, int accumulate, T* __restrict__ d) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    d[k] = accumulate ? d[k] + (T)(s[k] * scale) : (T)(s[k] * scale);
}

__global__ void k_cp_zero_d(double* p, long long n) {
  for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
    p[k] = 0.0;
}

// ---------------------------------------------------------------- host side
int make_geom(const double* box_host, long long natoms, double cutoff, double cell_size, CpGeom& g) {
  if (!box_host || natoms < 1 || natoms > 0x3fffffffLL || !(cutoff > 0) || !(cell_size > 0)) {
    tff_set_error("cluster pairs: invalid arguments (box, 1 <= natoms < 2^30, cutoff > 0, cell_size > 0)");
    return TFF_ERR_ARG;
  }
  g.box = tff_make_box(box_host);
  if (!(g.box.det > 0)) {
    tff_set_error("cluster pairs: the box must have a positive volume");
    return TFF_ERR_ARG;
  }
  double h[3];
  for (int a = 0; a < 3; ++a) {
    const double c0 = g.box.inv[a], c1 = g.box.inv[3 + a], c2 = g.box.inv[6 + a];
    h[a] = 1.0 / sqrt(c0 * c0 + c1 * c1 + c2 * c2);
  }
  g.half_h = 0.5 * (h[0] < h[1] ? (h[0] < h[2] ? h[0] : h[2]) : (h[1] < h[2] ? h[1] : h[2]));
  if (cutoff > g.half_h) {
    tff_set_error("cluster pairs: cutoff %g exceeds half the smallest box height %g", cutoff, g.half_h);
    return TFF_ERR_ARG;
  }
  const double w[3] = {cell_size, cell_size, cell_size / 10.0};
  long long nc[3];
  for (int a = 0; a < 3; ++a) {
    nc[a] = (long long)floor(h[a] / w[a]);
    if (nc[a] < 1) nc[a] = 1;
    if (nc[a] > 4096) nc[a] = 4096;
  }
  // refine along z while the cells would hold more than 8 atoms on average
  while (natoms > 8LL * nc[0] * nc[1] * nc[2] && nc[2] < 1 << 20) nc[2] *= 2;
  if (nc[0] * nc[1] * nc[2] > (1LL << 28)) {
    tff_set_error("cluster pairs: cell grid too large (cell_size %g)", cell_size);
    return TFF_ERR_ARG;
  }
  for (int a = 0; a < 3; ++a) g.nc[a] = (int)nc[a];
  g.rc = cutoff;
  g.rc2 = cutoff * cutoff;
  g.natoms = (int)natoms;
  g.nclusters = (int)((natoms + CPW - 1) / CPW);
  return TFF_OK;
}

struct BuildWs {
  int* cell_of;
  int* count;    // ncell
  int* start;    // ncell + 1
  int* cursor;   // ncell
  int* seg;
  int* sorted;
  int* inv;
  double4* sphere;
  int* xlist;    // C * CP_XCL
  int* xcount;   // C
  int* xoff;     // C + 1
  int* rows;     // C
  long long* row_off;  // C + 1
  int* flags;
  size_t bytes;
};

BuildWs carve_build(void* base, const CpGeom& g) {
  BuildWs w{};
  size_t off = 0;
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* r = p ? p + off : nullptr;
    off += tff_align(b > 0 ? b : 1);
    return r;
  };
  const size_t n = (size_t)g.natoms, c = (size_t)g.nclusters;
  const size_t ncell = (size_t)g.nc[0] * g.nc[1] * g.nc[2];
  w.cell_of = (int*)take(4 * n);
  w.count = (int*)take(4 * ncell);
  w.start = (int*)take(4 * (ncell + 1));
  w.cursor = (int*)take(4 * ncell);
  w.seg = (int*)take(4 * n);
  w.sorted = (int*)take(4 * n);
  w.inv = (int*)take(4 * n);
  w.sphere = (double4*)take(sizeof(double4) * c);
  w.xlist = (int*)take(4 * c * CP_XCL);
  w.xcount = (int*)take(4 * c);
  w.xoff = (int*)take(4 * (c + 1));
  w.rows = (int*)take(4 * c);
  w.row_off = (long long*)take(8 * (c + 1));
  w.flags = (int*)take(4 * 4);
  w.bytes = off;
  return w;
}

unsigned grid_for(long long items_per_half, long long n) {
  long long b = (n * items_per_half + (CPB / 32) - 1) / (CPB / 32);
  if (b < 1) b = 1;
  if (b > CP_MAXBLK) b = CP_MAXBLK;
  return (unsigned)b;
}

template <typename T>
int run_build(const T* coords, const CpGeom& g, const long long* rp, const long long* col, const BuildWs& w,
              long long* counts_out, hipStream_t st) {
  const int n = g.natoms;
  const long long ncell = (long long)g.nc[0] * g.nc[1] * g.nc[2];
  const unsigned nb = (unsigned)((n + 255) / 256);
  auto fill = [&](int* p, long long cnt) {
    long long b = (cnt + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_cp_fill, dim3((unsigned)b), dim3(256), 0, st, p, 0LL, cnt, 0);
  };
  fill(w.count, ncell);
  fill(w.cursor, ncell);
  fill(w.flags, 4);
  hipLaunchKernelGGL(k_cp_keys<T>, dim3(nb), dim3(256), 0, st, coords, g, w.cell_of, w.count);
  hipLaunchKernelGGL((k_cp_scan<int, int>), dim3(1), dim3(1024), 0, st, w.count, ncell, w.start);
  hipLaunchKernelGGL(k_cp_scatter, dim3(nb), dim3(256), 0, st, n, w.cell_of, w.start, w.cursor, w.seg);
  hipLaunchKernelGGL(k_cp_rank, dim3(nb), dim3(256), 0, st, n, w.cell_of, w.start, w.seg, w.sorted, w.inv);
  const unsigned gc = grid_for(1, g.nclusters);
  hipLaunchKernelGGL(k_cp_bounds<T>, dim3(gc), dim3(CPB), 0, st, coords, g, w.sorted, w.sphere);
  hipLaunchKernelGGL(k_cp_excl, dim3(gc), dim3(CPB), 0, st, g, w.sorted, w.inv, rp, col, w.xlist, w.xcount,
                     w.flags);
  hipLaunchKernelGGL((k_cp_scan<int, int>), dim3(1), dim3(1024), 0, st, w.xcount, (long long)g.nclusters, w.xoff);
  hipLaunchKernelGGL((k_cp_rows<T, false>), dim3(gc), dim3(CPB), 0, st, coords, g, w.sorted, w.sphere, w.xlist,
                     w.xcount, w.rows, (const long long*)nullptr, 0LL, (int*)nullptr, (int*)nullptr);
  hipLaunchKernelGGL((k_cp_scan<int, long long>), dim3(1), dim3(1024), 0, st, w.rows, (long long)g.nclusters,
                     w.row_off);
  hipLaunchKernelGGL(k_cp_totals, dim3(1), dim3(1), 0, st, w.xoff, w.row_off, g.nclusters, w.flags, counts_out);
  TFF_CHECK_LAUNCH("cluster pairs build");
  return TFF_OK;
}

template <typename T>
int run_emit(const T* coords, const CpGeom& g, const long long* rp, const long long* col, const BuildWs& w,
             int* sorted_out, int* coo_out, unsigned* mask_out, long long ecap, int* ic_out, int* ia_out,
             long long mcap, hipStream_t st) {
  const unsigned gc = grid_for(1, g.nclusters);
  auto fill = [&](int* p, long long from, long long to, int v) {
    long long b = (to - from + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_cp_fill, dim3((unsigned)b), dim3(256), 0, st, p, from, to, v);
  };
  hipMemcpyAsync(sorted_out, w.sorted, 4 * (size_t)g.natoms, hipMemcpyDeviceToDevice, st);
  // padding first (-1 everywhere), then the entries over it, in stream order
  if (ecap > 0) {
    fill(coo_out, 0, 2 * ecap, -1);
    fill((int*)mask_out, 0, ecap * CPW, -1);
  }
  if (mcap > 0) {
    fill(ic_out, 0, mcap, -1);
    fill(ia_out, 0, mcap * CPW, -1);
  }
  hipLaunchKernelGGL(k_cp_excl_emit, dim3(gc), dim3(CPB), 0, st, g, w.sorted, rp, col, w.xlist, w.xcount, w.xoff,
                     (int)ecap, coo_out, mask_out);
  hipLaunchKernelGGL((k_cp_rows<T, true>), dim3(gc), dim3(CPB), 0, st, coords, g, w.sorted, w.sphere, w.xlist,
                     w.xcount, (int*)nullptr, (const long long*)w.row_off, mcap, ic_out, ia_out);
  TFF_CHECK_LAUNCH("cluster pairs emit");
  return TFF_OK;
}

struct TravWs {
  int* cnt;
  long long* off;
  double* e_part;
  double* g_xyz;
  double* g_q;
  double* g_sig;
  double* g_eps;
  size_t bytes;
};

TravWs carve_trav(void* base, long long natoms, long long entries) {
  TravWs w{};
  size_t off = 0;
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* r = p ? p + off : nullptr;
    off += tff_align(b > 0 ? b : 1);
    return r;
  };
  w.cnt = (int*)take(4 * (size_t)entries);
  w.off = (long long*)take(8 * ((size_t)entries + 1));
  w.e_part = (double*)take(8 * (size_t)CP_MAXBLK);
  w.g_xyz = (double*)take(24 * (size_t)natoms);
  w.g_q = (double*)take(8 * (size_t)natoms);
  w.g_sig = (double*)take(8 * (size_t)natoms);
  w.g_eps = (double*)take(8 * (size_t)natoms);
  w.bytes = off;
  return w;
}

int check_entries(const Entries& en, const char* name) {
  if (!en.sorted || en.E < 0 || en.M < 0 || (en.E > 0 && (!en.coo || !en.mask)) || (en.M > 0 && (!en.ic || !en.ia)) ||
      en.E + en.M > 0x7fffffffLL) {
    tff_set_error("%s: invalid cluster-pair arguments", name);
    return TFF_ERR_ARG;
  }
  return TFF_OK;
}

}  // namespace

extern "C" size_t tff_cluster_pairs_workspace_bytes(long long natoms, const double* box_host, double cutoff,
                                                    double cell_size) {
  CpGeom g{};
  if (make_geom(box_host, natoms, cutoff, cell_size, g)) return 0;
  return carve_build(nullptr, g).bytes;
}

extern "C" int tff_cluster_pairs_build(int dtype, const void* coords, long long natoms, const double* box_host,
                                       double cutoff, double cell_size, const int64_t* excl_row_ptr,
                                       const int64_t* excl_col, long long* counts_out, void* workspace,
                                       size_t workspace_bytes, hipStream_t stream) {
  CpGeom g{};
  int rc = make_geom(box_host, natoms, cutoff, cell_size, g);
  if (rc) return rc;
  if (!coords || !counts_out || (!excl_row_ptr) != (!excl_col) || (dtype != TFF_F32 && dtype != TFF_F64)) {
    tff_set_error("tff_cluster_pairs_build: invalid arguments");
    return TFF_ERR_ARG;
  }
  const BuildWs w = carve_build(workspace, g);
  if (!workspace || workspace_bytes < w.bytes) {
    tff_set_error("tff_cluster_pairs_build: workspace too small (%zu < %zu)", workspace_bytes, w.bytes);
    return TFF_ERR_WORKSPACE;
  }
  const long long* rp = reinterpret_cast<const long long*>(excl_row_ptr);
  const long long* col = reinterpret_cast<const long long*>(excl_col);
  if (dtype == TFF_F64) return run_build<double>((const double*)coords, g, rp, col, w, counts_out, stream);
  return run_build<float>((const float*)coords, g, rp, col, w, counts_out, stream);
}

extern "C" int tff_cluster_pairs_emit(int dtype, const void* coords, long long natoms, const double* box_host,
                                      double cutoff, double cell_size, const int64_t* excl_row_ptr,
                                      const int64_t* excl_col, int32_t* sorted_atom_indices,
                                      int32_t* cluster_exclusions, uint32_t* bitmask_exclusions,
                                      long long excl_capacity, int32_t* interacting_clusters,
                                      int32_t* interacting_atoms, long long row_capacity, void* workspace,
                                      size_t workspace_bytes, hipStream_t stream) {
  CpGeom g{};
  int rc = make_geom(box_host, natoms, cutoff, cell_size, g);
  if (rc) return rc;
  if (!coords || !sorted_atom_indices || excl_capacity < 0 || row_capacity < 0 ||
      (excl_capacity > 0 && (!cluster_exclusions || !bitmask_exclusions)) ||
      (row_capacity > 0 && (!interacting_clusters || !interacting_atoms)) || excl_capacity > 0x7fffffffLL ||
      (!excl_row_ptr) != (!excl_col) || (dtype != TFF_F32 && dtype != TFF_F64)) {
    tff_set_error("tff_cluster_pairs_emit: invalid arguments");
    return TFF_ERR_ARG;
  }
  const BuildWs w = carve_build(workspace, g);
  if (!workspace || workspace_bytes < w.bytes) {
    tff_set_error("tff_cluster_pairs_emit: workspace too small (%zu < %zu)", workspace_bytes, w.bytes);
    return TFF_ERR_WORKSPACE;
  }
  const long long* rp = reinterpret_cast<const long long*>(excl_row_ptr);
  const long long* col = reinterpret_cast<const long long*>(excl_col);
  if (dtype == TFF_F64)
    return run_emit<double>((const double*)coords, g, rp, col, w, sorted_atom_indices, cluster_exclusions,
                            bitmask_exclusions, excl_capacity, interacting_clusters, interacting_atoms, row_capacity,
                            stream);
  return run_emit<float>((const float*)coords, g, rp, col, w, sorted_atom_indices, cluster_exclusions,
                         bitmask_exclusions, excl_capacity, interacting_clusters, interacting_atoms, row_capacity,
                         stream);
}

extern "C" size_t tff_cluster_pairs_traverse_workspace_bytes(long long natoms, long long num_exclusion_blocks,
                                                             long long num_rows) {
  if (natoms < 0 || num_exclusion_blocks < 0 || num_rows < 0) return 0;
  return carve_trav(nullptr, natoms, num_exclusion_blocks + num_rows).bytes;
}

extern "C" int tff_cluster_pairs_decode(int dtype, const void* coords, long long natoms, const double* box_host,
                                        double cutoff, const int32_t* sorted_atom_indices,
                                        const int32_t* cluster_exclusions, const uint32_t* bitmask_exclusions,
                                        long long num_exclusion_blocks, const int32_t* interacting_clusters,
                                        const int32_t* interacting_atoms, long long num_rows, int32_t* pairs_out,
                                        long long pair_capacity, int32_t* npairs_out, void* workspace,
                                        size_t workspace_bytes, hipStream_t stream) {
  CpGeom g{};
  int rc = make_geom(box_host, natoms, cutoff, cutoff, g);
  if (rc) return rc;
  const Entries en{(const int*)sorted_atom_indices, (const int*)cluster_exclusions, num_exclusion_blocks,
                   (const unsigned*)bitmask_exclusions, (const int*)interacting_clusters,
                   (const int*)interacting_atoms, num_rows};
  if ((rc = check_entries(en, "tff_cluster_pairs_decode"))) return rc;
  if (!coords || !npairs_out || pair_capacity < 0 || (pair_capacity > 0 && !pairs_out) ||
      (dtype != TFF_F32 && dtype != TFF_F64)) {
    tff_set_error("tff_cluster_pairs_decode: invalid arguments");
    return TFF_ERR_ARG;
  }
  const long long total = num_exclusion_blocks + num_rows;
  const TravWs w = carve_trav(workspace, natoms, total);
  if (!workspace || workspace_bytes < w.bytes) {
    tff_set_error("tff_cluster_pairs_decode: workspace too small (%zu < %zu)", workspace_bytes, w.bytes);
    return TFF_ERR_WORKSPACE;
  }
  const unsigned gt = grid_for(1, total);
  if (dtype == TFF_F64) {
    hipLaunchKernelGGL((k_cp_decode<double, TR_COUNT>), dim3(gt), dim3(CPB), 0, stream, (const double*)coords, g, en,
                       w.cnt, (const long long*)nullptr, 0LL, (int*)nullptr);
  } else {
    hipLaunchKernelGGL((k_cp_decode<float, TR_COUNT>), dim3(gt), dim3(CPB), 0, stream, (const float*)coords, g, en,
                       w.cnt, (const long long*)nullptr, 0LL, (int*)nullptr);
  }
  hipLaunchKernelGGL((k_cp_scan<int, long long>), dim3(1), dim3(1024), 0, stream, w.cnt, total, w.off);
  if (pair_capacity > 0) {
    if (dtype == TFF_F64)
      hipLaunchKernelGGL((k_cp_decode<double, TR_WRITE>), dim3(gt), dim3(CPB), 0, stream, (const double*)coords, g,
                         en, w.cnt, (const long long*)w.off, pair_capacity, (int*)pairs_out);
    else
      hipLaunchKernelGGL((k_cp_decode<float, TR_WRITE>), dim3(gt), dim3(CPB), 0, stream, (const float*)coords, g, en,
                         w.cnt, (const long long*)w.off, pair_capacity, (int*)pairs_out);
  }
  long long pb = (pair_capacity + 255) / 256;
  if (pb > 4096) pb = 4096;
  if (pb < 1) pb = 1;
  hipLaunchKernelGGL(k_cp_npairs, dim3((unsigned)pb), dim3(256), 0, stream, (const long long*)w.off, total,
                     pair_capacity, (int*)npairs_out, (int*)pairs_out);
  TFF_CHECK_LAUNCH("cluster pairs decode");
  return TFF_OK;
}

extern "C" int tff_cluster_pairs_nonbonded(int dtype, const void* coords, long long natoms, const double* box_host,
                                           const void* charges, const void* sigma, const void* epsilon,
                                           double coul_constant, double cutoff, int do_shift,
                                           const int32_t* sorted_atom_indices, const int32_t* cluster_exclusions,
                                           const uint32_t* bitmask_exclusions, long long num_exclusion_blocks,
                                           const int32_t* interacting_clusters, const int32_t* interacting_atoms,
                                           long long num_rows, double scale, void* energy_out, void* grad_coords,
                                           void* grad_charges, void* grad_sigma, void* grad_epsilon,
                                           void* forces_accum, void* workspace, size_t workspace_bytes,
                                           hipStream_t stream) {
  CpGeom g{};
  int rc = make_geom(box_host, natoms, cutoff, cutoff, g);
  if (rc) return rc;
  const Entries en{(const int*)sorted_atom_indices, (const int*)cluster_exclusions, num_exclusion_blocks,
                   (const unsigned*)bitmask_exclusions, (const int*)interacting_clusters,
                   (const int*)interacting_atoms, num_rows};
  if ((rc = check_entries(en, "tff_cluster_pairs_nonbonded"))) return rc;
  if (!coords || !charges || !sigma || !epsilon || (dtype != TFF_F32 && dtype != TFF_F64)) {
    tff_set_error("tff_cluster_pairs_nonbonded: invalid arguments");
    return TFF_ERR_ARG;
  }
  const long long total = num_exclusion_blocks + num_rows;
  const TravWs w = carve_trav(workspace, natoms, total);
  if (!workspace || workspace_bytes < w.bytes) {
    tff_set_error("tff_cluster_pairs_nonbonded: workspace too small (%zu < %zu)", workspace_bytes, w.bytes);
    return TFF_ERR_WORKSPACE;
  }
  const bool need_xyz = grad_coords || forces_accum;
  auto zero = [&](double* p, long long n) {
    long long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_cp_zero_d, dim3((unsigned)b), dim3(256), 0, stream, p, n);
  };
  if (need_xyz) zero(w.g_xyz, 3 * natoms);
  if (grad_charges) zero(w.g_q, natoms);
  if (grad_sigma) zero(w.g_sig, natoms);
  if (grad_epsilon) zero(w.g_eps, natoms);
  const unsigned gt = grid_for(1, total);
  double* gx = need_xyz ? w.g_xyz : nullptr;
  double* gq = grad_charges ? w.g_q : nullptr;
  double* gs = grad_sigma ? w.g_sig : nullptr;
  double* ge = grad_epsilon ? w.g_eps : nullptr;
  auto cast = [&](auto tag, const double* s, long long n, double sc, int acc, void* d) {
    using TT = decltype(tag);
    long long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_cp_cast<TT>, dim3((unsigned)b), dim3(256), 0, stream, s, n, sc, acc, (TT*)d);
  };
  auto finish = [&](auto tag) {
    using TT = decltype(tag);
    if (energy_out)
      hipLaunchKernelGGL(k_cp_energy_final<TT>, dim3(1), dim3(256), 0, stream, (const double*)w.e_part, (int)gt,
                         scale, (TT*)energy_out);
    if (grad_coords) cast(tag, w.g_xyz, 3 * natoms, scale, 0, grad_coords);
    if (forces_accum) cast(tag, w.g_xyz, 3 * natoms, -scale, 1, forces_accum);
    if (grad_charges) cast(tag, w.g_q, natoms, scale, 0, grad_charges);
    if (grad_sigma) cast(tag, w.g_sig, natoms, scale, 0, grad_sigma);
    if (grad_epsilon) cast(tag, w.g_eps, natoms, scale, 0, grad_epsilon);
  };
  if (dtype == TFF_F64) {
    hipLaunchKernelGGL(k_cp_nonbonded<double>, dim3(gt), dim3(CPB), 0, stream, (const double*)coords, g, en,
                       (const double*)charges, (const double*)sigma, (const double*)epsilon, coul_constant, do_shift,
                       w.e_part, gx, gq, gs, ge);
    finish(double{});
  } else {
    hipLaunchKernelGGL(k_cp_nonbonded<float>, dim3(gt), dim3(CPB), 0, stream, (const float*)coords, g, en,
                       (const float*)charges, (const float*)sigma, (const float*)epsilon, coul_constant, do_shift,
                       w.e_part, gx, gq, gs, ge);
    finish(float{});
  }
  TFF_CHECK_LAUNCH("cluster pairs nonbonded");
  return TFF_OK;
}
