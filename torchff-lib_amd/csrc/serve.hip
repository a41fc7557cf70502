// Batched energy + forces of many small periodic water systems in one fixed sequence of launches
// (the serving path).
//
// A batch lives in one input block: a head with the totals, one SysDesc record per system
// (atom range, grid offset, image range, FFT pass layouts and plans, influence-function
// parameters) and the atom records (x, y, z, q, sigma, epsilon, molecule, 0 as float64).  The
// host fills the same layout in pinned memory and uploads it with one copy; every kernel is
// launched with a grid sized for the capacities the workspace was made for and reads the batch
// from the block, so the launch sequence can be captured once into a HIP graph and replayed for
// any batch that fits.  Results go to one output block (energies of every system, then the
// forces of every atom), read back with one copy.
//
// Real space (any box, also boxes shorter than twice the cutoff): the full double-counted sum
// over all atoms j and periodic images n within the cutoff,
//   E_real = 1/2 sum_i sum_{j, n: (j, n) != (i, 0), |d_ijn| <= rc} phi_ijn,
//   d_ijn = minimg(r_i - r_j) - n L,
// phi = k q_i q_j erfc(a r)/r + 4 eps_ij (s^12 - s^6) (Lorentz-Berthelot), and on the nearest
// image (n = 0) of a pair of one molecule k q_i q_j (erfc(a r) - 1)/r instead (the exclusion
// correction, torchff/multipoles.py:256-257).  A 64-atom block of i atoms is one task; its
// units (64-atom chunk of j atoms x image) are spread over the system's splits (workgroups of
// RW waves, as many as give each wave about UNITS_PER_WAVE units), each unit staged through the
// wave's LDS and evaluated branch-free; each workgroup sums its waves in a fixed order, and the
// gather kernel adds the splits' partial sums in a fixed order.  Persistent workgroups drain
// the work items of the batch.
// Reciprocal space: smooth PME per system (order P spread with int64 fixed-point atomics into
// the system's grid; 2D transforms of the x-slabs in LDS; the x-line pass of the shared FFT code
// with the influence function; inverse slab transforms; gather).  The forward slab transform
// zeroes the charge grid as it reads it, so the grid is clean for the next batch without a
// clearing launch.  Orthorhombic boxes.

#include "pair_math.h"
#include "pme_core.h"
#include "tff_common.h"
#include "torchff_hip.h"

namespace {

using namespace tffpme;

constexpr double SQ_SCALE = 1099511627776.0;  // 2^40: charge grids
constexpr int SERVE_LINES = 16;               // lines per FFT tile
constexpr int FFT_BLOCKS = 512;               // workgroups of an FFT pass (each loops over tiles)
constexpr int REC = 8;                        // float64 values per atom record
constexpr int STAGE_BLOCKS = 128;             // workgroups of the table staging kernel
constexpr int RW = 8;                         // waves per workgroup
constexpr int MAX_SPLITS = 32;                // real-space workgroups per 64-atom task, at most
constexpr int UNITS_PER_WAVE = 1;             // target (chunk x image) units per wave
constexpr int PERSIST_BLOCKS = 1024;          // workgroups of the real-space + spread launch
constexpr size_t HEAD_BYTES = 256;            // the head, padded: records start 256 bytes in

struct SysDesc {
  int atom_off, natoms;       // atom records atom_off .. atom_off + natoms - 1
  int grid_off;               // first point of the system's grids (K0 K1 K2 points)
  int task_off, ntask;        // 64-atom tasks
  int work_off, nwork, splits;  // real-space work items: ntask x splits, split fastest
  int slab_off;               // first x-slab of the system among the batch's slabs (K0 each)
  int tile_off[3], ntile[3];  // FFT tiles of the z, y and x passes
  int mod_off[3];             // B-spline moduli of the three axes in the moduli table
  int nimg[3];                // image range -nimg..nimg per axis around the minimum image
  double L[3], invL[3];
  double cutoff, alpha;
  PassArgs pa[3];             // z, y, x pass layouts
  FftPlan plan[3];
  ConvArgs cv;
};

struct BatchHead {
  int nsys, natoms, ntask, ntile[3];
  int atom_end;  // one past the highest atom record used
  int nwork;     // real-space work items
  int nslab;     // x-slabs (summed K0)
};
static_assert(sizeof(BatchHead) <= HEAD_BYTES, "batch head");

struct ServeCaps {
  int max_sys, max_atoms, max_k;
  long long max_grid;
  int max_tasks() const { return max_atoms / TFF_WAVE + max_sys; }
  long long max_work() const { return (long long)MAX_SPLITS * max_tasks(); }
  // an FFT pass of a K^3 grid has at most K^2 / 16 + 2 K + 1 tiles of 16 lines
  long long max_tiles() const { return max_grid / 16 + (long long)max_sys * (2 * max_k + 1); }
};

size_t atoms_offset(int max_sys) { return tff_align(HEAD_BYTES + sizeof(SysDesc) * (size_t)max_sys); }
size_t input_bytes(const ServeCaps& c) { return atoms_offset(c.max_sys) + (size_t)REC * 8 * c.max_atoms; }

__device__ __forceinline__ int find_sys(const SysDesc* __restrict__ d, int nsys, int key, int which) {
  // last system whose offset (which: 0 work items, 1 tasks, 2..4 tiles of pass which - 2, 5 slabs)
  // is <= key
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int off = which == 0 ? d[mid].work_off : which == 1 ? d[mid].task_off
                  : which == 5 ? d[mid].slab_off : d[mid].tile_off[which - 2];
    if (off <= key) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// the atom of lane `lane` of task `task`, or -1 (s: its system)
__device__ __forceinline__ int task_atom(const BatchHead* __restrict__ hd, const SysDesc* __restrict__ d, int task,
                                         int lane, int& s) {
  if (task >= hd->ntask) return -1;
  s = find_sys(d, hd->nsys, task, 1);
  const int il = (task - d[s].task_off) * TFF_WAVE + lane;
  return il < d[s].natoms ? d[s].atom_off + il : -1;
}

__device__ __forceinline__ double wrap(double x, double L, double invL) {
  double f = x * invL;
  f -= floor(f);
  return f * L;
}

template <typename T> struct P4 { T x, y, z, w; };
template <typename T> struct L4 {
  T s, e;   // sigma / 2, 2 sqrt(epsilon)
  int m, j; // molecule, atom index in the system (-1: padding)
};

// ---------------------------------------------------------------- real space + spread
// one spread term q M_a M_b M_c in units of 2^-40.  float: rounded to 2^-28 in float arithmetic
// (a 32-bit conversion; |q| < 8 keeps it in range) and scaled by 2^12; double: rounded at 2^-40.
__device__ __forceinline__ long long grid_term(double wab, float mc) {
  const float v = (float)wab * mc * 268435456.0f;
  return (long long)__float2int_rn(v) * 4096LL;
}
__device__ __forceinline__ long long grid_term(double wab, double mc) {
  return __double2ll_rn(wab * mc * SQ_SCALE);
}

// spread of one task's atoms (a workgroup: lane = atom, the RW waves share the P^2 (a, b) stencil
// columns): order-P B-spline weights, int64 fixed-point atomics
template <typename TC, int P>
__device__ __forceinline__ void spread_task(const BatchHead* __restrict__ hd, const SysDesc* __restrict__ d,
                                            const double* __restrict__ atoms, unsigned long long* __restrict__ grid,
                                            int task, int lane, int wv) {
  int s = 0;
  const int i = task_atom(hd, d, task, lane, s);
  if (i < 0) return;
  const SysDesc& sd = d[s];
  const int K0 = sd.cv.K[0], K1 = sd.cv.K[1], K2 = sd.cv.K[2];
  const double* rec = atoms + (size_t)REC * i;
  TC Mv[3][P], dM[3][P];
  int base[3];
  for (int a = 0; a < 3; ++a) {
    double t;
    anchor_axis(rec[a] * sd.invL[a] * (double)sd.cv.K[a], sd.cv.K[a], P, base[a], t);
    spline_weights<TC, P>((TC)t, Mv[a], dM[a]);
  }
  const double qi = rec[3];
  unsigned long long* g = grid + sd.grid_off;
  for (int ab = wv; ab < P * P; ab += RW) {
    const int a = ab / P, b = ab - a * P;
    int gx = base[0] + a, gy = base[1] + b;
    while (gx >= K0) gx -= K0;  // a, b < P: a few steps even when K < P
    while (gy >= K1) gy -= K1;
    TC ma = Mv[0][0], mb = Mv[1][0];
#pragma unroll
    for (int k = 1; k < P; ++k) {
      ma = a == k ? Mv[0][k] : ma;
      mb = b == k ? Mv[1][k] : mb;
    }
    const double wab = qi * (double)ma * (double)mb;
    unsigned long long* row = g + ((long long)gx * K1 + gy) * K2;
    int gz = base[2];
#pragma unroll
    for (int c = 0; c < P; ++c) {
      atomicAdd(&row[gz], (unsigned long long)grid_term(wab, Mv[2][c]));
      gz = gz + 1 == K2 ? 0 : gz + 1;
    }
  }
}

// Real space of one work item (task = 64 i atoms, split sp of the system's splits): the split's
// waves take the units gw, gw + splits RW, ... (gw = sp RW + wave) of the task's nch x nim
// units (chunk fastest); the RW wave sums are added in a fixed order into freal[item].
template <typename TC>
__device__ __forceinline__ void real_item(const BatchHead* __restrict__ hd, const SysDesc* __restrict__ d,
                                          const double* __restrict__ atoms, double coul, int item,
                                          unsigned char* __restrict__ lds, int stage_bytes,
                                          double* __restrict__ freal, double* __restrict__ epart) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int wv = threadIdx.x / TFF_WAVE;
  const int s = find_sys(d, hd->nsys, item, 0);
  const SysDesc& sd = d[s];
  const int local = item - sd.work_off;
  const int task = sd.task_off + local / sd.splits, split = local % sd.splits;
  const int n = sd.natoms, a0 = sd.atom_off;
  const int il = (task - sd.task_off) * TFF_WAVE + lane;
  const bool vi = il < n;
  const TC L[3] = {(TC)sd.L[0], (TC)sd.L[1], (TC)sd.L[2]};
  const TC iL[3] = {(TC)sd.invL[0], (TC)sd.invL[1], (TC)sd.invL[2]};
  const TC rc2 = (TC)(sd.cutoff * sd.cutoff), alpha = (TC)sd.alpha;
  TC xi[3] = {0, 0, 0}, qi = 0, si = 0, ei = 0;
  int mi = -1;
  if (vi) {
    const double* rec = atoms + (size_t)REC * (a0 + il);
    for (int a = 0; a < 3; ++a) xi[a] = (TC)wrap(rec[a], sd.L[a], sd.invL[a]);
    qi = (TC)(rec[3] * coul);
    si = (TC)(0.5 * rec[4]);
    ei = (TC)(2.0 * sqrt(rec[5] > 0 ? rec[5] : 0.0));
    mi = (int)rec[6];
  }
  P4<TC>* s_p = reinterpret_cast<P4<TC>*>(lds + wv * stage_bytes);
  L4<TC>* s_l = reinterpret_cast<L4<TC>*>(s_p + TFF_WAVE);
  const int nch = (n + TFF_WAVE - 1) / TFF_WAVE;
  const int wx = 2 * sd.nimg[0] + 1, wy = 2 * sd.nimg[1] + 1;
  const int nunits = nch * wx * wy * (2 * sd.nimg[2] + 1);
  const TC c2 = (TC)1.1283791670955126;  // 2 / sqrt(pi)
  double fx = 0, fy = 0, fz = 0, e_acc = 0;
  for (int u = split * RW + wv; u < nunits; u += sd.splits * RW) {
    const int c = u % nch, m = u / nch;
    const int ix = m % wx - sd.nimg[0], iy = (m / wx) % wy - sd.nimg[1], iz = m / (wx * wy) - sd.nimg[2];
    const bool zero = ix == 0 && iy == 0 && iz == 0;
    const int j0 = c * TFF_WAVE;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    {
      // lanes past the system's end stage a zero-charge, zero-epsilon atom that never counts
      const bool vj = j0 + lane < n;
      const double* rec = atoms + (size_t)REC * (a0 + (vj ? j0 + lane : 0));
      s_p[lane] = {(TC)wrap(rec[0], sd.L[0], sd.invL[0]), (TC)wrap(rec[1], sd.L[1], sd.invL[1]),
                   (TC)wrap(rec[2], sd.L[2], sd.invL[2]), vj ? (TC)rec[3] : (TC)0};
      s_l[lane] = {(TC)(0.5 * rec[4]), vj ? (TC)(2.0 * sqrt(rec[5] > 0 ? rec[5] : 0.0)) : (TC)0,
                   vj ? (int)rec[6] : -2, vj ? j0 + lane : -1};
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const TC sx = (TC)ix * L[0], sy = (TC)iy * L[1], sz = (TC)iz * L[2];
    // the own-atom pair (j = i on the zero image) and padded j never count
    const int self = zero ? il : -3;
    TC ex = 0, ey = 0, ez = 0, ee = 0;
#pragma unroll 4
    for (int k = 0; k < TFF_WAVE; ++k) {
      const P4<TC> pj = s_p[k];
      const L4<TC> lj = s_l[k];
      TC dx = xi[0] - pj.x, dy = xi[1] - pj.y, dz = xi[2] - pj.z;
      dx -= L[0] * rint(dx * iL[0]);
      dy -= L[1] * rint(dy * iL[1]);
      dz -= L[2] * rint(dz * iL[2]);
      dx -= sx;
      dy -= sy;
      dz -= sz;
      const TC r2 = dx * dx + dy * dy + dz * dz;
      const bool count = vi && r2 <= rc2 && lj.j != self && lj.j >= 0;
      if (__builtin_amdgcn_ballot_w64(count) == 0) continue;  // uniform: no lane has j in range
      const bool excl = count && zero && lj.m == mi;
      // pairs that do not count are moved out to r^2 = 1e30, where every term vanishes
      const TC r2c = count ? r2 : (TC)1e30;
      const TC rinv = rsqrt_t(r2c);
      const TC rinv2 = rinv * rinv;
      const TC ua = alpha * r2c * rinv;
      TC ec, em;
      erfc_exp(ua, ec, em);
      ec = excl ? ec - (TC)1 : ec;
      const TC qq = qi * pj.w;
      TC fs = qq * (ec + c2 * ua * em) * rinv * rinv2;
      ee += qq * ec * rinv;
      const TC ep4 = excl ? (TC)0 : ei * lj.e;
      const TC sg = si + lj.s;
      const TC s2 = sg * sg * rinv2;
      const TC s6 = s2 * s2 * s2;
      ee += ep4 * s6 * (s6 - (TC)1);
      fs += ep4 * s6 * ((TC)12 * s6 - (TC)6) * rinv2;
      ex += fs * dx;
      ey += fs * dy;
      ez += fs * dz;
    }
    fx += (double)ex;
    fy += (double)ey;
    fz += (double)ez;
    e_acc += (double)ee;
  }
  // the RW partial sums of each atom, pairwise in a fixed order
  __syncthreads();
  double* red = reinterpret_cast<double*>(lds);
  for (int h = RW / 2; h >= 1; h /= 2) {
    if (wv >= h && wv < 2 * h) {
      double* r = red + (size_t)(wv - h) * 4 * TFF_WAVE;
      r[lane] = fx;
      r[TFF_WAVE + lane] = fy;
      r[2 * TFF_WAVE + lane] = fz;
      r[3 * TFF_WAVE + lane] = e_acc;
    }
    __syncthreads();
    if (wv < h) {
      const double* r = red + (size_t)wv * 4 * TFF_WAVE;
      fx += r[lane];
      fy += r[TFF_WAVE + lane];
      fz += r[2 * TFF_WAVE + lane];
      e_acc += r[3 * TFF_WAVE + lane];
    }
    __syncthreads();
  }
  if (wv == 0) {
    const size_t slot = (size_t)item * TFF_WAVE + lane;
    freal[3 * slot + 0] = fx;
    freal[3 * slot + 1] = fy;
    freal[3 * slot + 2] = fz;
    const double tot = tff_wave_sum(0.5 * e_acc);  // every pair is visited from both ends
    if (lane == 0) epart[item] = tot;
  }
}

// Persistent workgroups over the items [0, ntask) (spread of one task each) and
// [ntask, ntask + nwork) (real space, one work item each).
template <typename TC, int P>
__global__ __launch_bounds__(RW * TFF_WAVE) void k_serve_real_spread(const BatchHead* __restrict__ hd,
                                                                     const SysDesc* __restrict__ d,
                                                                     const double* __restrict__ atoms, double coul,
                                                                     unsigned long long* __restrict__ grid,
                                                                     double* __restrict__ freal,
                                                                     double* __restrict__ epart) {
  // per wave: staged j chunk (position + charge; sigma / 2, 2 sqrt(eps), molecule, index);
  // reused by the final reduction as RW / 2 x 4 x 64 doubles
  constexpr int STAGE = TFF_WAVE * (sizeof(P4<TC>) + sizeof(L4<TC>));
  constexpr int RED = RW / 2 * 4 * TFF_WAVE * sizeof(double);
  __shared__ __align__(16) unsigned char lds[RW * STAGE > RED ? RW * STAGE : RED];
  const int nsp = hd->ntask;
  const int nitems = nsp + hd->nwork;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {  // uniform over the workgroup
    if (it < nsp) {
      spread_task<TC, P>(hd, d, atoms, grid, it, threadIdx.x & (TFF_WAVE - 1), threadIdx.x / TFF_WAVE);
    } else {
      real_item<TC>(hd, d, atoms, coul, it - nsp, lds, STAGE, freal, epart);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------- FFT passes
// one pass of the line FFTs over every system of the batch: workgroup -> (system, tile), looping
template <typename T, int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_serve_fft(const BatchHead* __restrict__ hd,
                                                           const SysDesc* __restrict__ d,
                                                           unsigned long long* __restrict__ grid,
                                                           cx<T>* __restrict__ work, T* __restrict__ phi,
                                                           const T* __restrict__ moduli,
                                                           double* __restrict__ kpart) {
  constexpr int PI = MODE == PASS_Z_FWD || MODE == PASS_Z_INV ? 0 : (MODE == PASS_X_CONV ? 2 : 1);
  const int nt = hd->ntile[PI];
  for (int b = blockIdx.x; b < nt; b += gridDim.x) {  // uniform over the workgroup
    const int s = find_sys(d, hd->nsys, b, 2 + PI);
    const SysDesc& sd = d[s];
    const int tile = b - sd.tile_off[PI];
    const void* in;
    void* out;
    if (MODE == PASS_Z_FWD) in = grid + sd.grid_off; else in = work + sd.grid_off;
    if (MODE == PASS_Z_INV) out = phi + sd.grid_off; else out = work + sd.grid_off;
    fft_pass_tile<T, MODE>(in, out, sd.pa[PI], sd.plan[PI], sd.cv, moduli + sd.mod_off[0],
                           moduli + sd.mod_off[1], moduli + sd.mod_off[2], 0, tile,
                           MODE == PASS_X_CONV ? kpart + b : nullptr);
    __syncthreads();  // the next tile may have another K, moving the LDS buffers
  }
}

// Two-dimensional transforms of the x-slabs (one workgroup per slab, in LDS): forward, z lines
// then y lines, from the int64 fixed-point charge grid (zeroed as it is read) into the complex
// work grid; inverse, y lines then z lines, from the work grid into the real potential grid.
// Lines are padded to a power-of-two count with zeros; the transpose between the two directions
// goes through LDS.
template <typename T, int DIR>
__global__ __launch_bounds__(FFT_THREADS) void k_serve_slab(const BatchHead* __restrict__ hd,
                                                            const SysDesc* __restrict__ d,
                                                            unsigned long long* __restrict__ grid,
                                                            cx<T>* __restrict__ work, T* __restrict__ phi) {
  extern __shared__ unsigned char lds_raw[];
  const int nslab = hd->nslab;
  for (int b = blockIdx.x; b < nslab; b += gridDim.x) {  // uniform over the workgroup
    const SysDesc& sd = d[find_sys(d, hd->nsys, b, 5)];
    const int x = b - sd.slab_off;
    const int K1 = sd.cv.K[1], K2 = sd.cv.K[2];
    int sz = 0, sy = 0;  // lines of the z phase (y values) and of the y phase (z values), as shifts
    while ((1 << sz) < K1) ++sz;
    while ((1 << sy) < K2) ++sy;
    const int LZ = K2 + 1, LY = K1 + 1;  // LDS line strides of the z and y phases
    cx<T>* twz = reinterpret_cast<cx<T>*>(lds_raw);
    cx<T>* twy = twz + K2;
    cx<T>* bufA = twy + K1;
    const int span = ((1 << sz) * LZ > (1 << sy) * LY) ? (1 << sz) * LZ : (1 << sy) * LY;
    cx<T>* bufB = bufA + span;
    fill_twiddles(twz, K2);
    fill_twiddles(twy, K1);
    const long long plane = (long long)x * K1 * K2 + sd.grid_off;
    const float inv_k2 = 1.0f / (float)K2, inv_k1 = 1.0f / (float)K1;
    if (DIR < 0) {
      long long* gi = reinterpret_cast<long long*>(grid);
      for (int idx = threadIdx.x; idx < (K2 << sz); idx += blockDim.x) {
        const int y = fdiv(idx, inv_k2), z = idx - y * K2;
        cx<T> v = {(T)0, (T)0};
        if (y < K1) {
          const long long g = plane + (long long)y * K2 + z;
          v.x = (T)((double)gi[g] * sd.cv.grid_inv_scale);
          gi[g] = 0;  // clean for the next batch; read by this thread only
        }
        bufA[y * LZ + z] = v;
      }
      __syncthreads();
      cx<T>* r = fft_lines<T, -1>(bufA, bufB, twz, K2, LZ, sz, sd.plan[0]);
      cx<T>* o = r == bufA ? bufB : bufA;
      for (int idx = threadIdx.x; idx < (K1 << sy); idx += blockDim.x) {
        const int z = fdiv(idx, inv_k1), y = idx - z * K1;
        o[z * LY + y] = z < K2 ? r[y * LZ + z] : cx<T>{(T)0, (T)0};
      }
      __syncthreads();
      cx<T>* r2 = fft_lines<T, -1>(o, r, twy, K1, LY, sy, sd.plan[1]);
      for (int idx = threadIdx.x; idx < K1 * K2; idx += blockDim.x) {
        const int y = fdiv(idx, inv_k2), z = idx - y * K2;
        work[plane + (long long)y * K2 + z] = r2[z * LY + y];
      }
    } else {
      for (int idx = threadIdx.x; idx < (K1 << sy); idx += blockDim.x) {
        const int z = fdiv(idx, inv_k1), y = idx - z * K1;
        bufA[z * LY + y] = z < K2 ? work[plane + (long long)y * K2 + z] : cx<T>{(T)0, (T)0};
      }
      __syncthreads();
      cx<T>* r = fft_lines<T, 1>(bufA, bufB, twy, K1, LY, sy, sd.plan[1]);
      cx<T>* o = r == bufA ? bufB : bufA;
      for (int idx = threadIdx.x; idx < (K2 << sz); idx += blockDim.x) {
        const int y = fdiv(idx, inv_k2), z = idx - y * K2;
        o[y * LZ + z] = y < K1 ? r[z * LY + y] : cx<T>{(T)0, (T)0};
      }
      __syncthreads();
      cx<T>* r2 = fft_lines<T, 1>(o, r, twz, K2, LZ, sz, sd.plan[0]);
      for (int idx = threadIdx.x; idx < K1 * K2; idx += blockDim.x) {
        const int y = fdiv(idx, inv_k2), z = idx - y * K2;
        phi[plane + (long long)y * K2 + z] = r2[y * LZ + z].x;
      }
    }
    __syncthreads();  // the next slab may have other K, moving the LDS buffers
  }
}

// ---------------------------------------------------------------- gather + results
// Persistent workgroups over the items [0, ntask): reciprocal-space force of one task's atoms by
// the gather (lane = atom, the RW waves share the P^2 (a, b) stencil columns and are added in a
// fixed order), plus the task's real-space partial forces (one per split, in order), into the
// output; and [ntask, ntask + ceil(nsys / RW)): one system per wave, its energies [real space
// (LJ + Coulomb incl. the exclusion correction), 0, reciprocal, self, total].
template <typename TC, int P>
__global__ __launch_bounds__(RW * TFF_WAVE) void k_serve_gather(const BatchHead* __restrict__ hd,
                                                                const SysDesc* __restrict__ d,
                                                                const double* __restrict__ atoms,
                                                                const TC* __restrict__ phi, double coul,
                                                                const double* __restrict__ freal,
                                                                const double* __restrict__ epart,
                                                                const double* __restrict__ kpart, int max_sys,
                                                                double* __restrict__ out) {
  __shared__ double red[RW][3][TFF_WAVE];
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int wv = threadIdx.x / TFF_WAVE;
  const int ntask = hd->ntask;
  const int nitems = ntask + (hd->nsys + RW - 1) / RW;
  for (int it = blockIdx.x; it < nitems; it += gridDim.x) {  // uniform over the workgroup
    if (it >= ntask) {
      const int s = (it - ntask) * RW + wv;
      if (s < hd->nsys) {
        const SysDesc& sd = d[s];
        double er = 0, ek = 0, q2 = 0;
        for (int t = lane; t < sd.nwork; t += TFF_WAVE) er += epart[sd.work_off + t];
        for (int t = lane; t < sd.ntile[2]; t += TFF_WAVE) ek += kpart[sd.tile_off[2] + t];
        for (int t = lane; t < sd.natoms; t += TFF_WAVE) {
          const double v = atoms[(size_t)REC * (sd.atom_off + t) + 3];
          q2 += v * v;
        }
        er = tff_wave_sum(er);
        ek = tff_wave_sum(ek);
        q2 = tff_wave_sum(q2);
        if (lane == 0) {
          // out may be pinned host memory: stores only, no read-back
          double* o = out + 5 * (size_t)s;
          const double erec = coul * ek, eself = -coul * sd.alpha * 0.5641895835477563 * q2;
          o[0] = er;
          o[1] = 0.0;
          o[2] = erec;
          o[3] = eself;
          o[4] = er + erec + eself;
        }
      }
      continue;  // uniform: every wave of the workgroup takes this branch
    }
    const int task = it;
    const int s = find_sys(d, hd->nsys, task, 1);
    const SysDesc& sd = d[s];
    const int il = (task - sd.task_off) * TFF_WAVE + lane;
    const bool vi = il < sd.natoms;
    const int i = sd.atom_off + (vi ? il : 0);
    const int K0 = sd.cv.K[0], K1 = sd.cv.K[1], K2 = sd.cv.K[2];
    const double* rec = atoms + (size_t)REC * i;
    TC Mv[3][P], dM[3][P];
    int base[3];
    for (int a = 0; a < 3; ++a) {
      double t;
      anchor_axis(rec[a] * sd.invL[a] * (double)sd.cv.K[a], sd.cv.K[a], P, base[a], t);
      spline_weights<TC, P>((TC)t, Mv[a], dM[a]);
    }
    const TC* g = phi + sd.grid_off;
    double du0 = 0, du1 = 0, du2 = 0;
    for (int ab = wv; ab < P * P; ab += RW) {
      const int a = ab / P, b = ab - a * P;
      int gx = base[0] + a, gy = base[1] + b;
      while (gx >= K0) gx -= K0;  // a, b < P: a few steps even when K < P
      while (gy >= K1) gy -= K1;
      TC ma = Mv[0][0], mb = Mv[1][0], da = dM[0][0], db = dM[1][0];
#pragma unroll
      for (int k = 1; k < P; ++k) {
        ma = a == k ? Mv[0][k] : ma;
        mb = b == k ? Mv[1][k] : mb;
        da = a == k ? dM[0][k] : da;
        db = b == k ? dM[1][k] : db;
      }
      const TC* row = g + ((long long)gx * K1 + gy) * K2;
      int gz = base[2];
      double s0 = 0, s1 = 0;
#pragma unroll
      for (int c = 0; c < P; ++c) {
        const double v = (double)row[gz];
        s0 += v * (double)Mv[2][c];
        s1 += v * (double)dM[2][c];
        gz = gz + 1 == K2 ? 0 : gz + 1;
      }
      du0 += (double)da * (double)mb * s0;
      du1 += (double)ma * (double)db * s0;
      du2 += (double)ma * (double)mb * s1;
    }
    red[wv][0][lane] = du0;
    red[wv][1][lane] = du1;
    red[wv][2][lane] = du2;
    __syncthreads();
    if (wv == 0 && vi) {
      du0 = du1 = du2 = 0;
      for (int w = 0; w < RW; ++w) {
        du0 += red[w][0][lane];
        du1 += red[w][1][lane];
        du2 += red[w][2][lane];
      }
      // force = -k q grad phi; d/ds of the weights is -d/dt, d s_a / d r_a = K_a / L_a
      const double kq = coul * rec[3];
      double f[3] = {kq * du0 * K0 * sd.invL[0], kq * du1 * K1 * sd.invL[1], kq * du2 * K2 * sd.invL[2]};
      const int item0 = sd.work_off + (task - sd.task_off) * sd.splits;
      for (int sp = 0; sp < sd.splits; ++sp) {
        const size_t slot = (size_t)(item0 + sp) * TFF_WAVE + lane;
        for (int c = 0; c < 3; ++c) f[c] += freal[3 * slot + c];
      }
      double* fo = out + 5 * (size_t)max_sys + 3 * (size_t)i;
      for (int c = 0; c < 3; ++c) fo[c] = f[c];
    }
    __syncthreads();
  }
}

// The batch table (head, system records, atom records up to the highest one used) copied from
// pinned host memory into the workspace.  The size is read from the table's head when the kernel
// runs, so a captured graph that holds this launch serves batches of any size.  Sizes are
// multiples of 16 bytes (the records start 256-byte aligned, atom records are 64 bytes).
__global__ __launch_bounds__(256) void k_serve_stage(const uint4* __restrict__ table, uint4* __restrict__ input,
                                                     long long rec_bytes, int max_atoms) {
  __shared__ long long s_n;
  if (threadIdx.x == 0) {
    int ae = ((const BatchHead*)table)->atom_end;
    ae = ae < 0 ? 0 : (ae > max_atoms ? max_atoms : ae);
    s_n = (rec_bytes + (long long)REC * 8 * ae) / 16;
  }
  __syncthreads();
  const long long n = s_n;
  for (long long k = blockIdx.x * 256LL + threadIdx.x; k < n; k += (long long)gridDim.x * 256) input[k] = table[k];
}

__global__ void k_serve_zero(unsigned long long* __restrict__ p, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// ---------------------------------------------------------------- host side
struct ServeWs {
  unsigned char* input;  // head, records, atom records: the host table's layout
  BatchHead* head;
  SysDesc* desc;
  double* atoms;
  unsigned long long* grid;
  void* work;
  void* phi;
  double *freal, *epart, *kpart;
  size_t bytes;
};

ServeWs serve_carve(void* base, const ServeCaps& c, size_t tc) {
  ServeWs w{};
  size_t off = 0;
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* r = p ? p + off : nullptr;
    off += tff_align(b > 0 ? b : 1);
    return r;
  };
  w.input = (unsigned char*)take(input_bytes(c));
  if (w.input) {
    w.head = (BatchHead*)w.input;
    w.desc = (SysDesc*)(w.input + HEAD_BYTES);
    w.atoms = (double*)(w.input + atoms_offset(c.max_sys));
  }
  w.grid = (unsigned long long*)take(8 * (size_t)c.max_grid);
  w.work = take(2 * tc * (size_t)c.max_grid);
  w.phi = take(tc * (size_t)c.max_grid);
  w.freal = (double*)take(24 * TFF_WAVE * (size_t)c.max_work());
  w.epart = (double*)take(8 * (size_t)c.max_work());
  w.kpart = (double*)take(8 * (size_t)c.max_tiles());
  w.bytes = off;
  return w;
}

size_t serve_lds(int max_k, size_t celem) {
  return (size_t)max_k * celem + 2 * (size_t)SERVE_LINES * (max_k + 1) * celem;
}

// LDS of a slab transform with edges up to max_k: two sets of twiddles and two buffers of
// (power-of-two lines) x (edge + 1); the largest over edges <= max_k
size_t slab_lds(int max_k, size_t celem) {
  size_t best = 0;
  for (int k = 1; k <= max_k; ++k) {
    size_t p2 = 1;
    while ((int)p2 < k) p2 *= 2;
    const size_t b = 2 * (size_t)k * celem + 2 * p2 * (size_t)(k + 1) * celem;
    best = b > best ? b : best;
  }
  return best;
}

int check_caps(const ServeCaps& c, int compute_dtype) {
  if (c.max_sys < 1 || c.max_atoms < 1 || c.max_k < 1 || c.max_grid < 1 ||
      (compute_dtype != TFF_F32 && compute_dtype != TFF_F64)) {
    tff_set_error("serve: capacities must be positive, compute dtype 0 or 1");
    return TFF_ERR_ARG;
  }
  if (c.max_atoms > (1 << 26) || c.max_sys > (1 << 20) || c.max_grid > (1LL << 31) - 1) {
    tff_set_error("serve: capacities above 2^26 atoms, 2^20 systems or 2^31 grid points");
    return TFF_ERR_ARG;
  }
  const size_t ce = 2 * (compute_dtype == TFF_F64 ? 8 : 4);
  const size_t lds = serve_lds(c.max_k, ce), lds2 = slab_lds(c.max_k, ce);
  if (lds > 64 * 1024 || lds2 > 64 * 1024) {
    tff_set_error("serve: max_k %d needs %zu bytes of LDS per FFT workgroup (64 KiB limit)", c.max_k,
                  lds > lds2 ? lds : lds2);
    return TFF_ERR_UNSUPPORTED;
  }
  return TFF_OK;
}

unsigned ceil_div(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

template <typename TC>
int run_serve(const ServeWs& w, const ServeCaps& c, double coul, int order, const TC* moduli, double* out,
              hipStream_t st) {
  const int mt = c.max_tasks();
  const long long items = mt + c.max_work();
  const unsigned nrs = (unsigned)(items < PERSIST_BLOCKS ? items : PERSIST_BLOCKS);
  const unsigned nfft = (unsigned)(c.max_tiles() < FFT_BLOCKS ? c.max_tiles() : FFT_BLOCKS);
  // slabs: at most max_grid / 1 (K0 per system, each slab at least one point)
  const unsigned nslab = (unsigned)(c.max_grid < FFT_BLOCKS ? c.max_grid : FFT_BLOCKS);
  const long long gitems = mt + (long long)ceil_div(c.max_sys, RW);
  const unsigned ngt = (unsigned)(gitems < PERSIST_BLOCKS ? gitems : PERSIST_BLOCKS);
  const size_t lds = serve_lds(c.max_k, 2 * sizeof(TC)), lds2 = slab_lds(c.max_k, 2 * sizeof(TC));
#define TFF_SERVE(P)                                                                                            \
  {                                                                                                             \
    hipLaunchKernelGGL((k_serve_real_spread<TC, P>), dim3(nrs), dim3(RW * TFF_WAVE), 0, st, w.head, w.desc,     \
                       w.atoms, coul, w.grid, w.freal, w.epart);                                                \
    hipLaunchKernelGGL((k_serve_slab<TC, -1>), dim3(nslab), dim3(FFT_THREADS), lds2, st, w.head, w.desc, w.grid, \
                       (cx<TC>*)w.work, (TC*)w.phi);                                                            \
    hipLaunchKernelGGL((k_serve_fft<TC, PASS_X_CONV>), dim3(nfft), dim3(FFT_THREADS), lds, st, w.head, w.desc,   \
                       w.grid, (cx<TC>*)w.work, (TC*)w.phi, moduli, w.kpart);                                   \
    hipLaunchKernelGGL((k_serve_slab<TC, 1>), dim3(nslab), dim3(FFT_THREADS), lds2, st, w.head, w.desc, w.grid,   \
                       (cx<TC>*)w.work, (TC*)w.phi);                                                            \
    hipLaunchKernelGGL((k_serve_gather<TC, P>), dim3(ngt), dim3(RW * TFF_WAVE), 0, st, w.head, w.desc, w.atoms, \
                       (const TC*)w.phi, coul, w.freal, w.epart, w.kpart, c.max_sys, out);                     \
  }
  switch (order) {
    case 3: TFF_SERVE(3); break;
    case 4: TFF_SERVE(4); break;
    case 5: TFF_SERVE(5); break;
    case 6: TFF_SERVE(6); break;
    case 7: TFF_SERVE(7); break;
    case 8: TFF_SERVE(8); break;
    default: tff_set_error("serve: unsupported spline order %d (3..8)", order); return TFF_ERR_UNSUPPORTED;
  }
#undef TFF_SERVE
  TFF_CHECK_LAUNCH("serve energy+forces");
  return TFF_OK;
}

}  // namespace

extern "C" size_t tff_serve_atoms_offset(int max_systems) {
  return atoms_offset(max_systems > 0 ? max_systems : 0);
}

extern "C" size_t tff_serve_table_bytes(int max_systems, int max_atoms) {
  if (max_systems < 1 || max_atoms < 1) return 0;
  return atoms_offset(max_systems) + (size_t)REC * 8 * max_atoms;
}

extern "C" void tff_serve_table_reset(void* table_host) {
  if (table_host) *(BatchHead*)table_host = BatchHead{};
}

extern "C" int tff_serve_table_totals(const void* table_host, long long* totals_host) {
  if (!table_host || !totals_host) {
    tff_set_error("serve: null table");
    return TFF_ERR_ARG;
  }
  const BatchHead* h = (const BatchHead*)table_host;
  long long grid = 0;
  if (h->nsys > 0) {
    const SysDesc& last = ((const SysDesc*)((const char*)table_host + HEAD_BYTES))[h->nsys - 1];
    grid = (long long)last.grid_off + (long long)last.cv.K[0] * last.cv.K[1] * last.cv.K[2];
  }
  const long long t[8] = {h->nsys, h->natoms, h->ntask, h->ntile[0], h->ntile[1], h->ntile[2], grid, h->atom_end};
  for (int k = 0; k < 8; ++k) totals_host[k] = t[k];
  return TFF_OK;
}

// Appends one system to the host table (see the header for the argument meaning).
extern "C" int tff_serve_table_add(void* table_host, int compute_dtype, int max_systems, int max_atoms, int max_k,
                                   long long max_grid_points, int atom_off, int natoms, const double* L,
                                   const int* K, const int* mod_off, double cutoff, double alpha) {
  const ServeCaps c{max_systems, max_atoms, max_k, max_grid_points};
  int rc = check_caps(c, compute_dtype);
  if (rc) return rc;
  if (!table_host || !L || !K || !mod_off || natoms < 1 || atom_off < 0 || !(cutoff > 0) || !(alpha > 0)) {
    tff_set_error("serve: invalid system arguments");
    return TFF_ERR_ARG;
  }
  BatchHead* h = (BatchHead*)table_host;
  if (h->nsys < 0 || h->nsys > max_systems || h->natoms < 0 || h->ntask < 0 || h->nwork < 0 || h->nslab < 0 ||
      h->ntile[0] < 0 || h->ntile[1] < 0 || h->ntile[2] < 0 || h->atom_end < 0 || h->atom_end > max_atoms) {
    tff_set_error("serve: table head holds %d systems / %d atom records for capacity %d / %d "
                  "(a table must be reset before its first system)", h->nsys, h->atom_end, max_systems, max_atoms);
    return TFF_ERR_ARG;
  }
  SysDesc* recs = (SysDesc*)((char*)table_host + HEAD_BYTES);
  SysDesc d{};
  for (int a = 0; a < 3; ++a) {
    if (!(L[a] > 0) || K[a] < 1 || K[a] > max_k || mod_off[a] < 0) {
      tff_set_error("serve: box edge %g or grid %d outside (0, inf) x [1, max_k = %d]", L[a], K[a], max_k);
      return TFF_ERR_ARG;
    }
    d.L[a] = L[a];
    d.invL[a] = 1.0 / L[a];
    d.mod_off[a] = mod_off[a];
    // the minimum-image difference lies in [-L/2, L/2], so images up to |n| <= 1/2 + rc / L can
    // be in range (0, the minimum image alone, when L >= 2 rc)
    const double ni = floor(0.5 + cutoff / L[a]);
    if (ni > 16) {
      tff_set_error("serve: box edge %g is too short for cutoff %g (more than 16 images)", L[a], cutoff);
      return TFF_ERR_ARG;
    }
    d.nimg[a] = (int)ni;
    if (!make_plan(K[a], d.plan[a])) {
      tff_set_error("serve: PME grid size %d must be 2^a 3^b 5^c", K[a]);
      return TFF_ERR_UNSUPPORTED;
    }
  }
  const SysDesc* prev = h->nsys > 0 ? &recs[h->nsys - 1] : nullptr;
  d.atom_off = atom_off;
  d.natoms = natoms;
  d.grid_off = prev ? prev->grid_off + prev->cv.K[0] * prev->cv.K[1] * prev->cv.K[2] : 0;
  d.task_off = h->ntask;
  d.ntask = (natoms + TFF_WAVE - 1) / TFF_WAVE;
  {
    // enough workgroups per task for about UNITS_PER_WAVE (chunk x image) units per wave
    const long long units = (long long)d.ntask * (2 * d.nimg[0] + 1) * (2 * d.nimg[1] + 1) * (2 * d.nimg[2] + 1);
    const long long sp = (units + RW * UNITS_PER_WAVE - 1) / (RW * UNITS_PER_WAVE);
    d.splits = (int)(sp < 1 ? 1 : sp > MAX_SPLITS ? MAX_SPLITS : sp);
  }
  d.work_off = h->nwork;
  d.nwork = d.ntask * d.splits;
  d.slab_off = h->nslab;
  d.cutoff = cutoff;
  d.alpha = alpha;
  // passes: z (lines along K2), y (lines along K1), x (lines along K0, fused influence function)
  const size_t ce = 2 * (compute_dtype == TFF_F64 ? 8 : 4);
  const PassCfg pc[3] = {pass_cfg(K[0] * K[1], K[2], 1, ce), pass_cfg(K[0], K[1], K[2], ce),
                         pass_cfg(1, K[0], K[1] * K[2], ce)};
  const FftPlan pl[3] = {d.plan[2], d.plan[1], d.plan[0]};
  for (int p = 0; p < 3; ++p) {
    PassArgs pa = pc[p].pa;  // at most SERVE_LINES lines per tile (the LDS size is set for that)
    if (pa.AB > SERVE_LINES) { pa.AB = SERVE_LINES; pa.ab_shift = 4; }
    if (pa.CB > SERVE_LINES) { pa.CB = SERVE_LINES; pa.cb_shift = 4; }
    d.pa[p] = pa;
    d.plan[p] = pl[p];
    d.ntile[p] = ((pa.A + pa.AB - 1) / pa.AB) * ((pa.C + pa.CB - 1) / pa.CB);
    d.tile_off[p] = h->ntile[p];
  }
  for (int a = 0; a < 3; ++a) d.cv.K[a] = K[a];
  for (int k = 0; k < 9; ++k) d.cv.inv[k] = 0.0;
  d.cv.inv[0] = 1.0 / L[0];
  d.cv.inv[4] = 1.0 / L[1];
  d.cv.inv[8] = 1.0 / L[2];
  d.cv.vol = L[0] * L[1] * L[2];
  d.cv.alpha = alpha;
  d.cv.grid_fixed = 1;
  d.cv.grid_inv_scale = 1.0 / SQ_SCALE;
  d.cv.nz = K[2];
  d.cv.half = 0;
  d.cv.clear_in = 1;
  const long long grid_end = (long long)d.grid_off + (long long)K[0] * K[1] * K[2];
  if (h->nsys + 1 > max_systems || (long long)atom_off + natoms > max_atoms || grid_end > max_grid_points ||
      h->ntask + d.ntask > c.max_tasks() || h->nwork + d.nwork > c.max_work() ||
      h->ntile[0] + d.ntile[0] > c.max_tiles() ||
      h->ntile[1] + d.ntile[1] > c.max_tiles() || h->ntile[2] + d.ntile[2] > c.max_tiles())
    return TFF_FULL;
  recs[h->nsys] = d;
  h->nsys += 1;
  h->natoms += natoms;
  h->ntask += d.ntask;
  h->nwork += d.nwork;
  h->nslab += K[0];
  for (int p = 0; p < 3; ++p) h->ntile[p] += d.ntile[p];
  if (atom_off + natoms > h->atom_end) h->atom_end = atom_off + natoms;
  return TFF_OK;
}

extern "C" size_t tff_serve_workspace_bytes(int compute_dtype, int max_systems, int max_atoms, int max_k,
                                            long long max_grid_points) {
  const ServeCaps c{max_systems, max_atoms, max_k, max_grid_points};
  if (check_caps(c, compute_dtype)) return 0;
  return serve_carve(nullptr, c, compute_dtype == TFF_F64 ? 8 : 4).bytes;
}

extern "C" int tff_serve_workspace_init(int compute_dtype, int max_systems, int max_atoms, int max_k,
                                        long long max_grid_points, void* workspace, size_t workspace_bytes,
                                        hipStream_t stream) {
  const ServeCaps c{max_systems, max_atoms, max_k, max_grid_points};
  int rc = check_caps(c, compute_dtype);
  if (rc) return rc;
  const ServeWs w = serve_carve(workspace, c, compute_dtype == TFF_F64 ? 8 : 4);
  if (!workspace || workspace_bytes < w.bytes) {
    tff_set_error("serve: null workspace or workspace too small");
    return TFF_ERR_ARG;
  }
  hipLaunchKernelGGL(k_serve_zero, dim3(1024), dim3(256), 0, stream, (unsigned long long*)workspace,
                     (long long)(w.bytes / 8));
  TFF_CHECK_LAUNCH("serve workspace init");
  return TFF_OK;
}

extern "C" int tff_serve_upload(const void* table_host, int compute_dtype, int max_systems, int max_atoms,
                                int max_k, long long max_grid_points, void* workspace, size_t workspace_bytes,
                                hipStream_t stream) {
  const ServeCaps c{max_systems, max_atoms, max_k, max_grid_points};
  int rc = check_caps(c, compute_dtype);
  if (rc) return rc;
  const ServeWs w = serve_carve(workspace, c, compute_dtype == TFF_F64 ? 8 : 4);
  if (!table_host || !workspace || workspace_bytes < w.bytes) {
    tff_set_error("serve: null table or workspace too small");
    return TFF_ERR_ARG;
  }
  const BatchHead* h = (const BatchHead*)table_host;
  if (h->nsys < 0 || h->nsys > max_systems || h->atom_end < 0 || h->atom_end > max_atoms) {
    tff_set_error("serve: table holds %d systems / %d atom records, capacity %d / %d", h->nsys, h->atom_end,
                  max_systems, max_atoms);
    return TFF_ERR_ARG;
  }
  if ((uintptr_t)table_host % 16 != 0) {
    const size_t bytes = atoms_offset(max_systems) + (size_t)REC * 8 * h->atom_end;
    if (hipMemcpyAsync(w.input, table_host, bytes, hipMemcpyHostToDevice, stream) != hipSuccess) {
      tff_set_error("serve: table upload failed");
      return TFF_ERR_LAUNCH;
    }
    return TFF_OK;
  }
  // a kernel reads the pinned table directly (no copy command; the size comes from the head at
  // run time, so the launch can be captured together with the evaluation)
  hipLaunchKernelGGL(k_serve_stage, dim3(STAGE_BLOCKS), dim3(256), 0, stream, (const uint4*)table_host,
                     (uint4*)w.input, (long long)atoms_offset(max_systems), max_atoms);
  TFF_CHECK_LAUNCH("serve upload");
  return TFF_OK;
}

extern "C" int tff_serve_energy_forces(int compute_dtype, double coul_constant, int order, const void* moduli,
                                       int max_systems, int max_atoms, int max_k, long long max_grid_points,
                                       double* out, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  const ServeCaps c{max_systems, max_atoms, max_k, max_grid_points};
  int rc = check_caps(c, compute_dtype);
  if (rc) return rc;
  const size_t tc = compute_dtype == TFF_F64 ? 8 : 4;
  const ServeWs w = serve_carve(workspace, c, tc);
  if (!workspace || workspace_bytes < w.bytes || !moduli || !out) {
    tff_set_error("serve: invalid arguments or workspace too small (%zu < %zu)", workspace_bytes, w.bytes);
    return TFF_ERR_ARG;
  }
  if (compute_dtype == TFF_F32) return run_serve<float>(w, c, coul_constant, order, (const float*)moduli, out, stream);
  return run_serve<double>(w, c, coul_constant, order, (const double*)moduli, out, stream);
}
