// Batched energy + forces of many small periodic water systems in one fixed sequence of launches
// (the serving path).
//
// A batch is described by device tables written by the host before the launches: one SysDesc
// per system (atom range, grid offset, image range, FFT pass layouts and plans, influence
// function parameters) and a header with the totals.  Every kernel is launched with a grid
// sized for the capacity the workspace was made for and reads the batch from the tables, so
// the same launch sequence can be captured once into a HIP graph and replayed for any batch
// that fits.
//
// Real space (any box, also boxes shorter than twice the cutoff): the full double-counted sum
// over all atoms j and all periodic images n within the cutoff, one wave per 64-atom block of
// i atoms, j atoms staged through LDS:
//   E_real = 1/2 sum_i sum_{j, n: (j, n) != (i, 0), |r_i - r_j - n L| <= rc} phi_ijn
// with phi = k q_i q_j erfc(a r)/r + 4 eps_ij (s^12 - s^6) (Lorentz-Berthelot), and for the
// excluded pairs (same molecule, n = 0) k q_i q_j (erfc(a r) - 1)/r instead (the reference's
// exclusion correction, multipoles.py:256-257).  Each atom's force is summed by its own lane.
// Reciprocal space: smooth PME per system (order P spread with int64 fixed-point atomics into
// the system's grid, the shared line-FFT passes with per-system layouts, influence function,
// gather).  Orthorhombic boxes.

#include "pme_core.h"
#include "tff_common.h"
#include "torchff_hip.h"

namespace {

using namespace tffpme;

constexpr double SQ_SCALE = 1099511627776.0;  // 2^40: charge grids
constexpr int SERVE_LINES = 16;               // lines per FFT tile
constexpr int FFT_BLOCKS = 2048;              // workgroups of an FFT pass (each loops over tiles)

struct SysDesc {
  int atom_off, natoms;   // packed atom range
  int grid_off;           // first point of the system's grids (K0 K1 K2 points)
  int task_off, ntask;    // real-space tasks (64-atom i blocks)
  int tile_off[3], ntile[3];  // FFT tiles of the z, y and x passes
  int mod_off[3];         // B-spline moduli of the three axes in the moduli table
  int nimg[3];            // image range -nimg..nimg per axis around the minimum image
  double L[3], invL[3];
  double cutoff, alpha;
  PassArgs pa[3];         // z, y, x pass layouts
  FftPlan plan[3];
  ConvArgs cv;
};

struct BatchHead {
  int nsys, natoms, ntask, ntile[3];
};

__device__ __forceinline__ int find_sys(const SysDesc* __restrict__ d, int nsys, int key, int which) {
  // last system whose offset (which: 0 atoms, 1 tasks, 2..4 tiles of pass which - 2) is <= key
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int off = which == 0 ? d[mid].atom_off : which == 1 ? d[mid].task_off : d[mid].tile_off[which - 2];
    if (off <= key) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void k_serve_clear(const BatchHead* __restrict__ hd, const SysDesc* __restrict__ d,
                              unsigned long long* __restrict__ grid) {
  const int nsys = hd->nsys;
  if (nsys <= 0) return;
  const SysDesc& last = d[nsys - 1];
  const long long npts = (long long)last.grid_off + (long long)last.cv.K[0] * last.cv.K[1] * last.cv.K[2];
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < npts; i += (long long)gridDim.x * blockDim.x)
    grid[i] = 0ull;
}

template <typename T> struct P4 { T x, y, z, w; };

// ---------------------------------------------------------------- real space
constexpr int RW = 4;  // waves per real-space task: each takes every RW-th 64-atom chunk of j

template <typename TI, typename TC>
__global__ __launch_bounds__(RW * TFF_WAVE) void k_serve_real(const BatchHead* __restrict__ hd,
                                                              const SysDesc* __restrict__ d,
                                                              const TI* __restrict__ coords, const TI* __restrict__ q,
                                                              const TI* __restrict__ sig, const TI* __restrict__ eps,
                                                              const int* __restrict__ mol, double coul,
                                                              double* __restrict__ freal,
                                                              double* __restrict__ epart) {
  __shared__ P4<TC> s_p[RW][TFF_WAVE];    // wrapped position, charge
  __shared__ TC s_s[RW][TFF_WAVE];        // sigma / 2
  __shared__ TC s_e[RW][TFF_WAVE];        // 2 sqrt(eps)
  __shared__ int s_m[RW][TFF_WAVE];       // molecule
  __shared__ double s_red[RW][4][TFF_WAVE];
  const int task = blockIdx.x;
  if (task >= hd->ntask) return;  // uniform over the block
  const int s = find_sys(d, hd->nsys, task, 1);
  const SysDesc& sd = d[s];
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int wv = threadIdx.x / TFF_WAVE;
  const int n = sd.natoms, a0 = sd.atom_off;
  const int il = (task - sd.task_off) * TFF_WAVE + lane;
  const bool vi = il < n;
  const TC L[3] = {(TC)sd.L[0], (TC)sd.L[1], (TC)sd.L[2]};
  const TC iL[3] = {(TC)sd.invL[0], (TC)sd.invL[1], (TC)sd.invL[2]};
  const TC rc2 = (TC)(sd.cutoff * sd.cutoff), alpha = (TC)sd.alpha;
  TC xi[3] = {0, 0, 0}, qi = 0, si = 0, ei = 0;
  int mi = -1;
  if (vi) {
    const long long i = a0 + il;
    for (int a = 0; a < 3; ++a) {
      double f = (double)coords[3 * i + a] * sd.invL[a];
      f -= floor(f);
      xi[a] = (TC)(f * sd.L[a]);
    }
    qi = (TC)((double)q[i] * coul);
    si = (TC)(0.5 * (double)sig[i]);
    ei = (TC)(2.0 * sqrt((double)eps[i] > 0 ? (double)eps[i] : 0.0));
    mi = mol[i];
  }
  const TC c2 = (TC)1.1283791670955126;  // 2 / sqrt(pi)
  double fx = 0, fy = 0, fz = 0, e_acc = 0;
  for (int j0 = wv * TFF_WAVE; j0 < n; j0 += RW * TFF_WAVE) {
    const int jl = j0 + lane;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (jl < n) {
      const long long j = a0 + jl;
      double w[3];
      for (int a = 0; a < 3; ++a) {
        double f = (double)coords[3 * j + a] * sd.invL[a];
        f -= floor(f);
        w[a] = f * sd.L[a];
      }
      s_p[wv][lane] = {(TC)w[0], (TC)w[1], (TC)w[2], (TC)q[j]};
      s_s[wv][lane] = (TC)(0.5 * (double)sig[j]);
      s_e[wv][lane] = (TC)(2.0 * sqrt((double)eps[j] > 0 ? (double)eps[j] : 0.0));
      s_m[wv][lane] = mol[j];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int nj = n - j0 < TFF_WAVE ? n - j0 : TFF_WAVE;
    TC ex = 0, ey = 0, ez = 0, ee = 0;
    for (int k = 0; k < nj; ++k) {
      const P4<TC> pj = s_p[wv][k];
      const bool same_atom = j0 + k == il;
      const bool same_mol = s_m[wv][k] == mi;
      // minimum image first: n = 0 is then the nearest image, the one an exclusion removes
      TC d0x = xi[0] - pj.x, d0y = xi[1] - pj.y, d0z = xi[2] - pj.z;
      d0x -= L[0] * rint(d0x * iL[0]);
      d0y -= L[1] * rint(d0y * iL[1]);
      d0z -= L[2] * rint(d0z * iL[2]);
      for (int nx = -sd.nimg[0]; nx <= sd.nimg[0]; ++nx)
        for (int ny = -sd.nimg[1]; ny <= sd.nimg[1]; ++ny)
          for (int nz = -sd.nimg[2]; nz <= sd.nimg[2]; ++nz) {
            const TC dx = d0x - (TC)nx * L[0], dy = d0y - (TC)ny * L[1], dz = d0z - (TC)nz * L[2];
            const TC r2 = dx * dx + dy * dy + dz * dz;
            const bool zero_img = nx == 0 && ny == 0 && nz == 0;
            if (!vi || r2 > rc2 || (same_atom && zero_img)) continue;
            const bool excl = same_mol && zero_img;
            const TC rinv = (TC)1 / sqrt(r2);
            const TC rinv2 = rinv * rinv;
            const TC u = alpha * r2 * rinv;
            const TC em = exp(-u * u);
            TC ec = erfc(u);
            if (excl) ec -= (TC)1;
            const TC qq = qi * pj.w;
            ee += qq * ec * rinv;
            TC fs = qq * (ec + c2 * u * em) * rinv * rinv2;
            if (!excl) {
              const TC sg = si + s_s[wv][k];
              const TC ep4 = ei * s_e[wv][k];
              const TC s2 = sg * sg * rinv2;
              const TC s6 = s2 * s2 * s2;
              ee += ep4 * s6 * (s6 - (TC)1);
              fs += ep4 * s6 * ((TC)12 * s6 - (TC)6) * rinv2;
            }
            ex += fs * dx;
            ey += fs * dy;
            ez += fs * dz;
          }
    }
    fx += (double)ex;
    fy += (double)ey;
    fz += (double)ez;
    e_acc += (double)ee;
  }
  // the RW partial sums of each atom, added in wave order
  s_red[wv][0][lane] = fx;
  s_red[wv][1][lane] = fy;
  s_red[wv][2][lane] = fz;
  s_red[wv][3][lane] = e_acc;
  __syncthreads();
  if (wv != 0) return;
  for (int w = 1; w < RW; ++w) {
    fx += s_red[w][0][lane];
    fy += s_red[w][1][lane];
    fz += s_red[w][2][lane];
    e_acc += s_red[w][3][lane];
  }
  if (vi) {
    const long long slot = (long long)task * TFF_WAVE + lane;  // batch-local scratch index
    freal[3 * slot + 0] = fx;
    freal[3 * slot + 1] = fy;
    freal[3 * slot + 2] = fz;
  }
  const double tot = tff_wave_sum(0.5 * e_acc);  // every pair is visited from both ends
  if (lane == 0) epart[task] = tot;
}

// ---------------------------------------------------------------- reciprocal space
// the atom of lane `lane` of task `task` (64-atom blocks of the systems), or -1
__device__ __forceinline__ long long task_atom(const BatchHead* __restrict__ hd, const SysDesc* __restrict__ d,
                                               int task, int lane, int& s) {
  if (task >= hd->ntask) return -1;
  s = find_sys(d, hd->nsys, task, 1);
  const int il = (task - d[s].task_off) * TFF_WAVE + lane;
  return il < d[s].natoms ? (long long)d[s].atom_off + il : -1;
}

template <typename TI, typename TC, int P>
__global__ __launch_bounds__(TFF_WAVE) void k_serve_spread(const BatchHead* __restrict__ hd,
                                                           const SysDesc* __restrict__ d,
                                                           const TI* __restrict__ coords, const TI* __restrict__ q,
                                                           unsigned long long* __restrict__ grid) {
  int s = 0;
  const long long i = task_atom(hd, d, blockIdx.x, threadIdx.x, s);
  if (i < 0) return;
  const SysDesc& sd = d[s];
  const int K0 = sd.cv.K[0], K1 = sd.cv.K[1], K2 = sd.cv.K[2];
  TC Mv[3][P], dM[3][P];
  int base[3];
  for (int a = 0; a < 3; ++a) {
    const double sa = (double)coords[3 * (long long)i + a] * sd.invL[a] * (double)sd.cv.K[a];
    double t;
    anchor_axis(sa, sd.cv.K[a], P, base[a], t);
    spline_weights<TC, P>((TC)t, Mv[a], dM[a]);
  }
  const double qi = (double)q[i];
  unsigned long long* g = grid + sd.grid_off;
  int gx = base[0];
  for (int a = 0; a < P; ++a) {
    int gy = base[1];
    for (int b = 0; b < P; ++b) {
      const double wab = qi * (double)Mv[0][a] * (double)Mv[1][b];
      unsigned long long* row = g + ((long long)gx * K1 + gy) * K2;
      int gz = base[2];
#pragma unroll
      for (int c = 0; c < P; ++c) {
        atomicAdd(&row[gz], (unsigned long long)__double2ll_rn(wab * (double)Mv[2][c] * SQ_SCALE));
        gz = gz + 1 == K2 ? 0 : gz + 1;
      }
      gy = gy + 1 == K1 ? 0 : gy + 1;
    }
    gx = gx + 1 == K0 ? 0 : gx + 1;
  }
}

// one pass of the line FFTs over every system of the batch: block -> (system, tile)
template <typename T, int MODE>
__global__ __launch_bounds__(FFT_THREADS) void k_serve_fft(const BatchHead* __restrict__ hd,
                                                           const SysDesc* __restrict__ d,
                                                           const unsigned long long* __restrict__ grid,
                                                           cx<T>* __restrict__ work, T* __restrict__ phi,
                                                           const T* __restrict__ moduli,
                                                           double* __restrict__ kpart) {
  constexpr int PI = MODE == PASS_Z_FWD || MODE == PASS_Z_INV ? 0 : (MODE == PASS_X_CONV ? 2 : 1);
  const int nt = hd->ntile[PI];
  for (int b = blockIdx.x; b < nt; b += gridDim.x) {  // uniform over the block
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

template <typename TI, typename TC, int P>
__global__ __launch_bounds__(TFF_WAVE) void k_serve_gather(const BatchHead* __restrict__ hd,
                                                           const SysDesc* __restrict__ d,
                                                           const TI* __restrict__ coords, const TI* __restrict__ q,
                                                           const TC* __restrict__ phi, double coul,
                                                           double* __restrict__ fpme) {
  int s = 0;
  const long long i = task_atom(hd, d, blockIdx.x, threadIdx.x, s);
  if (i < 0) return;
  const SysDesc& sd = d[s];
  const long long slot = (long long)blockIdx.x * TFF_WAVE + threadIdx.x;  // batch-local scratch index
  const int K1 = sd.cv.K[1], K2 = sd.cv.K[2];
  TC Mv[3][P], dM[3][P];
  int base[3];
  for (int a = 0; a < 3; ++a) {
    const double sa = (double)coords[3 * (long long)i + a] * sd.invL[a] * (double)sd.cv.K[a];
    double t;
    anchor_axis(sa, sd.cv.K[a], P, base[a], t);
    spline_weights<TC, P>((TC)t, Mv[a], dM[a]);
  }
  const TC* g = phi + sd.grid_off;
  double du0 = 0, du1 = 0, du2 = 0;
  int gx = base[0];
  for (int a = 0; a < P; ++a) {
    int gy = base[1];
    double t0 = 0, t1 = 0, t2 = 0;
    for (int b = 0; b < P; ++b) {
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
      t0 += (double)Mv[1][b] * s0;
      t1 += (double)dM[1][b] * s0;
      t2 += (double)Mv[1][b] * s1;
      gy = gy + 1 == K1 ? 0 : gy + 1;
    }
    du0 += (double)dM[0][a] * t0;
    du1 += (double)Mv[0][a] * t1;
    du2 += (double)Mv[0][a] * t2;
    gx = gx + 1 == sd.cv.K[0] ? 0 : gx + 1;
  }
  // force = -k q grad phi; d/ds of the weights is -d/dt, d s_a / d r_a = K_a / L_a
  const double kq = coul * (double)q[i];
  fpme[3 * slot + 0] = kq * du0 * sd.cv.K[0] * sd.invL[0];
  fpme[3 * slot + 1] = kq * du1 * sd.cv.K[1] * sd.invL[1];
  fpme[3 * slot + 2] = kq * du2 * sd.cv.K[2] * sd.invL[2];
}

// forces (real + reciprocal) into the caller's atom arrays; blocks >= max_tasks: one system each,
// its energies out[5 s ..] = [real space (LJ + Coulomb incl. exclusion correction), 0, reciprocal,
// self, total] (s = position in the batch)
template <typename TI>
__global__ __launch_bounds__(TFF_WAVE) void k_serve_finalize(const BatchHead* __restrict__ hd,
                                                             const SysDesc* __restrict__ d, const TI* __restrict__ q,
                                                             const double* __restrict__ freal,
                                                             const double* __restrict__ fpme,
                                                             const double* __restrict__ epart,
                                                             const double* __restrict__ kpart, double coul,
                                                             int max_tasks, TI* __restrict__ forces,
                                                             double* __restrict__ energies) {
  if ((int)blockIdx.x < max_tasks) {
    int s = 0;
    const long long i = task_atom(hd, d, blockIdx.x, threadIdx.x, s);
    if (i < 0) return;
    const long long slot = (long long)blockIdx.x * TFF_WAVE + threadIdx.x;
    for (int a = 0; a < 3; ++a) forces[3 * i + a] = (TI)(freal[3 * slot + a] + fpme[3 * slot + a]);
    return;
  }
  const int s = blockIdx.x - max_tasks;
  if (s >= hd->nsys) return;
  const SysDesc& sd = d[s];
  double er = 0, ek = 0, q2 = 0;
  for (int t = threadIdx.x; t < sd.ntask; t += blockDim.x) er += epart[sd.task_off + t];
  for (int t = threadIdx.x; t < sd.ntile[2]; t += blockDim.x) ek += kpart[sd.tile_off[2] + t];
  for (int t = threadIdx.x; t < sd.natoms; t += blockDim.x) {
    const double v = (double)q[sd.atom_off + t];
    q2 += v * v;
  }
  er = tff_wave_sum(er);
  ek = tff_wave_sum(ek);
  q2 = tff_wave_sum(q2);
  if (threadIdx.x == 0) {
    double* o = energies + 5 * (long long)s;
    o[0] = er;
    o[1] = 0.0;
    o[2] = coul * ek;
    o[3] = -coul * sd.alpha * 0.5641895835477563 * q2;
    o[4] = o[0] + o[2] + o[3];
  }
}

struct ServeCaps {
  int max_sys, max_tasks, max_tiles, max_k;
  long long max_grid;
};

struct ServeWs {
  BatchHead* head;  // the head and the records are contiguous, laid out as the host table
  SysDesc* desc;
  unsigned long long* grid;
  void* work;
  void* phi;
  double *freal, *fpme, *epart, *kpart;
  size_t bytes;
};

constexpr size_t HEAD_BYTES = 256;  // the head, padded: records start 256 bytes in
static_assert(sizeof(BatchHead) <= HEAD_BYTES, "batch head");

ServeWs serve_carve(void* base, const ServeCaps& c, size_t tc) {
  ServeWs w{};
  size_t off = 0;
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* r = p ? p + off : nullptr;
    off += tff_align(b > 0 ? b : 1);
    return r;
  };
  w.head = (BatchHead*)take(HEAD_BYTES + sizeof(SysDesc) * (size_t)c.max_sys);
  w.desc = w.head ? (SysDesc*)((char*)w.head + HEAD_BYTES) : nullptr;
  w.grid = (unsigned long long*)take(8 * (size_t)c.max_grid);
  w.work = take(2 * tc * (size_t)c.max_grid);
  w.phi = take(tc * (size_t)c.max_grid);
  w.freal = (double*)take(24 * TFF_WAVE * (size_t)c.max_tasks);
  w.fpme = (double*)take(24 * TFF_WAVE * (size_t)c.max_tasks);
  w.epart = (double*)take(8 * (size_t)c.max_tasks);
  w.kpart = (double*)take(8 * (size_t)c.max_tiles);
  w.bytes = off;
  return w;
}

size_t serve_lds(int max_k, size_t celem) {
  return (size_t)max_k * celem + 2 * (size_t)SERVE_LINES * (max_k + 1) * celem;
}

int check_caps(const ServeCaps& c, int compute_dtype) {
  if (c.max_sys < 1 || c.max_tasks < 1 || c.max_tiles < 1 || c.max_k < 1 || c.max_grid < 1 ||
      (compute_dtype != TFF_F32 && compute_dtype != TFF_F64)) {
    tff_set_error("serve: capacities must be positive, compute dtype 0 or 1");
    return TFF_ERR_ARG;
  }
  const size_t lds = serve_lds(c.max_k, 2 * (compute_dtype == TFF_F64 ? 8 : 4));
  if (lds > 64 * 1024) {
    tff_set_error("serve: max_k %d needs %zu bytes of LDS per FFT tile (64 KiB limit)", c.max_k, lds);
    return TFF_ERR_UNSUPPORTED;
  }
  return TFF_OK;
}

template <typename TI, typename TC>
int run_serve(const ServeWs& w, const ServeCaps& c, const TI* coords, const TI* q, const TI* sig, const TI* eps,
              const int* mol, double coul, int order, const TC* moduli, TI* forces, double* energies,
              hipStream_t st) {
  const unsigned nt = (unsigned)c.max_tasks;
  const size_t lds = serve_lds(c.max_k, 2 * sizeof(TC));
  hipLaunchKernelGGL(k_serve_clear, dim3(1024), dim3(256), 0, st, w.head, w.desc, w.grid);
  hipLaunchKernelGGL((k_serve_real<TI, TC>), dim3(nt), dim3(RW * TFF_WAVE), 0, st, w.head, w.desc, coords, q, sig, eps,
                     mol, coul, w.freal, w.epart);
#define TFF_SERVE_PME(P)                                                                                         \
  {                                                                                                              \
    hipLaunchKernelGGL((k_serve_spread<TI, TC, P>), dim3(nt), dim3(TFF_WAVE), 0, st, w.head, w.desc, coords, q,  \
                       w.grid);                                                                                  \
    const unsigned nb = (unsigned)(c.max_tiles < FFT_BLOCKS ? c.max_tiles : FFT_BLOCKS);                        \
    hipLaunchKernelGGL((k_serve_fft<TC, PASS_Z_FWD>), dim3(nb), dim3(FFT_THREADS), lds, st, w.head, w.desc,       \
                       w.grid, (cx<TC>*)w.work, (TC*)w.phi, moduli, w.kpart);                                    \
    hipLaunchKernelGGL((k_serve_fft<TC, PASS_Y_FWD>), dim3(nb), dim3(FFT_THREADS), lds, st, w.head, w.desc,       \
                       w.grid, (cx<TC>*)w.work, (TC*)w.phi, moduli, w.kpart);                                    \
    hipLaunchKernelGGL((k_serve_fft<TC, PASS_X_CONV>), dim3(nb), dim3(FFT_THREADS), lds, st, w.head, w.desc,      \
                       w.grid, (cx<TC>*)w.work, (TC*)w.phi, moduli, w.kpart);                                    \
    hipLaunchKernelGGL((k_serve_fft<TC, PASS_Y_INV>), dim3(nb), dim3(FFT_THREADS), lds, st, w.head, w.desc,       \
                       w.grid, (cx<TC>*)w.work, (TC*)w.phi, moduli, w.kpart);                                    \
    hipLaunchKernelGGL((k_serve_fft<TC, PASS_Z_INV>), dim3(nb), dim3(FFT_THREADS), lds, st, w.head, w.desc,       \
                       w.grid, (cx<TC>*)w.work, (TC*)w.phi, moduli, w.kpart);                                    \
    hipLaunchKernelGGL((k_serve_gather<TI, TC, P>), dim3(nt), dim3(TFF_WAVE), 0, st, w.head, w.desc, coords, q,  \
                       (const TC*)w.phi, coul, w.fpme);                                                          \
  }
  switch (order) {
    case 3: TFF_SERVE_PME(3); break;
    case 4: TFF_SERVE_PME(4); break;
    case 5: TFF_SERVE_PME(5); break;
    case 6: TFF_SERVE_PME(6); break;
    case 7: TFF_SERVE_PME(7); break;
    case 8: TFF_SERVE_PME(8); break;
    default: tff_set_error("serve: unsupported spline order %d (3..8)", order); return TFF_ERR_UNSUPPORTED;
  }
#undef TFF_SERVE_PME
  hipLaunchKernelGGL(k_serve_finalize<TI>, dim3(nt + (unsigned)c.max_sys), dim3(TFF_WAVE), 0, st, w.head, w.desc, q,
                     w.freal, w.fpme, w.epart, w.kpart, coul, (int)nt, forces, energies);
  TFF_CHECK_LAUNCH("serve energy+forces");
  return TFF_OK;
}

}  // namespace

extern "C" size_t tff_serve_table_bytes(int max_systems) {
  return HEAD_BYTES + sizeof(SysDesc) * (size_t)(max_systems > 0 ? max_systems : 0);
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
  const long long t[7] = {h->nsys, h->natoms, h->ntask, h->ntile[0], h->ntile[1], h->ntile[2], grid};
  for (int k = 0; k < 7; ++k) totals_host[k] = t[k];
  return TFF_OK;
}

// Appends one system to the host table (see the header for the argument meaning).
extern "C" int tff_serve_table_add(void* table_host, int compute_dtype, int max_systems, int max_tasks,
                                   int max_tiles, int max_k, long long max_grid_points, int atom_off, int natoms,
                                   const double* L, const int* K, const int* mod_off, double cutoff,
                                   double alpha) {
  const ServeCaps c{max_systems, max_tasks, max_tiles, max_k, max_grid_points};
  int rc = check_caps(c, compute_dtype);
  if (rc) return rc;
  if (!table_host || !L || !K || !mod_off || natoms < 1 || atom_off < 0 || !(cutoff > 0) || !(alpha > 0)) {
    tff_set_error("serve: invalid system arguments");
    return TFF_ERR_ARG;
  }
  BatchHead* h = (BatchHead*)table_host;
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
    d.nimg[a] = (int)floor(0.5 + cutoff / L[a]);
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
  const long long grid_end = (long long)d.grid_off + (long long)K[0] * K[1] * K[2];
  if (h->nsys + 1 > max_systems || h->ntask + d.ntask > max_tasks || grid_end > max_grid_points ||
      h->ntile[0] + d.ntile[0] > max_tiles || h->ntile[1] + d.ntile[1] > max_tiles ||
      h->ntile[2] + d.ntile[2] > max_tiles || (long long)h->natoms + natoms > 0x7fffffffLL)
    return TFF_FULL;
  recs[h->nsys] = d;
  h->nsys += 1;
  h->natoms += natoms;
  h->ntask += d.ntask;
  for (int p = 0; p < 3; ++p) h->ntile[p] += d.ntile[p];
  return TFF_OK;
}

extern "C" size_t tff_serve_workspace_bytes(int compute_dtype, int max_systems, int max_tasks, int max_tiles,
                                            int max_k, long long max_grid_points) {
  const ServeCaps c{max_systems, max_tasks, max_tiles, max_k, max_grid_points};
  if (check_caps(c, compute_dtype)) return 0;
  return serve_carve(nullptr, c, compute_dtype == TFF_F64 ? 8 : 4).bytes;
}

// Copies the host table (head and records) into the workspace on `stream`.
extern "C" int tff_serve_upload(const void* table_host, int compute_dtype, int max_systems, int max_tasks,
                                int max_tiles, int max_k, long long max_grid_points, void* workspace,
                                size_t workspace_bytes, hipStream_t stream) {
  const ServeCaps c{max_systems, max_tasks, max_tiles, max_k, max_grid_points};
  int rc = check_caps(c, compute_dtype);
  if (rc) return rc;
  const ServeWs w = serve_carve(workspace, c, compute_dtype == TFF_F64 ? 8 : 4);
  if (!table_host || !workspace || workspace_bytes < w.bytes) {
    tff_set_error("serve: null table or workspace too small");
    return TFF_ERR_ARG;
  }
  const BatchHead* h = (const BatchHead*)table_host;
  if (h->nsys < 0 || h->nsys > max_systems) {
    tff_set_error("serve: table holds %d systems, capacity %d", h->nsys, max_systems);
    return TFF_ERR_ARG;
  }
  const size_t bytes = HEAD_BYTES + sizeof(SysDesc) * (size_t)h->nsys;
  if (hipMemcpyAsync(w.head, table_host, bytes, hipMemcpyHostToDevice, stream) != hipSuccess) {
    tff_set_error("serve: table upload failed");
    return TFF_ERR_LAUNCH;
  }
  return TFF_OK;
}

extern "C" int tff_serve_energy_forces(int io_dtype, int compute_dtype, const void* coords, const void* charges,
                                       const void* sigma, const void* epsilon, const int32_t* molecule,
                                       double coul_constant, int order, const void* moduli, int max_systems,
                                       int max_tasks, int max_tiles, int max_k, long long max_grid_points,
                                       double* energies_out, void* forces_out, void* workspace,
                                       size_t workspace_bytes, hipStream_t stream) {
  const ServeCaps c{max_systems, max_tasks, max_tiles, max_k, max_grid_points};
  int rc = check_caps(c, compute_dtype);
  if (rc) return rc;
  const size_t tc = compute_dtype == TFF_F64 ? 8 : 4;
  const size_t need = serve_carve(nullptr, c, tc).bytes;
  if (!workspace || workspace_bytes < need || !coords || !charges || !sigma || !epsilon || !molecule ||
      !moduli || !energies_out || !forces_out) {
    tff_set_error("serve: invalid arguments or workspace too small (%zu < %zu)", workspace_bytes, need);
    return TFF_ERR_ARG;
  }
  ServeWs w = serve_carve(workspace, c, tc);
#define TFF_RUN(TI, TC)                                                                                          \
  return run_serve<TI, TC>(w, c, (const TI*)coords, (const TI*)charges, (const TI*)sigma, (const TI*)epsilon,   \
                           molecule, coul_constant, order, (const TC*)moduli, (TI*)forces_out,                   \
                           energies_out, stream)
  if (io_dtype == TFF_F64 && compute_dtype == TFF_F32) TFF_RUN(double, float);
  if (io_dtype == TFF_F32 && compute_dtype == TFF_F32) TFF_RUN(float, float);
  if (io_dtype == TFF_F64 && compute_dtype == TFF_F64) TFF_RUN(double, double);
  if (io_dtype == TFF_F32 && compute_dtype == TFF_F64) TFF_RUN(float, double);
#undef TFF_RUN
  tff_set_error("serve: unsupported dtype combination");
  return TFF_ERR_ARG;
}
