// Direct (FFT-free) Ewald reciprocal sum: the reference's ewald_long_range / ewald_long_range_all
// (csrc/ewald/ewald.cu:9-171, 304-610; torchff/ewald.py:60-110).  Point charges below; point
// dipoles and quadrupoles in the *_mp kernels further down.
//
//   integer vectors m of the half space (m_x > 0, or m_x = 0 and m_y > 0, or m_x = m_y = 0 and
//   m_z > 0) with |m_a| <= kmax; k = B^-1 m (the columns of B^-1 are the reciprocal lattice
//   vectors of the row-major lattice B; k . r = m . s, s = r B^-1 the fractional coordinates);
//   G(k) = exp(-pi^2 k^2 / alpha^2) / k^2;  S(k) = sum_j q_j exp(2 pi i m . s_j);
//   E   = 1/(pi V) sum_k G |S|^2 - alpha/sqrt(pi) sum_j q_j^2,
//   phi_i = 2/(pi V) sum_k G (Re S cos theta_i + Im S sin theta_i) - 2 alpha/sqrt(pi) q_i,
//   F_i = -dE/dr_i = 4 q_i / V sum_k G k (Re S sin theta_i - Im S cos theta_i),  theta_i = 2 pi m . s_i.
//
// Layout: per atom and axis the phases exp(2 pi i h s_a), h = -kmax..kmax (one sincos each, in the
// input precision); one workgroup of 16 waves per (m_x, m_y) line, the lanes of a wave taking the
// m_z values and the waves a share of the atoms (fixed-order reduction), G S stored per vector; one wave per
// atom for the potential and force, the lanes striding over the vectors (reading G S through
// L1 / L2) and a wave sum at the end.  Every sum runs in a fixed order, in float64.

#include "tff_common.h"
#include "torchff_hip.h"

namespace {

constexpr int EW_WAVES = 16;  // waves per line workgroup
constexpr int EW_MAXK = 31;  // 2 kmax + 1 lanes

template <typename T> struct CX { T x, y; };

__device__ __forceinline__ void sincos_pi2(double v, double& s, double& c) { sincospi(2.0 * v, &s, &c); }
__device__ __forceinline__ void sincos_pi2(float v, float& s, float& c) { sincospif(2.0f * v, &s, &c); }

// fractional coordinates and the per-axis phase tables tab[(a * N + j) * W + (h + kmax)]
template <typename T>
__global__ void k_ewald_phases(const T* __restrict__ coords, long long n, TffBox box, int kmax,
                               CX<T>* __restrict__ tab) {
  const int W = 2 * kmax + 1;
  const long long total = n * 3 * W;
  for (long long idx = blockIdx.x * (long long)blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int h = (int)(idx % W) - kmax;
    const long long ja = idx / W;
    const int a = (int)(ja / n);
    const long long j = ja - (long long)a * n;
    const double r[3] = {(double)coords[3 * j], (double)coords[3 * j + 1], (double)coords[3 * j + 2]};
    double s = r[0] * box.inv[a] + r[1] * box.inv[3 + a] + r[2] * box.inv[6 + a];
    s -= floor(s);  // the phase of an integer multiple does not change; keeps h s small
    // h s reduced to [-1/2, 1/2) before the transcendental (exact: h is an integer)
    double v = (double)h * s;
    v -= rint(v);
    T sn, cs;
    sincos_pi2((T)v, sn, cs);
    tab[idx] = {cs, sn};
  }
}

__device__ __forceinline__ void line_of(int line, int kmax, int& mx, int& my) {
  if (line <= kmax) {
    mx = 0;
    my = line;
  } else {
    const int t = line - (kmax + 1);
    mx = 1 + t / (2 * kmax + 1);
    my = t % (2 * kmax + 1) - kmax;
  }
}

// S(k) of one (m_x, m_y) line per workgroup; lane = m_z + kmax; energy partial of the line
template <typename T>
__global__ __launch_bounds__(EW_WAVES * TFF_WAVE) void k_ewald_structure(const T* __restrict__ q, long long n,
                                                                          int kmax, TffBox box, double alpha,
                                                                          const CX<T>* __restrict__ tab,
                                                                          double* __restrict__ S,
                                                                          double* __restrict__ eline) {
  __shared__ double red[EW_WAVES][2][TFF_WAVE];
  const int line = blockIdx.x;
  const int lane = threadIdx.x & (TFF_WAVE - 1), wv = threadIdx.x / TFF_WAVE;
  const int W = 2 * kmax + 1;
  int mx, my;
  line_of(line, kmax, mx, my);
  const bool lane_ok = lane < W;
  const int hz = lane_ok ? lane : 0;
  double sr = 0, si = 0;
  for (long long j = wv; j < n; j += EW_WAVES) {
    const CX<T> tx = tab[(0 * n + j) * W + mx + kmax];
    const CX<T> ty = tab[(1 * n + j) * W + my + kmax];
    const CX<T> tz = tab[(2 * n + j) * W + hz];
    const double qj = (double)q[j];
    const double ax = (double)tx.x * ty.x - (double)tx.y * ty.y, ay = (double)tx.x * ty.y + (double)tx.y * ty.x;
    sr += qj * (ax * tz.x - ay * tz.y);
    si += qj * (ax * tz.y + ay * tz.x);
  }
  red[wv][0][lane] = sr;
  red[wv][1][lane] = si;
  __syncthreads();
  if (wv != 0) return;
  sr = si = 0;
  for (int w = 0; w < EW_WAVES; ++w) {
    sr += red[w][0][lane];
    si += red[w][1][lane];
  }
  const int mz = lane - kmax;
  // half space: m_x = m_y = 0 keeps m_z > 0 only
  const bool valid = lane_ok && !(mx == 0 && my == 0 && mz <= 0);
  double e = 0, g = 0;
  if (valid) {
    const double kx = box.inv[0] * mx + box.inv[1] * my + box.inv[2] * mz;
    const double ky = box.inv[3] * mx + box.inv[4] * my + box.inv[5] * mz;
    const double kz = box.inv[6] * mx + box.inv[7] * my + box.inv[8] * mz;
    const double k2 = kx * kx + ky * ky + kz * kz;
    g = exp(-M_PI * M_PI * k2 / (alpha * alpha)) / k2;
    e = g * (sr * sr + si * si);
  }
  // G S: the per-atom pass then needs no transcendental per k
  if (lane_ok) {
    S[2 * ((size_t)line * W + lane)] = g * sr;
    S[2 * ((size_t)line * W + lane) + 1] = g * si;
  }
  e = tff_wave_sum(e);
  if (lane == 0) eline[line] = e;
}

// potential (incl. the self term) and force of one atom per wave: the lanes stride over the
// half-space vectors (G S from the structure pass), then a wave sum; the energy by block 0
template <typename T>
__global__ __launch_bounds__(TFF_WAVE) void k_ewald_atoms(const T* __restrict__ q, long long n, int kmax, TffBox box,
                                                          double alpha, int nlines, const CX<T>* __restrict__ tab,
                                                          const double* __restrict__ S,
                                                          const double* __restrict__ eline,
                                                          double* __restrict__ energy, T* __restrict__ pot,
                                                          T* __restrict__ forces) {
  const long long i = blockIdx.x;
  const int lane = threadIdx.x;
  const int W = 2 * kmax + 1;
  const double V = box.det;
  if (blockIdx.x == 0 && energy) {
    // energy: line partials and q^2, each lane a fixed stride, then the wave sum
    double e = 0, q2 = 0;
    for (int l = lane; l < nlines; l += TFF_WAVE) e += eline[l];
    for (long long j = lane; j < n; j += TFF_WAVE) q2 += (double)q[j] * (double)q[j];
    e = tff_wave_sum(e);
    q2 = tff_wave_sum(q2);
    if (lane == 0) *energy = e / (M_PI * V) - alpha / sqrt(M_PI) * q2;
  }
  if (i >= n || (!pot && !forces)) return;
  const CX<T>* tx = tab + (0 * n + i) * W;
  const CX<T>* ty = tab + (1 * n + i) * W;
  const CX<T>* tz = tab + (2 * n + i) * W;
  double ph = 0, fx = 0, fy = 0, fz = 0;
  const int total = nlines * W;
  for (int kk = lane; kk < total; kk += TFF_WAVE) {
    const int line = kk / W, h = kk - line * W;
    int mx, my;
    line_of(line, kmax, mx, my);
    const int mz = h - kmax;
    if (mx == 0 && my == 0 && mz <= 0) continue;
    const CX<T> ax_ = tx[mx + kmax], ay_ = ty[my + kmax], az_ = tz[h];
    const double ax = (double)ax_.x * ay_.x - (double)ax_.y * ay_.y, ay = (double)ax_.x * ay_.y + (double)ax_.y * ay_.x;
    const double c = ax * az_.x - ay * az_.y, s = ax * az_.y + ay * az_.x;
    const double kx = box.inv[0] * mx + box.inv[1] * my + box.inv[2] * mz;
    const double ky = box.inv[3] * mx + box.inv[4] * my + box.inv[5] * mz;
    const double kz = box.inv[6] * mx + box.inv[7] * my + box.inv[8] * mz;
    const double Sr = S[2 * (size_t)kk], Si = S[2 * (size_t)kk + 1];  // G S
    ph += Sr * c + Si * s;
    const double fi = Sr * s - Si * c;
    fx += fi * kx;
    fy += fi * ky;
    fz += fi * kz;
  }
  ph = tff_wave_sum(ph);
  fx = tff_wave_sum(fx);
  fy = tff_wave_sum(fy);
  fz = tff_wave_sum(fz);
  if (lane != 0) return;
  const double qi = (double)q[i];
  if (pot) pot[i] = (T)(2.0 / (M_PI * V) * ph - 2.0 * alpha / sqrt(M_PI) * qi);
  if (forces) {
    const double f = 4.0 * qi / V;
    forces[3 * i] = (T)(f * fx);
    forces[3 * i + 1] = (T)(f * fy);
    forces[3 * i + 2] = (T)(f * fz);
  }
}

// ------------------------------------------------------------------ multipoles (ranks 1, 2)
// Point dipoles p and Cartesian quadrupoles theta (3 x 3 row-major) enter through
// L(k) = q - (4 pi^2/3) k.theta.k + i 2 pi k.p and S(k) = sum_j L_j exp(2 pi i k.r_j)
// (torchff/ewald.py:75-97).  With A_i = G S* exp(2 pi i k.r_i) summed over the half space,
//   potential  = dE/dq      = 2/(pi V) sum Re A - 2 a q/sqrt(pi)
//   field      = -dE/dp     = 4/V sum k Im A + 4 a^3/(3 sqrt(pi)) p
//   field grad = -3 dE/dth  = 8 pi/V sum Re A k k^T + 16 a^5/(15 sqrt(pi)) theta
//   force      = -dE/dr     = 4/V sum k Im(L_i A_i),
// the force contracting a second multipole set X in place of L when one is given (the
// backward of ewald_long_range_all contracts one structure factor with two sets).
struct EwMp {
  const void* q;   // N, NULL = zero
  const void* p;   // N x 3, NULL = zero
  const void* th;  // N x 9, NULL = zero
};

template <typename T>
__device__ __forceinline__ double ew_load(const void* a, long long k) {
  return (double)((const T*)a)[k];
}

// real and imaginary parts of L(k) for atom j
template <typename T>
__device__ __forceinline__ void ew_L(const EwMp& m, long long j, double kx, double ky, double kz, double& lr,
                                     double& li) {
  lr = m.q ? ew_load<T>(m.q, j) : 0.0;
  li = 0.0;
  if (m.p)
    li = 2.0 * M_PI * (kx * ew_load<T>(m.p, 3 * j) + ky * ew_load<T>(m.p, 3 * j + 1) + kz * ew_load<T>(m.p, 3 * j + 2));
  if (m.th) {
    const double k[3] = {kx, ky, kz};
    double acc = 0.0;
#pragma unroll
    for (int x = 0; x < 3; ++x)
#pragma unroll
      for (int y = 0; y < 3; ++y) acc += k[x] * ew_load<T>(m.th, 9 * j + 3 * x + y) * k[y];
    lr -= 4.0 * M_PI * M_PI / 3.0 * acc;
  }
}

__device__ __forceinline__ void ew_k(const TffBox& box, int mx, int my, int mz, double& kx, double& ky, double& kz) {
  kx = box.inv[0] * mx + box.inv[1] * my + box.inv[2] * mz;
  ky = box.inv[3] * mx + box.inv[4] * my + box.inv[5] * mz;
  kz = box.inv[6] * mx + box.inv[7] * my + box.inv[8] * mz;
}

template <typename T>
__global__ __launch_bounds__(EW_WAVES * TFF_WAVE) void k_ewald_structure_mp(EwMp src, long long n, int kmax,
                                                                             TffBox box, double alpha,
                                                                             const CX<T>* __restrict__ tab,
                                                                             double* __restrict__ S,
                                                                             double* __restrict__ eline) {
  __shared__ double red[EW_WAVES][2][TFF_WAVE];
  const int line = blockIdx.x;
  const int lane = threadIdx.x & (TFF_WAVE - 1), wv = threadIdx.x / TFF_WAVE;
  const int W = 2 * kmax + 1;
  int mx, my;
  line_of(line, kmax, mx, my);
  const bool lane_ok = lane < W;
  const int hz = lane_ok ? lane : 0;
  const int mz = lane - kmax;
  double kx, ky, kz;
  ew_k(box, mx, my, mz, kx, ky, kz);
  double sr = 0, si = 0;
  for (long long j = wv; j < n; j += EW_WAVES) {
    const CX<T> tx = tab[(0 * n + j) * W + mx + kmax];
    const CX<T> ty = tab[(1 * n + j) * W + my + kmax];
    const CX<T> tz = tab[(2 * n + j) * W + hz];
    const double ax = (double)tx.x * ty.x - (double)tx.y * ty.y, ay = (double)tx.x * ty.y + (double)tx.y * ty.x;
    const double c = ax * tz.x - ay * tz.y, s = ax * tz.y + ay * tz.x;
    double lr, li;
    ew_L<T>(src, j, kx, ky, kz, lr, li);
    sr += lr * c - li * s;
    si += lr * s + li * c;
  }
  red[wv][0][lane] = sr;
  red[wv][1][lane] = si;
  __syncthreads();
  if (wv != 0) return;
  sr = si = 0;
  for (int w = 0; w < EW_WAVES; ++w) {
    sr += red[w][0][lane];
    si += red[w][1][lane];
  }
  const bool valid = lane_ok && !(mx == 0 && my == 0 && mz <= 0);
  double e = 0, g = 0;
  if (valid) {
    const double k2 = kx * kx + ky * ky + kz * kz;
    g = exp(-M_PI * M_PI * k2 / (alpha * alpha)) / k2;
    e = g * (sr * sr + si * si);
  }
  if (lane_ok) {
    S[2 * ((size_t)line * W + lane)] = g * sr;
    S[2 * ((size_t)line * W + lane) + 1] = g * si;
  }
  e = tff_wave_sum(e);
  if (lane == 0) eline[line] = e;
}

// one atom's multipoles in registers
struct EwAtom {
  double q, p[3], th[9];
};

template <typename T>
__device__ __forceinline__ EwAtom ew_atom(const EwMp& m, long long j) {
  EwAtom a{};
  if (m.q) a.q = ew_load<T>(m.q, j);
  if (m.p)
    for (int k = 0; k < 3; ++k) a.p[k] = ew_load<T>(m.p, 3 * j + k);
  if (m.th)
    for (int k = 0; k < 9; ++k) a.th[k] = ew_load<T>(m.th, 9 * j + k);
  return a;
}

template <typename T>
__global__ __launch_bounds__(TFF_WAVE) void k_ewald_atoms_mp(EwMp src, EwMp con, long long n, int kmax, TffBox box,
                                                             double alpha, int nlines, const CX<T>* __restrict__ tab,
                                                             const double* __restrict__ S,
                                                             const double* __restrict__ eline,
                                                             double* __restrict__ energy, T* __restrict__ pot,
                                                             T* __restrict__ field, T* __restrict__ fgrad,
                                                             T* __restrict__ forces) {
  const long long i = blockIdx.x;
  const int lane = threadIdx.x;
  const int W = 2 * kmax + 1;
  const double V = box.det;
  const double rpi = 1.0 / sqrt(M_PI), a2 = alpha * alpha;
  if (blockIdx.x == 0 && energy) {
    double e = 0, self = 0;
    for (int l = lane; l < nlines; l += TFF_WAVE) e += eline[l];
    for (long long j = lane; j < n; j += TFF_WAVE) {
      const EwAtom a = ew_atom<T>(src, j);
      self += a.q * a.q;
      for (int k = 0; k < 3; ++k) self += 2.0 * a2 / 3.0 * a.p[k] * a.p[k];
      for (int k = 0; k < 9; ++k) self += 8.0 * a2 * a2 / 45.0 * a.th[k] * a.th[k];
    }
    e = tff_wave_sum(e);
    self = tff_wave_sum(self);
    if (lane == 0) *energy = e / (M_PI * V) - alpha * rpi * self;
  }
  if (i >= n || (!pot && !field && !fgrad && !forces)) return;
  const EwAtom X = ew_atom<T>(con, i);  // the contracted set, loaded once
  const bool xp = con.p != nullptr, xt = con.th != nullptr;
  const CX<T>* tx = tab + (0 * n + i) * W;
  const CX<T>* ty = tab + (1 * n + i) * W;
  const CX<T>* tz = tab + (2 * n + i) * W;
  double sar = 0, fa[3] = {0, 0, 0}, eg[6] = {0, 0, 0, 0, 0, 0}, fk[3] = {0, 0, 0};
  const int total = nlines * W;
  for (int kk = lane; kk < total; kk += TFF_WAVE) {
    const int line = kk / W, h = kk - line * W;
    int mx, my;
    line_of(line, kmax, mx, my);
    const int mz = h - kmax;
    if (mx == 0 && my == 0 && mz <= 0) continue;
    const CX<T> ax_ = tx[mx + kmax], ay_ = ty[my + kmax], az_ = tz[h];
    const double ax = (double)ax_.x * ay_.x - (double)ax_.y * ay_.y, ay = (double)ax_.x * ay_.y + (double)ax_.y * ay_.x;
    const double c = ax * az_.x - ay * az_.y, s = ax * az_.y + ay * az_.x;
    double kx, ky, kz;
    ew_k(box, mx, my, mz, kx, ky, kz);
    const double Sr = S[2 * (size_t)kk], Si = S[2 * (size_t)kk + 1];  // G S
    const double Ar = Sr * c + Si * s, Ai = Sr * s - Si * c;
    sar += Ar;
    fa[0] += kx * Ai;
    fa[1] += ky * Ai;
    fa[2] += kz * Ai;
    if (fgrad) {
      eg[0] += Ar * kx * kx;
      eg[1] += Ar * ky * ky;
      eg[2] += Ar * kz * kz;
      eg[3] += Ar * kx * ky;
      eg[4] += Ar * kx * kz;
      eg[5] += Ar * ky * kz;
    }
    if (forces) {
      double lr = X.q, li = 0.0;
      if (xp) li = 2.0 * M_PI * (kx * X.p[0] + ky * X.p[1] + kz * X.p[2]);
      if (xt) {
        const double kv[3] = {kx, ky, kz};
        double acc = 0.0;
#pragma unroll
        for (int x = 0; x < 3; ++x)
#pragma unroll
          for (int y = 0; y < 3; ++y) acc += kv[x] * X.th[3 * x + y] * kv[y];
        lr -= 4.0 * M_PI * M_PI / 3.0 * acc;
      }
      const double im = lr * Ai + li * Ar;
      fk[0] += kx * im;
      fk[1] += ky * im;
      fk[2] += kz * im;
    }
  }
  sar = tff_wave_sum(sar);
  for (int k = 0; k < 3; ++k) {
    fa[k] = tff_wave_sum(fa[k]);
    fk[k] = tff_wave_sum(fk[k]);
  }
  if (fgrad)
    for (int k = 0; k < 6; ++k) eg[k] = tff_wave_sum(eg[k]);
  if (lane != 0) return;
  const EwAtom A = ew_atom<T>(src, i);
  if (pot) pot[i] = (T)(2.0 / (M_PI * V) * sar - 2.0 * alpha * rpi * A.q);
  if (field) {
    const double cf = 4.0 * alpha * a2 * rpi / 3.0;
    for (int x = 0; x < 3; ++x) field[3 * i + x] = (T)(4.0 / V * fa[x] + cf * A.p[x]);
  }
  if (fgrad) {
    const double cg = 16.0 * alpha * a2 * a2 * rpi / 15.0;
    const int map[9] = {0, 3, 4, 3, 1, 5, 4, 5, 2};
    for (int k = 0; k < 9; ++k) fgrad[9 * i + k] = (T)(8.0 * M_PI / V * eg[map[k]] + cg * A.th[k]);
  }
  if (forces)
    for (int x = 0; x < 3; ++x) forces[3 * i + x] = (T)(4.0 / V * fk[x]);
}

struct EwaldWs {
  void* tab;
  double* S;
  double* eline;
  size_t bytes;
};

EwaldWs ewald_carve(void* base, long long n, int kmax, size_t tsize) {
  EwaldWs w{};
  size_t off = 0;
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* r = p ? p + off : nullptr;
    off += tff_align(b > 0 ? b : 1);
    return r;
  };
  const size_t W = 2 * (size_t)kmax + 1;
  const size_t nlines = (size_t)(kmax + 1) + (size_t)kmax * W;
  w.tab = take(2 * tsize * 3 * (size_t)n * W);
  w.S = (double*)take(16 * nlines * W);
  w.eline = (double*)take(8 * nlines);
  w.bytes = off;
  return w;
}

template <typename T>
int run_ewald(const T* coords, const T* q, long long n, const TffBox& box, int kmax, double alpha, double* energy,
              T* pot, T* forces, const EwaldWs& w, hipStream_t st) {
  const int W = 2 * kmax + 1;
  const int nlines = (kmax + 1) + kmax * W;
  const long long total = n * 3 * W;
  const unsigned g1 = (unsigned)((total + 255) / 256 < 8192 ? (total + 255) / 256 : 8192);
  hipLaunchKernelGGL(k_ewald_phases<T>, dim3(g1 > 0 ? g1 : 1), dim3(256), 0, st, coords, n, box, kmax,
                     (CX<T>*)w.tab);
  hipLaunchKernelGGL(k_ewald_structure<T>, dim3(nlines), dim3(EW_WAVES * TFF_WAVE), 0, st, q, n, kmax, box, alpha,
                     (const CX<T>*)w.tab, w.S, w.eline);
  const unsigned g3 = (unsigned)(n > 0 ? n : 1);  // one wave per atom
  hipLaunchKernelGGL(k_ewald_atoms<T>, dim3(g3), dim3(TFF_WAVE), 0, st, q, n, kmax, box, alpha, nlines,
                     (const CX<T>*)w.tab, (const double*)w.S, (const double*)w.eline, energy, pot, forces);
  TFF_CHECK_LAUNCH("ewald");
  return TFF_OK;
}

template <typename T>
int run_ewald_mp(const T* coords, const EwMp& src, const EwMp& con, long long n, const TffBox& box, int kmax,
                 double alpha, double* energy, T* pot, T* field, T* fgrad, T* forces, const EwaldWs& w,
                 hipStream_t st) {
  const int W = 2 * kmax + 1;
  const int nlines = (kmax + 1) + kmax * W;
  const long long total = n * 3 * W;
  const unsigned g1 = (unsigned)((total + 255) / 256 < 8192 ? (total + 255) / 256 : 8192);
  hipLaunchKernelGGL(k_ewald_phases<T>, dim3(g1 > 0 ? g1 : 1), dim3(256), 0, st, coords, n, box, kmax,
                     (CX<T>*)w.tab);
  hipLaunchKernelGGL(k_ewald_structure_mp<T>, dim3(nlines), dim3(EW_WAVES * TFF_WAVE), 0, st, src, n, kmax, box,
                     alpha, (const CX<T>*)w.tab, w.S, w.eline);
  const unsigned g3 = (unsigned)(n > 0 ? n : 1);  // one wave per atom
  hipLaunchKernelGGL(k_ewald_atoms_mp<T>, dim3(g3), dim3(TFF_WAVE), 0, st, src, con, n, kmax, box, alpha, nlines,
                     (const CX<T>*)w.tab, (const double*)w.S, (const double*)w.eline, energy, pot, field, fgrad,
                     forces);
  TFF_CHECK_LAUNCH("ewald multipole");
  return TFF_OK;
}

}  // namespace

extern "C" size_t tff_ewald_workspace_bytes(int dtype, long long natoms, int kmax) {
  if (natoms < 1 || kmax < 1 || kmax > EW_MAXK) return 0;
  return ewald_carve(nullptr, natoms, kmax, dtype == TFF_F64 ? 8 : 4).bytes;
}

extern "C" int tff_ewald(int dtype, const void* coords, const void* charges, long long natoms, const double* box_host,
                         int kmax, double alpha, double* energy_out, void* potential_out, void* forces_out,
                         void* workspace, size_t workspace_bytes, hipStream_t stream) {
  if (!coords || !charges || natoms < 1 || !box_host || kmax < 1 || kmax > EW_MAXK || !(alpha > 0) ||
      (dtype != TFF_F32 && dtype != TFF_F64)) {
    tff_set_error("ewald: invalid arguments (natoms >= 1, box, 1 <= kmax <= %d, alpha > 0, dtype 0/1)", EW_MAXK);
    return TFF_ERR_ARG;
  }
  const TffBox box = tff_make_box(box_host);
  if (!(box.det > 0)) {
    tff_set_error("ewald: the box must have a positive volume");
    return TFF_ERR_ARG;
  }
  const EwaldWs w = ewald_carve(workspace, natoms, kmax, dtype == TFF_F64 ? 8 : 4);
  if (!workspace || workspace_bytes < w.bytes) {
    tff_set_error("ewald: workspace too small (%zu < %zu)", workspace_bytes, w.bytes);
    return TFF_ERR_WORKSPACE;
  }
  if (dtype == TFF_F64)
    return run_ewald<double>((const double*)coords, (const double*)charges, natoms, box, kmax, alpha, energy_out,
                             (double*)potential_out, (double*)forces_out, w, stream);
  return run_ewald<float>((const float*)coords, (const float*)charges, natoms, box, kmax, alpha, energy_out,
                          (float*)potential_out, (float*)forces_out, w, stream);
}

extern "C" int tff_ewald_multipole(int dtype, const void* coords, long long natoms, const double* box_host, int kmax,
                                   double alpha, const void* charges, const void* dipoles, const void* quadrupoles,
                                   const void* c_charges, const void* c_dipoles, const void* c_quadrupoles,
                                   double* energy_out, void* potential_out, void* field_out, void* field_grad_out,
                                   void* forces_out, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  if (!coords || natoms < 1 || !box_host || kmax < 1 || kmax > EW_MAXK || !(alpha > 0) ||
      (dtype != TFF_F32 && dtype != TFF_F64)) {
    tff_set_error("ewald: invalid arguments (natoms >= 1, box, 1 <= kmax <= %d, alpha > 0, dtype 0/1)", EW_MAXK);
    return TFF_ERR_ARG;
  }
  const TffBox box = tff_make_box(box_host);
  if (!(box.det > 0)) {
    tff_set_error("ewald: the box must have a positive volume");
    return TFF_ERR_ARG;
  }
  const EwaldWs w = ewald_carve(workspace, natoms, kmax, dtype == TFF_F64 ? 8 : 4);
  if (!workspace || workspace_bytes < w.bytes) {
    tff_set_error("ewald: workspace too small (%zu < %zu)", workspace_bytes, w.bytes);
    return TFF_ERR_WORKSPACE;
  }
  const EwMp src{charges, dipoles, quadrupoles};
  const bool own = !c_charges && !c_dipoles && !c_quadrupoles;
  const EwMp con = own ? src : EwMp{c_charges, c_dipoles, c_quadrupoles};
  if (dtype == TFF_F64)
    return run_ewald_mp<double>((const double*)coords, src, con, natoms, box, kmax, alpha, energy_out,
                                (double*)potential_out, (double*)field_out, (double*)field_grad_out,
                                (double*)forces_out, w, stream);
  return run_ewald_mp<float>((const float*)coords, src, con, natoms, box, kmax, alpha, energy_out,
                             (float*)potential_out, (float*)field_out, (float*)field_grad_out, (float*)forces_out, w,
                             stream);
}
