// Drop-in pair kernels over an explicit (P, 2) int64 pair list, float32 or float64.
//
// One kernel family serves the reference ops that consume a pair list:
//   TFF_PAIR_NONBONDED  compute_nonbonded_energy_from_atom_pairs /
//                       compute_nonbonded_forces_from_atom_pairs
//                       (reference csrc/nonbonded/nonbonded_atom_pairs_cuda.cu:16-140):
//                       k q_i q_j (1/r - [1/rc]) + 4 eps_ij s^6 (s^6 - 1), LB mixing;
//   TFF_PAIR_EWALD_LJ   the fused real-space erfc-Coulomb + LB Lennard-Jones with
//                       the (erfc - 1)/r exclusion correction (additive op);
//   TFF_PAIR_COUL_EWALD compute_multipolar_energy_from_atom_pairs, rank 0
//                       (reference csrc/multipoles/multipoles.cu:24-152,
//                       multipoles_forward.cuh:355-617, damps csrc/ewald/damps.cuh:5-42);
//   TFF_PAIR_LJ_TYPES / TFF_PAIR_LJ_PAIRS  compute_lennard_jones_energy with a
//                       (T, T) type table or per-pair sigma/eps (reference csrc/vdw/vdw.cu:82-195).
// Rows with a negative index are padding and are skipped; pairs with
// r^2 > rc^2 contribute nothing.  Energy partial sums are float64 per
// workgroup and are reduced in a fixed order (deterministic); gradients are
// accumulated in float64 with hardware float64 atomics and cast to the
// caller's dtype at the end.  eps-gradient of a pair whose mixed epsilon is
// zero (e.g. TIP3P hydrogens) is defined as zero; the exact derivative of
// sqrt(eps_i eps_j) is unbounded there.

#include "tff_common.h"
#include "torchff_hip.h"

namespace {

constexpr int PB = 256;
constexpr int PMAXBLK = 4096;

struct PairArgs {
  double coul;      // Coulomb constant / prefactor
  double cutoff;
  double cutoff2;
  double alpha;     // Ewald alpha (< 0: bare Coulomb)
  int do_shift;
  long long ntypes;
};

template <typename T, int MODE>
__global__ __launch_bounds__(PB) void k_pair_terms(const T* __restrict__ coords, const long long* __restrict__ pairs,
                                                   long long P, const long long* __restrict__ pexcl, long long PE,
                                                   TffBox box, PairArgs a, const T* __restrict__ q,
                                                   const T* __restrict__ sig, const T* __restrict__ eps,
                                                   const long long* __restrict__ types, long long natoms,
                                                   double* __restrict__ e_part,
                                                   double* __restrict__ g_xyz, double* __restrict__ g_q,
                                                   double* __restrict__ g_sig, double* __restrict__ g_eps) {
  double e_acc = 0.0;
  const long long total = P + PE;
  const T rc2 = (T)a.cutoff2;
  const T inv_rc = (T)(1.0 / a.cutoff);
  const T two_over_sqrt_pi = (T)1.1283791670955126;
  for (long long idx = blockIdx.x * (long long)PB + threadIdx.x; idx < total; idx += (long long)gridDim.x * PB) {
    const bool is_excl = idx >= P;
    const long long* pr = is_excl ? pexcl + 2 * (idx - P) : pairs + 2 * idx;
    long long i = pr[0], j = pr[1];
    // negative rows are padding; indices past natoms are ignored rather than dereferenced
    if (i < 0 || j < 0 || i >= natoms || j >= natoms) continue;
    T dx = coords[3 * i] - coords[3 * j];
    T dy = coords[3 * i + 1] - coords[3 * j + 1];
    T dz = coords[3 * i + 2] - coords[3 * j + 2];
    tff_min_image<T>(box, dx, dy, dz);
    T r2 = dx * dx + dy * dy + dz * dz;
    if (r2 > rc2) continue;
    T rinv = (T)1 / sqrt(r2);
    T rinv2 = rinv * rinv;
    T e = 0, dedr_over_r = 0;  // dE/dr / r, so grad_i = dedr_over_r * d
    // ---- Coulomb part
    if (MODE == TFF_PAIR_NONBONDED) {
      T qq = q[i] * q[j] * (T)a.coul;
      T shift = a.do_shift ? inv_rc : (T)0;
      e += qq * (rinv - shift);
      dedr_over_r += -qq * rinv * rinv2;
      if (g_q) {
        T tmp = (rinv - shift) * (T)a.coul;
        atomicAdd(&g_q[i], (double)(q[j] * tmp));
        atomicAdd(&g_q[j], (double)(q[i] * tmp));
      }
    } else if (MODE == TFF_PAIR_EWALD_LJ || MODE == TFF_PAIR_COUL_EWALD) {
      T d0, d1;
      if (a.alpha >= 0) {
        T r = r2 * rinv;
        T u = (T)a.alpha * r;
        T ec = erfc(u);
        T ex = exp(-u * u) * two_over_sqrt_pi * u;
        d0 = ec;
        d1 = ec + ex;
        if (is_excl) {
          d0 -= (T)1;
          d1 -= (T)1;
        }
      } else {
        d0 = 1;
        d1 = 1;
      }
      T qq = q[i] * q[j] * (T)a.coul;
      e += qq * d0 * rinv;
      dedr_over_r += -qq * d1 * rinv * rinv2;
      if (g_q) {
        T tmp = d0 * rinv * (T)a.coul;
        atomicAdd(&g_q[i], (double)(q[j] * tmp));
        atomicAdd(&g_q[j], (double)(q[i] * tmp));
      }
    }
    // ---- Lennard-Jones part (never on exclusion-correction pairs)
    if (MODE != TFF_PAIR_COUL_EWALD && !is_excl) {
      T sij, eij;
      long long tix = 0;
      if (MODE == TFF_PAIR_LJ_TYPES) {
        long long ti = types[i], tj = types[j];
        if (ti < 0 || tj < 0 || ti >= a.ntypes || tj >= a.ntypes) continue;
        tix = ti * a.ntypes + tj;
        sij = sig[tix];
        eij = eps[tix];
      } else if (MODE == TFF_PAIR_LJ_PAIRS) {
        sij = sig[idx];
        eij = eps[idx];
      } else {
        sij = (sig[i] + sig[j]) * (T)0.5;
        eij = sqrt(eps[i] * eps[j]);
      }
      T sr2 = sij * sij * rinv2;
      T s6 = sr2 * sr2 * sr2;
      e += (T)4 * eij * s6 * (s6 - (T)1);
      dedr_over_r += -(T)24 * eij * s6 * ((T)2 * s6 - (T)1) * rinv2;
      T dsig = (T)24 * eij * s6 * ((T)2 * s6 - (T)1) / sij;  // dE/dsigma_ij
      T deps = (T)4 * s6 * (s6 - (T)1);                        // dE/deps_ij
      if (MODE == TFF_PAIR_LJ_TYPES) {
        if (g_sig) atomicAdd(&g_sig[tix], (double)dsig);
        if (g_eps) atomicAdd(&g_eps[tix], (double)deps);
      } else if (MODE == TFF_PAIR_LJ_PAIRS) {
        if (g_sig) g_sig[idx] = (double)dsig;
        if (g_eps) g_eps[idx] = (double)deps;
      } else {
        if (g_sig) {
          atomicAdd(&g_sig[i], (double)(dsig * (T)0.5));
          atomicAdd(&g_sig[j], (double)(dsig * (T)0.5));
        }
        if (g_eps && eij > (T)0) {
          T h = deps * (T)0.5 / eij;
          atomicAdd(&g_eps[i], (double)(h * eps[j]));
          atomicAdd(&g_eps[j], (double)(h * eps[i]));
        }
      }
    }
    e_acc += (double)e;
    if (g_xyz) {
      double gx = (double)(dedr_over_r * dx), gy = (double)(dedr_over_r * dy), gz = (double)(dedr_over_r * dz);
      atomicAdd(&g_xyz[3 * i], gx);
      atomicAdd(&g_xyz[3 * i + 1], gy);
      atomicAdd(&g_xyz[3 * i + 2], gz);
      atomicAdd(&g_xyz[3 * j], -gx);
      atomicAdd(&g_xyz[3 * j + 1], -gy);
      atomicAdd(&g_xyz[3 * j + 2], -gz);
    }
  }
  double tot = tff_block_sum(e_acc);
  if (threadIdx.x == 0) e_part[blockIdx.x] = tot;
}

__global__ void k_zero_d(double* p, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) p[i] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, int n, double scale,
                                                   T* __restrict__ out) {
  double s = tff_sum_partials(part, n);
  if (threadIdx.x == 0) out[0] = (T)(s * scale);
}

// dst = scale * src (accumulate: dst += scale * src)
template <typename T>
__global__ void k_cast(const double* __restrict__ src, long long n, double scale, int accumulate, T* __restrict__ dst) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) {
    double v = src[i] * scale;
    dst[i] = accumulate ? (T)((double)dst[i] + v) : (T)v;
  }
}

struct PairWs {
  double* e_part;
  double* g_xyz;
  double* g_q;
  double* g_sig;
  double* g_eps;
  size_t bytes;
};

PairWs carve(void* base, long long natoms, long long nparam) {
  PairWs w{};
  size_t off = 0;
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* r = p ? p + off : nullptr;
    off += tff_align(b);
    return r;
  };
  w.e_part = (double*)take(sizeof(double) * PMAXBLK);
  w.g_xyz = (double*)take(sizeof(double) * 3 * (size_t)(natoms > 0 ? natoms : 1));
  w.g_q = (double*)take(sizeof(double) * (size_t)(natoms > 0 ? natoms : 1));
  w.g_sig = (double*)take(sizeof(double) * (size_t)(nparam > 0 ? nparam : 1));
  w.g_eps = (double*)take(sizeof(double) * (size_t)(nparam > 0 ? nparam : 1));
  w.bytes = off;
  return w;
}

long long param_len(int mode, long long natoms, long long npairs, long long ntypes) {
  if (mode == TFF_PAIR_LJ_TYPES) return ntypes * ntypes;
  if (mode == TFF_PAIR_LJ_PAIRS) return npairs;
  return natoms;
}

template <typename T>
int run(int mode, const T* coords, long long natoms, const TffBox& box, const long long* pairs, long long P,
        const long long* pexcl, long long PE, const T* q, const T* sig, const T* eps, const long long* types,
        const PairArgs& pa, T* energy, T* g_xyz, T* g_q, T* g_sig, T* g_eps, T* forces, double scale,
        const PairWs& w, hipStream_t st) {
  const long long np = param_len(mode, natoms, P, pa.ntypes);
  const bool need_xyz = g_xyz || forces;
  auto zero = [&](double* p, long long n) {
    long long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_zero_d, dim3((unsigned)b), dim3(256), 0, st, p, n);
  };
  if (need_xyz) zero(w.g_xyz, 3 * natoms);
  if (g_q) zero(w.g_q, natoms);
  if (g_sig) zero(w.g_sig, np);
  if (g_eps) zero(w.g_eps, np);
  long long blocks = (P + PE + PB - 1) / PB;
  if (blocks > PMAXBLK) blocks = PMAXBLK;
  if (blocks < 1) blocks = 1;
  const dim3 grid((unsigned)blocks), blk(PB);
  double* gx = need_xyz ? w.g_xyz : nullptr;
  double* gq = g_q ? w.g_q : nullptr;
  double* gs = g_sig ? w.g_sig : nullptr;
  double* ge = g_eps ? w.g_eps : nullptr;
#define TFF_LAUNCH_PAIR(M)                                                                                   \
  hipLaunchKernelGGL((k_pair_terms<T, M>), grid, blk, 0, st, coords, pairs, P, pexcl, PE, box, pa, q, sig, eps, \
                     types, natoms, w.e_part, gx, gq, gs, ge)
  void* tok = tff_prof_begin("pair_terms", st);
  switch (mode) {
    case TFF_PAIR_NONBONDED: TFF_LAUNCH_PAIR(TFF_PAIR_NONBONDED); break;
    case TFF_PAIR_EWALD_LJ: TFF_LAUNCH_PAIR(TFF_PAIR_EWALD_LJ); break;
    case TFF_PAIR_COUL_EWALD: TFF_LAUNCH_PAIR(TFF_PAIR_COUL_EWALD); break;
    case TFF_PAIR_LJ_TYPES: TFF_LAUNCH_PAIR(TFF_PAIR_LJ_TYPES); break;
    case TFF_PAIR_LJ_PAIRS: TFF_LAUNCH_PAIR(TFF_PAIR_LJ_PAIRS); break;
    default: tff_set_error("tff_pair_energy: unknown mode %d", mode); return TFF_ERR_ARG;
  }
#undef TFF_LAUNCH_PAIR
  tff_prof_end(tok, st);
  TFF_CHECK_LAUNCH("k_pair_terms");
  if (energy) hipLaunchKernelGGL(k_sum_parts<T>, dim3(1), dim3(256), 0, st, w.e_part, (int)blocks, scale, energy);
  auto cast = [&](const double* s, long long n, double sc, int acc, T* d) {
    long long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_cast<T>, dim3((unsigned)b), dim3(256), 0, st, s, n, sc, acc, d);
  };
  if (g_xyz) cast(w.g_xyz, 3 * natoms, scale, 0, g_xyz);
  if (forces) cast(w.g_xyz, 3 * natoms, -scale, 1, forces);
  if (g_q) cast(w.g_q, natoms, scale, 0, g_q);
  if (g_sig) cast(w.g_sig, np, scale, 0, g_sig);
  if (g_eps) cast(w.g_eps, np, scale, 0, g_eps);
  TFF_CHECK_LAUNCH("pair finalize");
  return TFF_OK;
}

}  // namespace

extern "C" size_t tff_pair_workspace_bytes(int mode, long long natoms, long long npairs, long long ntypes) {
  return carve(nullptr, natoms, param_len(mode, natoms, npairs, ntypes)).bytes;
}

extern "C" int tff_pair_energy(int dtype, int mode, const void* coords, long long natoms, const double* box_host,
                               const int64_t* pairs, long long npairs, const int64_t* pairs_excl,
                               long long npairs_excl, const void* charges, const void* sigma, const void* epsilon,
                               const int64_t* atom_types, long long ntypes, double coul_constant, double cutoff,
                               double ewald_alpha, int do_shift, double scale, void* energy_out,
                               void* grad_coords, void* grad_charges, void* grad_sigma, void* grad_epsilon,
                               void* forces_accum, void* workspace, size_t workspace_bytes, hipStream_t stream) {
  if (natoms < 0 || npairs < 0 || npairs_excl < 0 || (npairs > 0 && !pairs) || (npairs_excl > 0 && !pairs_excl)) {
    tff_set_error("tff_pair_energy: invalid pair arguments");
    return TFF_ERR_ARG;
  }
  const bool coul = mode == TFF_PAIR_NONBONDED || mode == TFF_PAIR_EWALD_LJ || mode == TFF_PAIR_COUL_EWALD;
  const bool lj = mode != TFF_PAIR_COUL_EWALD;
  if ((coul && !charges) || (lj && (!sigma || !epsilon)) || (mode == TFF_PAIR_LJ_TYPES && (!atom_types || ntypes <= 0))) {
    tff_set_error("tff_pair_energy: missing parameter arrays for mode %d", mode);
    return TFF_ERR_ARG;
  }
  if (!coul && grad_charges) {
    tff_set_error("tff_pair_energy: charge gradient requested for a Lennard-Jones-only mode");
    return TFF_ERR_ARG;
  }
  size_t need = tff_pair_workspace_bytes(mode, natoms, npairs, ntypes);
  if (workspace_bytes < need || !workspace) {
    tff_set_error("tff_pair_energy: workspace too small (%zu < %zu)", workspace_bytes, need);
    return TFF_ERR_WORKSPACE;
  }
  TffBox box = tff_make_box(box_host);
  PairArgs pa;
  pa.coul = coul_constant;
  pa.cutoff = cutoff;
  pa.cutoff2 = cutoff * cutoff;
  pa.alpha = ewald_alpha;
  pa.do_shift = do_shift;
  pa.ntypes = ntypes;
  // the (erfc - 1)/r exclusion correction exists only for the Ewald Coulomb modes with alpha >= 0
  if (!(mode == TFF_PAIR_COUL_EWALD || mode == TFF_PAIR_EWALD_LJ) || ewald_alpha < 0) npairs_excl = 0;
  PairWs w = carve(workspace, natoms, param_len(mode, natoms, npairs, ntypes));
  const long long* P = reinterpret_cast<const long long*>(pairs);
  const long long* PE = reinterpret_cast<const long long*>(pairs_excl);
  const long long* TY = reinterpret_cast<const long long*>(atom_types);
  if (dtype == TFF_F64)
    return run<double>(mode, (const double*)coords, natoms, box, P, npairs, PE, npairs_excl, (const double*)charges,
                       (const double*)sigma, (const double*)epsilon, TY, pa, (double*)energy_out,
                       (double*)grad_coords, (double*)grad_charges, (double*)grad_sigma, (double*)grad_epsilon,
                       (double*)forces_accum, scale, w, stream);
  if (dtype == TFF_F32)
    return run<float>(mode, (const float*)coords, natoms, box, P, npairs, PE, npairs_excl, (const float*)charges,
                      (const float*)sigma, (const float*)epsilon, TY, pa, (float*)energy_out, (float*)grad_coords,
                      (float*)grad_charges, (float*)grad_sigma, (float*)grad_epsilon, (float*)forces_accum, scale,
                      w, stream);
  tff_set_error("tff_pair_energy: unsupported dtype %d", dtype);
  return TFF_ERR_ARG;
}
