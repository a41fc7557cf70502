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
//   TFF_PAIR_COUL_EWALD_SYM  the same Coulomb sum with q_i q_j replaced by the symmetric
//                       product (a_i b_j + a_j b_i) / 2 of two charge arrays (the backward of
//                       compute_multipolar_energy_and_fields_from_atom_pairs);
//   TFF_PAIR_LJ_TYPES / TFF_PAIR_LJ_PAIRS  compute_lennard_jones_energy with a
//                       (T, T) type table or per-pair sigma/eps (reference csrc/vdw/vdw.cu:82-195).
// Rows with a negative index are padding and are skipped; pairs with
// r^2 > rc^2 contribute nothing.  Energy partial sums are float64 per
// workgroup and are reduced in a fixed order.  Gradients are accumulated as
// exact integers: each contribution v (float64) is split into h = round(v 2^24)
// and the remainder v - h 2^-24 (exact) scaled by 2^56, and both go into int64
// sums, so the result is the sum to better than float64 resolution and the same
// bit for bit on every run of the same pair list.  Global atomics cost one
// memory-side operation per 64-byte line an instruction touches, whatever the
// number of lanes on that line, so an atom's words sit in one line (x, y, z, q:
// four high and four low words) plus a 32-byte record for sigma and epsilon,
// and a wave issues them transposed: it compacts the atoms it updates into LDS
// and eight adjacent lanes write one atom's line.  Consecutive rows that share
// an atom (the i side of a cell-list pair list) are summed first by a
// segmented scan, so such a run costs one line; the j side goes through a
// per-wave LDS hash of atoms (exact integer words), so each distinct atom of a
// wave step costs one line.  The integer sums hold up to
// 2^39 per atom; contributions (run sums) of magnitude 2^30 or more
// (near-overlapping atoms) go to float64 atomics instead, so an atom's integer
// sum would need 512 such contributions to overflow.  The
// gradients of a Lennard-Jones type table (LJ_TYPES) sum over every pair of a
// type pair, far beyond that range: they are summed in float64, per wave in LDS
// (one writer per slot), per workgroup in a fixed order, and over the workgroups
// in a fixed order by a last kernel (type tables of up to TAB_MAX entries; larger
// ones take float64 atomics).  eps-gradient of a pair
// whose mixed epsilon is zero (e.g. TIP3P hydrogens) is defined as zero; the
// exact derivative of sqrt(eps_i eps_j) is unbounded there.

#include "tff_common.h"
#include "torchff_hip.h"

namespace {

constexpr int PB = 256;
constexpr int PMAXBLK = 4096;
constexpr int NCH = 6;  // per-atom gradient channels: x, y, z, charge, sigma, epsilon
constexpr double HI_SCALE = 16777216.0;            // 2^24
constexpr double LO_SCALE = 72057594037927936.0;   // 2^56
constexpr double OVF_LIMIT = 1073741824.0;         // 2^30
constexpr int TAB_MAX = 256;                       // type-table entries summed in LDS

struct PairArgs {
  double coul;      // Coulomb constant / prefactor
  double cutoff;
  double cutoff2;
  double alpha;     // Ewald alpha (< 0: bare Coulomb)
  int do_shift;
  long long ntypes;
};

__device__ __forceinline__ double wave_sum_all(double v) {
#pragma unroll
  for (int off = TFF_WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, TFF_WAVE);
  return v;
}

// Per-atom accumulator: a = 8 int64 words per atom [hi x, y, z, q, lo x, y, z, q] (one 64-byte
// line), b = 4 words [hi sigma, eps, lo sigma, eps], ovf = float64 per atom and channel (6).
struct Acc {
  long long* a;
  long long* b;
  double* ovf;
};
constexpr int ACC_WORDS = 8 + 4 + 6;  // int64-sized words per atom

// v -> (hi, lo) words; false (words zero) when |v| needs the float64 fallback
__device__ __forceinline__ bool split_words(double v, long long& h, long long& l) {
  h = 0;
  l = 0;
  if (!(fabs(v) < OVF_LIMIT)) return false;
  const double hd = rint(v * HI_SCALE);
  h = (long long)hd;
  l = __double2ll_rn((v - hd * (1.0 / HI_SCALE)) * LO_SCALE);  // remainder exact, |.| <= 2^-25
  return true;
}

// Lanes with adds set add their W words w[] into rec[key * W + c]: the lanes are compacted into the
// wave's LDS area (s_w: 64 * W words, s_k: 64 keys) and the words go out 64 / W atoms per
// instruction, W adjacent lanes per atom.  Wave-uniform control flow.
template <int W, typename K>
__device__ __forceinline__ void wave_flush(bool adds, long long key, const long long (&w)[W],
                                           long long* __restrict__ rec, long long* s_w, K* s_k) {
  const unsigned long long m = __ballot(adds);
  if (m == 0) return;
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const int n = __popcll(m);
  if (adds) {
    const int r = __popcll(m & ((1ull << lane) - 1ull));
#pragma unroll
    for (int c = 0; c < W; ++c) s_w[r * W + c] = w[c];
    s_k[r] = (K)key;
  }
  tff_wave_sync_lds();
  constexpr int KPI = TFF_WAVE / W;
  for (int base = 0; base < n; base += KPI) {
    const int idx = base + lane / W;
    if (idx < n) {
      const long long v = s_w[base * W + lane];
      if (v != 0) atomicAdd((unsigned long long*)(rec + (long long)s_k[idx] * W + (lane % W)), (unsigned long long)v);
    }
  }
  tff_wave_sync_lds();
}

// Adds v[c] (channels in chmask) of every lane with act set to atom key's accumulator (key < 0:
// nothing).  Runs of lanes with the same key (consecutive rows of one atom) are summed by a
// segmented scan and the run's last lane adds the sum.  Inactive lanes add zero and do not break
// a run.  Wave-uniform control flow.
template <int NC, typename K>
__device__ __forceinline__ void wave_accum(bool act, long long key, const double (&v)[NC], unsigned chmask,
                                           const Acc& acc, long long* s_w, K* s_k) {
  if (__ballot(act) == 0) return;
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const long long prev = __shfl_up(key, 1, TFF_WAVE);
  const long long next = __shfl_down(key, 1, TFF_WAVE);
  const bool head = lane == 0 || prev != key;
  const bool tail = lane == TFF_WAVE - 1 || next != key;
  const unsigned long long heads = __ballot(head);
  double s[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) s[c] = act ? v[c] : 0.0;
  if (heads != ~0ull) {
    // first lane of this lane's run: the highest head at or below it
    const unsigned long long below = heads & (lane == TFF_WAVE - 1 ? ~0ull : ((2ull << lane) - 1));
    const int seg = TFF_WAVE - 1 - __clzll(below);
#pragma unroll
    for (int off = 1; off < TFF_WAVE; off <<= 1) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        if (!((chmask >> c) & 1u)) continue;
        const double t = __shfl_up(s[c], off, TFF_WAVE);
        if (lane - off >= seg) s[c] += t;
      }
    }
  }
  const bool adds = tail && key >= 0;
  if (chmask & 15u) {
    long long w[8];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      w[c] = 0;
      w[4 + c] = 0;
      if (!adds || !((chmask >> c) & 1u)) continue;
      if (!split_words(s[c], w[c], w[4 + c])) atomicAdd(acc.ovf + key * NCH + c, s[c]);
      any |= (w[c] | w[4 + c]) != 0;
    }
    wave_flush<8>(any, key, w, acc.a, s_w, s_k);
  }
  if (chmask & 48u) {
    long long w[4];
    bool any = false;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      w[c] = 0;
      w[2 + c] = 0;
      if (!adds || !((chmask >> (4 + c)) & 1u)) continue;
      if (!split_words(s[4 + c], w[c], w[2 + c])) atomicAdd(acc.ovf + key * NCH + 4 + c, s[4 + c]);
      any |= (w[c] | w[2 + c]) != 0;
    }
    wave_flush<4>(any, key, w, acc.b, s_w, s_k);
  }
}

// Per-wave LDS hash of the atoms one wave step updates on the j side (a cell-list pair list repeats
// each j a few times within 64 rows): key[64] (atom index, -1 empty) and per slot HW int64 words
// (x, y, z, q high then low), exact integer sums (order independent).  At most 64 keys per step,
// so the table never fills.  The flush reuses the table's memory.  (The sigma / epsilon gradient
// channels, rarely requested, take the direct path.)
constexpr int HW = 8;
struct WaveHash {
  int* key;
  long long* w;  // 64 * HW
};

__device__ __forceinline__ void hash_clear(const WaveHash& h) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  h.key[lane] = -1;
  longlong2* wv = reinterpret_cast<longlong2*>(h.w + lane * HW);
#pragma unroll
  for (int c = 0; c < HW / 2; ++c) wv[c] = make_longlong2(0, 0);
  tff_wave_sync_lds();
}

// adds v[c] (channels in chmask) of this lane to atom key's slot; contributions too large for the
// integer words go to the float64 fallback of the atom directly
template <int NC>
__device__ __forceinline__ void hash_add(const WaveHash& h, bool act, int key, const double (&v)[NC],
                                         unsigned chmask, const Acc& acc) {
  if (!act) return;
  int p = (int)(((unsigned)key * 2654435761u) >> 26);
  int slot = -1;
  for (int t = 0; t < TFF_WAVE; ++t) {
    const int old = atomicCAS(&h.key[p], -1, key);
    if (old == -1 || old == key) {
      slot = p;
      break;
    }
    p = (p + 1) & (TFF_WAVE - 1);
  }
  if (slot < 0) {  // not reached: at most 64 keys per step
#pragma unroll
    for (int c = 0; c < NC; ++c)
      if ((chmask >> c) & 1u) atomicAdd(acc.ovf + (long long)key * NCH + c, v[c]);
    return;
  }
  long long* w = h.w + slot * HW;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (!((chmask >> c) & 1u)) continue;
    long long hi, lo;
    if (!split_words(v[c], hi, lo)) {
      atomicAdd(acc.ovf + (long long)key * NCH + c, v[c]);
      continue;
    }
    if (hi) atomicAdd((unsigned long long*)(w + c), (unsigned long long)hi);
    if (lo) atomicAdd((unsigned long long*)(w + 4 + c), (unsigned long long)lo);
  }
}

// every filled slot's words to its atom's lines (transposed flush), then the table is free
__device__ __forceinline__ void hash_flush(const WaveHash& h, unsigned chmask, const Acc& acc) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  tff_wave_sync_lds();
  const int key = h.key[lane];
  long long wa[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) wa[c] = h.w[lane * HW + c];
  tff_wave_sync_lds();  // the table's memory is reused by the flush
  bool any = false;
#pragma unroll
  for (int c = 0; c < 8; ++c) any |= wa[c] != 0;
  wave_flush<8>(key >= 0 && any, key, wa, acc.a, h.w, h.key);
}

// Adds v[c] of the lanes with act set into this wave's LDS slot tab[2 key + c] (wave-uniform
// order: the keys in order of their first lane, each key's sum by a butterfly); direct float64
// atomics into ovf[2 key + c] when the table is too large for LDS.
__device__ __forceinline__ void wave_accum_table(bool act, long long key, const double (&v)[2], unsigned chmask,
                                                 double* __restrict__ tab, double* __restrict__ ovf) {
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  unsigned long long todo = __ballot(act);
  while (todo) {
    const int leader = __ffsll((long long)todo) - 1;
    const long long k = __shfl(key, leader, TFF_WAVE);
    const bool mine = act && key == k;
    todo &= ~__ballot(mine);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (!((chmask >> c) & 1u)) continue;
      const double sum = wave_sum_all(mine ? v[c] : 0.0);
      if (lane == leader) {
        if (tab)
          tab[2 * k + c] += sum;
        else if (sum != 0.0)
          atomicAdd(ovf + 2 * k + c, sum);
      }
    }
  }
}

// chmask bits: channels of the per-atom accumulator to fill; tmask: sigma / epsilon gradients
// of the type table (LJ_TYPES), per-workgroup partials in tab_part (or float64 atomics into
// tab_ovf for tables above TAB_MAX entries); b: the second charge array of COUL_EWALD_SYM.
template <typename T, int MODE>
__global__ __launch_bounds__(PB) void k_pair_terms(const T* __restrict__ coords, const long long* __restrict__ pairs,
                                                   long long P, const long long* __restrict__ pexcl, long long PE,
                                                   TffBox box, PairArgs a, const T* __restrict__ q,
                                                   const T* __restrict__ sig, const T* __restrict__ eps,
                                                   const long long* __restrict__ types, long long natoms,
                                                   double* __restrict__ e_part, unsigned chmask, Acc atoms,
                                                   unsigned tmask, double* __restrict__ tab_part,
                                                   double* __restrict__ tab_ovf, T* __restrict__ g_sig_pairs,
                                                   T* __restrict__ g_eps_pairs) {
  constexpr int TAB_WORDS = MODE == TFF_PAIR_LJ_TYPES ? 2 * TAB_MAX : 1;
  __shared__ double s_tab[PB / TFF_WAVE][TAB_WORDS];
  __shared__ __align__(16) long long s_hash_w[PB / TFF_WAVE][TFF_WAVE * HW];
  __shared__ int s_hash_k[PB / TFF_WAVE][TFF_WAVE];
  const long long ntab = MODE == TFF_PAIR_LJ_TYPES ? a.ntypes * a.ntypes : 0;
  const bool tab_lds = MODE == TFF_PAIR_LJ_TYPES && tmask && ntab <= TAB_MAX;
  if (tab_lds) {
    for (int k = threadIdx.x; k < (PB / TFF_WAVE) * TAB_WORDS; k += PB) (&s_tab[0][0])[k] = 0.0;
    __syncthreads();
  }
  double e_acc = 0.0;
  const long long total = P + PE;
  const T rc2 = (T)a.cutoff2;
  const T inv_rc = (T)(1.0 / a.cutoff);
  const T two_over_sqrt_pi = (T)1.1283791670955126;
  const int lane = threadIdx.x & (TFF_WAVE - 1);
  const long long wave_base = blockIdx.x * (long long)PB + (threadIdx.x - lane);
  const long long stride = (long long)gridDim.x * PB;
  // every lane of a wave runs the same iterations (the accumulation is a wave-wide step)
  for (long long base = wave_base; base < total; base += stride) {
    const long long idx = base + lane;
    const bool is_excl = idx >= P;
    long long i = -1, j = -1;
    if (idx < total) {
      const long long* pr = is_excl ? pexcl + 2 * (idx - P) : pairs + 2 * idx;
      i = pr[0];
      j = pr[1];
    }
    // negative rows are padding; indices past natoms are ignored rather than dereferenced
    const bool valid = idx < total && i >= 0 && j >= 0 && i < natoms && j < natoms;
    bool act = valid;
    double vi[NCH] = {0, 0, 0, 0, 0, 0}, vj[NCH] = {0, 0, 0, 0, 0, 0}, vt[2] = {0, 0};
    long long tix = 0;
    if (act) {
      T dx = coords[3 * i] - coords[3 * j];
      T dy = coords[3 * i + 1] - coords[3 * j + 1];
      T dz = coords[3 * i + 2] - coords[3 * j + 2];
      tff_min_image<T>(box, dx, dy, dz);
      const T r2 = dx * dx + dy * dy + dz * dz;
      act = r2 <= rc2;
      if (MODE == TFF_PAIR_LJ_TYPES && act) {
        const long long ti = types[i], tj = types[j];
        act = ti >= 0 && tj >= 0 && ti < a.ntypes && tj < a.ntypes;
        tix = act ? ti * a.ntypes + tj : 0;
      }
      if (act) {
        const T rinv = (T)1 / sqrt(r2);
        const T rinv2 = rinv * rinv;
        T e = 0, dedr_over_r = 0;  // dE/dr / r, so grad_i = dedr_over_r * d
        // ---- Coulomb part
        if (MODE == TFF_PAIR_NONBONDED) {
          const T qq = q[i] * q[j] * (T)a.coul;
          const T shift = a.do_shift ? inv_rc : (T)0;
          e += qq * (rinv - shift);
          dedr_over_r += -qq * rinv * rinv2;
          const T tmp = (rinv - shift) * (T)a.coul;
          vi[3] = (double)(q[j] * tmp);
          vj[3] = (double)(q[i] * tmp);
        } else if (MODE == TFF_PAIR_EWALD_LJ || MODE == TFF_PAIR_COUL_EWALD || MODE == TFF_PAIR_COUL_EWALD_SYM) {
          T d0, d1;
          if (a.alpha >= 0) {
            const T r = r2 * rinv;
            const T u = (T)a.alpha * r;
            const T ec = erfc(u);
            const T ex = exp(-u * u) * two_over_sqrt_pi * u;
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
          // COUL_EWALD_SYM: the symmetric product of the charges q and the second array (sig)
          const T qq = (MODE == TFF_PAIR_COUL_EWALD_SYM ? (T)0.5 * (q[i] * sig[j] + sig[i] * q[j]) : q[i] * q[j]) *
                       (T)a.coul;
          e += qq * d0 * rinv;
          dedr_over_r += -qq * d1 * rinv * rinv2;
          if (MODE != TFF_PAIR_COUL_EWALD_SYM) {
            const T tmp = d0 * rinv * (T)a.coul;
            vi[3] = (double)(q[j] * tmp);
            vj[3] = (double)(q[i] * tmp);
          }
        }
        // ---- Lennard-Jones part (never on exclusion-correction pairs)
        if (MODE != TFF_PAIR_COUL_EWALD && MODE != TFF_PAIR_COUL_EWALD_SYM && !is_excl) {
          T sij, eij;
          if (MODE == TFF_PAIR_LJ_TYPES) {
            sij = sig[tix];
            eij = eps[tix];
          } else if (MODE == TFF_PAIR_LJ_PAIRS) {
            sij = sig[idx];
            eij = eps[idx];
          } else {
            sij = (sig[i] + sig[j]) * (T)0.5;
            eij = sqrt(eps[i] * eps[j]);
          }
          const T sr2 = sij * sij * rinv2;
          const T s6 = sr2 * sr2 * sr2;
          e += (T)4 * eij * s6 * (s6 - (T)1);
          dedr_over_r += -(T)24 * eij * s6 * ((T)2 * s6 - (T)1) * rinv2;
          const T dsig = (T)24 * eij * s6 * ((T)2 * s6 - (T)1) / sij;  // dE/dsigma_ij
          const T deps = (T)4 * s6 * (s6 - (T)1);                        // dE/deps_ij
          if (MODE == TFF_PAIR_LJ_TYPES) {
            vt[0] = (double)dsig;
            vt[1] = (double)deps;
          } else if (MODE == TFF_PAIR_LJ_PAIRS) {
            if (g_sig_pairs) g_sig_pairs[idx] = dsig;
            if (g_eps_pairs) g_eps_pairs[idx] = deps;
          } else {
            vi[4] = vj[4] = (double)(dsig * (T)0.5);
            if (eij > (T)0) {
              const T h = deps * (T)0.5 / eij;
              vi[5] = (double)(h * eps[j]);
              vj[5] = (double)(h * eps[i]);
            }
          }
        }
        e_acc += (double)e;
        const double gx = (double)(dedr_over_r * dx), gy = (double)(dedr_over_r * dy), gz = (double)(dedr_over_r * dz);
        vi[0] = gx;
        vi[1] = gy;
        vi[2] = gz;
        vj[0] = -gx;
        vj[1] = -gy;
        vj[2] = -gz;
      }
    }
    if (chmask) {
      // i side: runs of one atom summed by a segmented scan, flushed directly; j side: through the
      // wave's hash (both use the hash memory in turn)
      const WaveHash wh{s_hash_k[threadIdx.x / TFF_WAVE], s_hash_w[threadIdx.x / TFF_WAVE]};
      wave_accum<NCH>(act, valid ? i : -1, vi, chmask, atoms, wh.w, wh.key);
      if (chmask & 15u) {
        hash_clear(wh);
        hash_add<NCH>(wh, act && valid, (int)j, vj, chmask, atoms);
        hash_flush(wh, chmask, atoms);
      }
      if (chmask & 48u) wave_accum<NCH>(act, valid ? j : -1, vj, chmask & 48u, atoms, wh.w, wh.key);
    }
    if (MODE == TFF_PAIR_LJ_TYPES && tmask)
      wave_accum_table(act, tix, vt, tmask, tab_lds ? s_tab[threadIdx.x / TFF_WAVE] : nullptr, tab_ovf);
  }
  const double tot = tff_block_sum(e_acc);  // (synchronises the workgroup)
  if (threadIdx.x == 0) e_part[blockIdx.x] = tot;
  if (tab_lds) {
    // the workgroup's table partials: its waves' slots summed in wave order
    for (long long k = threadIdx.x; k < 2 * ntab; k += PB) {
      double v = 0.0;
      for (int wv = 0; wv < PB / TFF_WAVE; ++wv) v += s_tab[wv][k];
      tab_part[(long long)blockIdx.x * 2 * ntab + k] = v;
    }
  }
}

// type-table gradient entry k (of n = 2 ntab): the workgroup partials summed in workgroup order
// (plus the float64 atomics of oversized tables), scaled, into out[k / 2] of channel k % 2
template <typename T>
__global__ void k_tab_reduce(const double* __restrict__ tab_part, int nblocks, const double* __restrict__ tab_ovf,
                             long long n, int lds, double scale, T* __restrict__ g_sig, T* __restrict__ g_eps) {
  const long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (k >= n) return;
  double v = tab_ovf[k];
  if (lds)
    for (int b = 0; b < nblocks; ++b) v += tab_part[(long long)b * n + k];
  T* d = (k & 1) ? g_eps : g_sig;
  if (d) d[k / 2] = (T)(v * scale);
}

__global__ void k_zero_ll(long long* p, long long n) {
  long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long s = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += s) p[i] = 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ part, int n, double scale,
                                                   T* __restrict__ out) {
  double s = tff_sum_partials(part, n);
  if (threadIdx.x == 0) out[0] = (T)(s * scale);
}

// dst[k] = scale * (channel c of atom k) (accumulate: dst[k] += ...), n atoms
template <typename T>
__global__ void k_cast(Acc a, int c, long long n, double scale, int accumulate, T* __restrict__ dst,
                       int dst_stride, int dst_off) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long s = (long long)gridDim.x * blockDim.x;
  for (; k < n; k += s) {
    const long long hi = c < 4 ? a.a[8 * k + c] : a.b[4 * k + c - 4];
    const long long lo = c < 4 ? a.a[8 * k + 4 + c] : a.b[4 * k + c - 2];
    const double v = ((double)hi * (1.0 / HI_SCALE) + (double)lo * (1.0 / LO_SCALE) + a.ovf[NCH * k + c]) * scale;
    T* d = dst + k * dst_stride + dst_off;
    *d = accumulate ? (T)((double)*d + v) : (T)v;
  }
}

template <typename T>
__global__ void k_zero_t(T* __restrict__ p, long long n) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long s = (long long)gridDim.x * blockDim.x;
  for (; k < n; k += s) p[k] = (T)0;
}

template <typename T>
__global__ void k_scale(T* __restrict__ p, long long n, double scale) {
  long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  const long long s = (long long)gridDim.x * blockDim.x;
  for (; k < n; k += s) p[k] = (T)((double)p[k] * scale);
}

struct PairWs {
  double* e_part;
  Acc atoms;
  double *tab_part, *tab_ovf;
  size_t bytes;
};

PairWs carve(void* base, long long natoms, long long ntab) {
  PairWs w{};
  size_t off = 0;
  char* p = (char*)base;
  auto take = [&](size_t b) {
    char* r = p ? p + off : nullptr;
    off += tff_align(b);
    return r;
  };
  const size_t na = (size_t)(natoms > 0 ? natoms : 1), nt = (size_t)(ntab > 0 ? ntab : 1) * 2;
  w.e_part = (double*)take(sizeof(double) * PMAXBLK);
  // the a, b and ovf parts are contiguous (one clearing launch); a starts 256-byte aligned
  w.atoms.a = (long long*)take(8 * ACC_WORDS * na);
  w.atoms.b = w.atoms.a ? w.atoms.a + 8 * na : nullptr;
  w.atoms.ovf = w.atoms.a ? (double*)(w.atoms.a + 12 * na) : nullptr;
  w.tab_ovf = (double*)take(8 * nt);
  w.tab_part = (double*)take(8 * (ntab > 0 && ntab <= TAB_MAX ? (size_t)PMAXBLK * nt : 1));
  w.bytes = off;
  return w;
}

// entries of the type-table accumulator (LJ_TYPES), else 0
long long table_len(int mode, long long ntypes) { return mode == TFF_PAIR_LJ_TYPES ? ntypes * ntypes : 0; }

template <typename T>
int run(int mode, const T* coords, long long natoms, const TffBox& box, const long long* pairs, long long P,
        const long long* pexcl, long long PE, const T* q, const T* sig, const T* eps, const long long* types,
        const PairArgs& pa, T* energy, T* g_xyz, T* g_q, T* g_sig, T* g_eps, T* forces, double scale,
        const PairWs& w, hipStream_t st) {
  const long long nt = table_len(mode, pa.ntypes);
  const bool per_atom_lj = mode == TFF_PAIR_NONBONDED || mode == TFF_PAIR_EWALD_LJ;
  unsigned chmask = 0, tmask = 0;
  if (g_xyz || forces) chmask |= 7u;
  if (g_q) chmask |= 8u;
  if (per_atom_lj && g_sig) chmask |= 16u;
  if (per_atom_lj && g_eps) chmask |= 32u;
  if (mode == TFF_PAIR_LJ_TYPES) tmask = (g_sig ? 1u : 0u) | (g_eps ? 2u : 0u);
  auto zero = [&](long long* p, long long n) {
    long long b = (n + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_zero_ll, dim3((unsigned)b), dim3(256), 0, st, p, n);
  };
  if (chmask) zero(w.atoms.a, (long long)ACC_WORDS * natoms);
  if (tmask) zero((long long*)w.tab_ovf, 2 * nt);
  long long blocks = (P + PE + PB - 1) / PB;
  if (blocks > PMAXBLK) blocks = PMAXBLK;
  if (blocks < 1) blocks = 1;
  const dim3 grid((unsigned)blocks), blk(PB);
  T* gsp = mode == TFF_PAIR_LJ_PAIRS ? g_sig : nullptr;
  T* gep = mode == TFF_PAIR_LJ_PAIRS ? g_eps : nullptr;
  // per-pair gradients (LJ_PAIRS) are written only for the pairs that contribute
  if (gsp) hipLaunchKernelGGL(k_zero_t<T>, dim3(1024), dim3(256), 0, st, gsp, P);
  if (gep) hipLaunchKernelGGL(k_zero_t<T>, dim3(1024), dim3(256), 0, st, gep, P);
#define TFF_LAUNCH_PAIR(M)                                                                                       \
  hipLaunchKernelGGL((k_pair_terms<T, M>), grid, blk, 0, st, coords, pairs, P, pexcl, PE, box, pa, q, sig, eps,   \
                     types, natoms, w.e_part, chmask, w.atoms, tmask, w.tab_part, w.tab_ovf, gsp, gep)
  void* tok = tff_prof_begin("pair_terms", st);
  switch (mode) {
    case TFF_PAIR_NONBONDED: TFF_LAUNCH_PAIR(TFF_PAIR_NONBONDED); break;
    case TFF_PAIR_EWALD_LJ: TFF_LAUNCH_PAIR(TFF_PAIR_EWALD_LJ); break;
    case TFF_PAIR_COUL_EWALD: TFF_LAUNCH_PAIR(TFF_PAIR_COUL_EWALD); break;
    case TFF_PAIR_COUL_EWALD_SYM: TFF_LAUNCH_PAIR(TFF_PAIR_COUL_EWALD_SYM); break;
    case TFF_PAIR_LJ_TYPES: TFF_LAUNCH_PAIR(TFF_PAIR_LJ_TYPES); break;
    case TFF_PAIR_LJ_PAIRS: TFF_LAUNCH_PAIR(TFF_PAIR_LJ_PAIRS); break;
    default: tff_set_error("tff_pair_energy: unknown mode %d", mode); return TFF_ERR_ARG;
  }
#undef TFF_LAUNCH_PAIR
  tff_prof_end(tok, st);
  TFF_CHECK_LAUNCH("k_pair_terms");
  if (energy) hipLaunchKernelGGL(k_sum_parts<T>, dim3(1), dim3(256), 0, st, w.e_part, (int)blocks, scale, energy);
  auto cast = [&](int c, double sc, int add, T* d, int dstride, int doff) {
    long long b = (natoms + 255) / 256;
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    hipLaunchKernelGGL(k_cast<T>, dim3((unsigned)b), dim3(256), 0, st, w.atoms, c, natoms, sc, add, d, dstride, doff);
  };
  for (int c = 0; c < 3; ++c) {
    if (g_xyz) cast(c, scale, 0, g_xyz, 3, c);
    if (forces) cast(c, -scale, 1, forces, 3, c);
  }
  if (g_q) cast(3, scale, 0, g_q, 1, 0);
  if (per_atom_lj) {
    if (g_sig) cast(4, scale, 0, g_sig, 1, 0);
    if (g_eps) cast(5, scale, 0, g_eps, 1, 0);
  } else if (mode == TFF_PAIR_LJ_TYPES) {
    if (tmask)
      hipLaunchKernelGGL(k_tab_reduce<T>, dim3((unsigned)((2 * nt + 255) / 256)), dim3(256), 0, st, w.tab_part,
                         (int)blocks, w.tab_ovf, 2 * nt, nt <= TAB_MAX ? 1 : 0, scale, g_sig, g_eps);
  } else if (mode != TFF_PAIR_LJ_PAIRS) {
    // no Lennard-Jones term in this mode: its parameter gradients are zero
    if (g_sig) hipLaunchKernelGGL(k_zero_t<T>, dim3(64), dim3(256), 0, st, g_sig, natoms);
    if (g_eps) hipLaunchKernelGGL(k_zero_t<T>, dim3(64), dim3(256), 0, st, g_eps, natoms);
  } else if (scale != 1.0) {
    // per-pair gradients were written unscaled by the kernel
    if (g_sig) hipLaunchKernelGGL(k_scale<T>, dim3(1024), dim3(256), 0, st, g_sig, P, scale);
    if (g_eps) hipLaunchKernelGGL(k_scale<T>, dim3(1024), dim3(256), 0, st, g_eps, P, scale);
  }
  TFF_CHECK_LAUNCH("pair finalize");
  return TFF_OK;
}

}  // namespace

extern "C" size_t tff_pair_workspace_bytes(int mode, long long natoms, long long npairs, long long ntypes) {
  (void)npairs;
  return carve(nullptr, natoms, table_len(mode, ntypes)).bytes;
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
  const bool sym = mode == TFF_PAIR_COUL_EWALD_SYM;
  const bool coul = mode == TFF_PAIR_NONBONDED || mode == TFF_PAIR_EWALD_LJ || mode == TFF_PAIR_COUL_EWALD || sym;
  const bool lj = mode != TFF_PAIR_COUL_EWALD && !sym;
  if ((coul && !charges) || (lj && (!sigma || !epsilon)) || (sym && !sigma) ||
      (mode == TFF_PAIR_LJ_TYPES && (!atom_types || ntypes <= 0))) {
    tff_set_error("tff_pair_energy: missing parameter arrays for mode %d", mode);
    return TFF_ERR_ARG;
  }
  if ((!coul || sym) && grad_charges) {
    tff_set_error("tff_pair_energy: charge gradient not provided for mode %d", mode);
    return TFF_ERR_ARG;
  }
  if (sym && (grad_sigma || grad_epsilon)) {
    tff_set_error("tff_pair_energy: the symmetric Coulomb mode has coordinate gradients only");
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
  if (!(mode == TFF_PAIR_COUL_EWALD || mode == TFF_PAIR_EWALD_LJ || sym) || ewald_alpha < 0) npairs_excl = 0;
  PairWs w = carve(workspace, natoms, table_len(mode, ntypes));
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
