// torch.ops.torchff.* registered in C++ over the C ABI of libtorchff_hip.so.
//
// A model that calls these operators can be traced or scripted, saved, and run in a process
// that has only libtorch and this library loaded (no Python), e.g. behind OpenMM's TorchForce
// (the reference's examples/tip3p/md.py:451-520).  The schemas are the reference's
// (csrc/nblist/nblist_interface.cpp:4-9, csrc/nonbonded/nonbonded_interface.cpp:9-10,
// csrc/multipoles/multipoles_interface.cpp:6-7, csrc/vdw/vdw_interface.cpp:6,
// csrc/pme/pme_interface.cpp:6-9), plus a trailing `int order=6` on pme_long_range and the
// additive compute_nonbonded_ewald_energy_from_atom_pairs.  As in the reference, the
// differentiable energies register an AutogradCUDA kernel whose forward computes the energy and
// the gradients of every input that requires grad, and whose backward scales them
// (csrc/nonbonded/nonbonded_atom_pairs_cuda.cu:143-260, csrc/multipoles/multipoles.cu:24-152,
// csrc/vdw/vdw.cu:200-250); PME's backward runs the gather (csrc/pme/pme.cu:170-246).  A CUDA
// kernel without autograd serves inference mode.
//
// Point dipoles and quadrupoles in pme_long_range are forwarded to the operator
// torchff::_pme_long_range_multipole, which the Python package registers.
//
// The box is copied to the host on every call (its values size the cell grid and the PME
// geometry), so a box written in place by the caller between calls is always seen.

#include <ATen/ATen.h>
#include <ATen/core/dispatch/Dispatcher.h>
#include <c10/hip/HIPStream.h>
#include <torch/autograd.h>
#include <torch/library.h>

#include "torchff_hip.h"

namespace {

using at::Tensor;
using torch::autograd::AutogradContext;
using torch::autograd::variable_list;

void check(int rc, const char* what) {
  TORCH_CHECK(rc == 0, "torchff ", what, ": ", tff_last_error(), " (status ", rc, ")");
}

tff_stream_t stream_of(const Tensor& t) {
  return reinterpret_cast<tff_stream_t>(c10::hip::getCurrentHIPStream(t.device().index()).stream());
}

void require_cuda(const Tensor& t, const char* name) {
  TORCH_CHECK(!t.defined() || t.is_cuda(), "torchff (gfx950 build): ", name, " must be on the ROCm device");
}

// dtype codes of the C ABI
constexpr int kF32 = 0, kF64 = 1;
int dtype_code(const Tensor& t) {
  if (t.scalar_type() == at::kDouble) return kF64;
  TORCH_CHECK(t.scalar_type() == at::kFloat, "torchff: unsupported dtype ", t.scalar_type(),
              " (float32 or float64 expected)");
  return kF32;
}

void* ptr(const Tensor& t) { return t.defined() ? t.data_ptr() : nullptr; }

// 9 float64 values of a (3, 3) box in host memory, read on every call
struct HostBox {
  double v[9];
  bool given = false;
  const double* get() const { return given ? v : nullptr; }
};
HostBox host_box(const c10::optional<Tensor>& box) {
  HostBox h;
  if (!box.has_value() || !box->defined()) return h;
  TORCH_CHECK(box->numel() == 9, "torchff: the box must be a (3, 3) tensor");
  const Tensor b = box->detach().to(at::kCPU, at::kDouble).contiguous();
  const double* p = b.data_ptr<double>();
  for (int k = 0; k < 9; ++k) h.v[k] = p[k];
  h.given = true;
  return h;
}

Tensor workspace(size_t nbytes, const Tensor& like) {
  return at::empty({(int64_t)(nbytes > 0 ? nbytes : 1)}, like.options().dtype(at::kByte));
}

Tensor as_int64(const Tensor& t) { return t.defined() ? t.detach().to(at::kLong).contiguous() : t; }

// ---------------------------------------------------------------- neighbor list
Tensor nblist_into(const Tensor& coords_in, const c10::optional<Tensor>& box, double cutoff, const Tensor& pairs,
                   const c10::optional<Tensor>& excl_row_ptr, const c10::optional<Tensor>& excl_col,
                   bool include_self, int algorithm) {
  require_cuda(coords_in, "coords");
  require_cuda(pairs, "pairs");
  TORCH_CHECK(pairs.scalar_type() == at::kLong && pairs.dim() == 2 && pairs.size(1) == 2 && pairs.is_contiguous(),
              "pairs must be a contiguous (max_npairs, 2) int64 tensor");
  TORCH_CHECK(algorithm != TFF_NBLIST_CELL_LIST || (box.has_value() && box->defined()),
              "Cell-list neighbor list requires a periodic box");
  const Tensor coords = coords_in.detach().contiguous();
  const long long n = coords.size(0);
  const HostBox bh = host_box(box);
  Tensor rp, col;
  if (excl_row_ptr.has_value() && excl_row_ptr->defined()) {
    TORCH_CHECK(excl_col.has_value() && excl_col->defined(), "exclusion row_ptr and col must be given together");
    rp = as_int64(*excl_row_ptr);
    col = as_int64(*excl_col);
    require_cuda(rp, "excl_row_ptr");
    require_cuda(col, "excl_col_indices");
    TORCH_CHECK(rp.numel() == n + 1, "excl_row_ptr must have natoms + 1 = ", n + 1, " entries");
  }
  Tensor npairs = at::empty({1}, coords.options().dtype(at::kInt));
  const size_t nbytes = tff_nblist_workspace_bytes(n, bh.get(), cutoff, algorithm);
  Tensor ws = workspace(nbytes, coords);
  check(tff_build_neighbor_list(dtype_code(coords), coords.data_ptr(), n, bh.get(), cutoff, algorithm,
                                include_self ? 1 : 0, (const int64_t*)ptr(rp), (const int64_t*)ptr(col),
                                (int64_t*)pairs.data_ptr(), pairs.size(0), npairs.data_ptr<int32_t>(), ws.data_ptr(),
                                (size_t)ws.numel(), stream_of(coords)),
        "build_neighbor_list");
  return npairs;
}

int64_t capacity(int64_t n, const c10::Scalar& max_npairs) {
  const int64_t m = max_npairs.toLong();
  return m < 0 ? n * (n + 1) / 2 : m;
}

std::tuple<Tensor, Tensor> nblist_new(const Tensor& coords, const c10::optional<Tensor>& box,
                                      const c10::Scalar& cutoff, const c10::Scalar& max_npairs,
                                      const c10::optional<Tensor>& rp, const c10::optional<Tensor>& col,
                                      bool include_self, int algorithm) {
  if (algorithm == TFF_NBLIST_CELL_LIST)
    TORCH_CHECK(box.has_value() && box->defined(), "Cell-list neighbor list requires a periodic box");
  Tensor pairs = at::empty({capacity(coords.size(0), max_npairs), 2}, coords.options().dtype(at::kLong));
  Tensor np = nblist_into(coords, box, cutoff.toDouble(), pairs, rp, col, include_self, algorithm);
  return {pairs, np};
}

std::tuple<Tensor, Tensor> nsq(const Tensor& coords, const c10::optional<Tensor>& box, const c10::Scalar& cutoff,
                               const c10::Scalar& max_npairs, const c10::optional<Tensor>& rp,
                               const c10::optional<Tensor>& col, bool include_self) {
  return nblist_new(coords, box, cutoff, max_npairs, rp, col, include_self, TFF_NBLIST_NSQUARED);
}
Tensor nsq_out(const Tensor& coords, const c10::optional<Tensor>& box, const c10::Scalar& cutoff, Tensor& pairs,
               const c10::optional<Tensor>& rp, const c10::optional<Tensor>& col, bool include_self) {
  return nblist_into(coords, box, cutoff.toDouble(), pairs, rp, col, include_self, TFF_NBLIST_NSQUARED);
}
std::tuple<Tensor, Tensor> cl(const Tensor& coords, const c10::optional<Tensor>& box, const c10::Scalar& cutoff,
                              const c10::Scalar& max_npairs, const c10::optional<Tensor>& rp,
                              const c10::optional<Tensor>& col, bool include_self) {
  return nblist_new(coords, box, cutoff, max_npairs, rp, col, include_self, TFF_NBLIST_CELL_LIST);
}
std::tuple<Tensor, Tensor> cl_out(const Tensor& coords, const c10::optional<Tensor>& box, const c10::Scalar& cutoff,
                                  Tensor& pairs, const c10::optional<Tensor>& rp, const c10::optional<Tensor>& col,
                                  bool include_self) {
  Tensor np = nblist_into(coords, box, cutoff.toDouble(), pairs, rp, col, include_self, TFF_NBLIST_CELL_LIST);
  return {pairs, np};
}

// ---------------------------------------------------------------- pair terms
struct PairSpec {
  int mode;
  double coul, cutoff, alpha;
  bool do_shift;
  bool shape1;  // energy shaped (1,) (the nonbonded ops) rather than ()
};

// energy and, for each of coords / charges / sigma / epsilon flagged in want, dE/d(input)
Tensor pair_call(const PairSpec& s, const Tensor& coords_in, const Tensor& pairs_in, const Tensor& box,
                 const Tensor& charges, const Tensor& sigma, const Tensor& epsilon, const Tensor& atom_types,
                 const Tensor& pairs_excl, const bool want[4], Tensor grads[4], const Tensor& forces) {
  for (const Tensor* t : {&coords_in, &pairs_in, &box, &charges, &sigma, &epsilon, &atom_types, &pairs_excl, &forces})
    require_cuda(*t, "every tensor argument");
  const Tensor coords = coords_in.detach().contiguous();
  const auto dt = coords.scalar_type();
  const long long n = coords.size(0);
  const Tensor pairs = as_int64(pairs_in);
  const long long P = pairs.size(0);
  const Tensor pe = as_int64(pairs_excl);
  const long long PE = pe.defined() ? pe.size(0) : 0;
  auto prep = [&](const Tensor& x) { return x.defined() ? x.detach().to(dt).contiguous() : x; };
  const Tensor q = prep(charges), sg = prep(sigma), ep = prep(epsilon);
  const Tensor ty = as_int64(atom_types);
  const long long ntypes = (ty.defined() && sg.defined()) ? sg.size(0) : 0;
  Tensor energy = s.shape1 ? at::empty({1}, coords.options()) : at::empty({}, coords.options());
  const Tensor* refs[4] = {&coords, &q, &sg, &ep};
  for (int k = 0; k < 4; ++k) grads[k] = (want[k] && refs[k]->defined()) ? at::empty_like(*refs[k]) : Tensor();
  const HostBox bh = host_box(box);
  Tensor ws = workspace(tff_pair_workspace_bytes(s.mode, n, P, ntypes), coords);
  check(tff_pair_energy(dtype_code(coords), s.mode, coords.data_ptr(), n, bh.get(), (const int64_t*)ptr(pairs), P,
                        (const int64_t*)ptr(pe), PE, ptr(q), ptr(sg), ptr(ep), (const int64_t*)ptr(ty), ntypes,
                        s.coul, s.cutoff, s.alpha, s.do_shift ? 1 : 0, 1.0, energy.data_ptr(), ptr(grads[0]),
                        ptr(grads[1]), ptr(grads[2]), ptr(grads[3]), ptr(forces), ws.data_ptr(), (size_t)ws.numel(),
                        stream_of(coords)),
        "pair energy");
  return energy;
}

Tensor opt(const c10::optional<Tensor>& t) { return t.has_value() ? *t : Tensor(); }
c10::optional<Tensor> some(const Tensor& t) { return t.defined() ? c10::optional<Tensor>(t) : c10::nullopt; }

// The reference's autograd design: the forward computes the gradients, the backward scales them.
// Tensors that may be absent travel as optionals (an absent one is not an input of the graph).
using OptT = c10::optional<Tensor>;
struct PairEnergy : public torch::autograd::Function<PairEnergy> {
  static Tensor forward(AutogradContext* ctx, Tensor coords, OptT charges, OptT sigma, OptT epsilon, Tensor pairs,
                        Tensor box, OptT atom_types, OptT pairs_excl, int64_t mode, double coul, double cutoff,
                        double alpha, bool do_shift, bool shape1) {
    const PairSpec s{(int)mode, coul, cutoff, alpha, do_shift, shape1};
    const Tensor in[4] = {coords, opt(charges), opt(sigma), opt(epsilon)};
    bool want[4];
    for (int k = 0; k < 4; ++k) want[k] = in[k].defined() && in[k].requires_grad();
    Tensor g[4];
    Tensor e = pair_call(s, coords, pairs, box, in[1], in[2], in[3], opt(atom_types), opt(pairs_excl), want, g,
                         Tensor());
    ctx->save_for_backward({g[0], g[1], g[2], g[3]});
    return e;
  }
  static variable_list backward(AutogradContext* ctx, variable_list grad_out) {
    const auto saved = ctx->get_saved_variables();
    const Tensor s = grad_out[0].reshape({});
    variable_list out(14);
    for (int k = 0; k < 4; ++k)
      if (saved[k].defined()) out[k] = saved[k] * s.to(saved[k].scalar_type());
    return out;
  }
};

// autograd kernels
Tensor nonbonded_energy_ag(const Tensor& coords, const Tensor& pairs, const Tensor& box, const Tensor& sigma,
                           const Tensor& epsilon, const Tensor& charges, const c10::Scalar& coul,
                           const c10::Scalar& cutoff, bool do_shift) {
  return PairEnergy::apply(coords, some(charges), some(sigma), some(epsilon), pairs, box, OptT(), OptT(),
                           (int64_t)TFF_PAIR_NONBONDED, coul.toDouble(), cutoff.toDouble(), -1.0, do_shift, true);
}
Tensor nonbonded_ewald_energy_ag(const Tensor& coords, const Tensor& pairs, const Tensor& box, const Tensor& sigma,
                                 const Tensor& epsilon, const Tensor& charges, const c10::Scalar& coul,
                                 const c10::Scalar& cutoff, const c10::Scalar& alpha,
                                 const c10::optional<Tensor>& pairs_excl) {
  return PairEnergy::apply(coords, some(charges), some(sigma), some(epsilon), pairs, box, OptT(), some(opt(pairs_excl)),
                           (int64_t)TFF_PAIR_EWALD_LJ, coul.toDouble(), cutoff.toDouble(), alpha.toDouble(), false,
                           true);
}
void rank0_only(const c10::optional<Tensor>& p, const c10::optional<Tensor>& t) {
  if ((p.has_value() && p->defined()) || (t.has_value() && t->defined()))
    throw c10::NotImplementedError(
        {__func__, __FILE__, static_cast<uint32_t>(__LINE__)},
        "compute_multipolar_energy_from_atom_pairs: only rank 0 (point charges) is provided by the gfx950 build");
}
Tensor multipolar_energy_ag(const Tensor& coords, const Tensor& box, const Tensor& pairs,
                            const c10::optional<Tensor>& pairs_excl, const Tensor& q, const c10::optional<Tensor>& p,
                            const c10::optional<Tensor>& t, const c10::Scalar& cutoff, const c10::Scalar& alpha,
                            const c10::Scalar& prefactor) {
  rank0_only(p, t);
  return PairEnergy::apply(coords, some(q), OptT(), OptT(), pairs, box, OptT(), some(opt(pairs_excl)),
                           (int64_t)TFF_PAIR_COUL_EWALD, prefactor.toDouble(), cutoff.toDouble(), alpha.toDouble(),
                           false, false);
}
Tensor lj_energy_ag(const Tensor& coords, const Tensor& pairs, const Tensor& box, const Tensor& sigma,
                    const Tensor& epsilon, const c10::Scalar& cutoff, const c10::optional<Tensor>& atom_types) {
  const Tensor ty = opt(atom_types);
  return PairEnergy::apply(coords, OptT(), some(sigma), some(epsilon), pairs, box, some(ty), OptT(),
                           (int64_t)(ty.defined() ? TFF_PAIR_LJ_TYPES : TFF_PAIR_LJ_PAIRS), 0.0, cutoff.toDouble(),
                           -1.0, false, false);
}

// inference-mode kernels (no autograd key): the energy only
Tensor pair_plain(int mode, const Tensor& coords, const Tensor& pairs, const Tensor& box, const Tensor& q,
                  const Tensor& sig, const Tensor& eps, const Tensor& ty, const Tensor& pe, double coul, double cutoff,
                  double alpha, bool do_shift, bool shape1) {
  const bool want[4] = {false, false, false, false};
  Tensor g[4];
  return pair_call(PairSpec{mode, coul, cutoff, alpha, do_shift, shape1}, coords, pairs, box, q, sig, eps, ty, pe,
                   want, g, Tensor());
}
Tensor nonbonded_energy_cuda(const Tensor& coords, const Tensor& pairs, const Tensor& box, const Tensor& sigma,
                             const Tensor& epsilon, const Tensor& charges, const c10::Scalar& coul,
                             const c10::Scalar& cutoff, bool do_shift) {
  return pair_plain(TFF_PAIR_NONBONDED, coords, pairs, box, charges, sigma, epsilon, Tensor(), Tensor(),
                    coul.toDouble(), cutoff.toDouble(), -1.0, do_shift, true);
}
Tensor nonbonded_ewald_energy_cuda(const Tensor& coords, const Tensor& pairs, const Tensor& box, const Tensor& sigma,
                                   const Tensor& epsilon, const Tensor& charges, const c10::Scalar& coul,
                                   const c10::Scalar& cutoff, const c10::Scalar& alpha,
                                   const c10::optional<Tensor>& pairs_excl) {
  return pair_plain(TFF_PAIR_EWALD_LJ, coords, pairs, box, charges, sigma, epsilon, Tensor(), opt(pairs_excl),
                    coul.toDouble(), cutoff.toDouble(), alpha.toDouble(), false, true);
}
Tensor multipolar_energy_cuda(const Tensor& coords, const Tensor& box, const Tensor& pairs,
                              const c10::optional<Tensor>& pairs_excl, const Tensor& q, const c10::optional<Tensor>& p,
                              const c10::optional<Tensor>& t, const c10::Scalar& cutoff, const c10::Scalar& alpha,
                              const c10::Scalar& prefactor) {
  rank0_only(p, t);
  return pair_plain(TFF_PAIR_COUL_EWALD, coords, pairs, box, q, Tensor(), Tensor(), Tensor(), opt(pairs_excl),
                    prefactor.toDouble(), cutoff.toDouble(), alpha.toDouble(), false, false);
}
Tensor lj_energy_cuda(const Tensor& coords, const Tensor& pairs, const Tensor& box, const Tensor& sigma,
                      const Tensor& epsilon, const c10::Scalar& cutoff, const c10::optional<Tensor>& atom_types) {
  const Tensor ty = opt(atom_types);
  return pair_plain(ty.defined() ? TFF_PAIR_LJ_TYPES : TFF_PAIR_LJ_PAIRS, coords, pairs, box, Tensor(), sigma, epsilon,
                    ty, Tensor(), 0.0, cutoff.toDouble(), -1.0, false, false);
}

void nonbonded_forces(const Tensor& coords, const Tensor& pairs, const Tensor& box, const Tensor& sigma,
                      const Tensor& epsilon, const Tensor& charges, const c10::Scalar& coul, const c10::Scalar& cutoff,
                      const Tensor& forces) {
  TORCH_CHECK(forces.scalar_type() == coords.scalar_type() && forces.is_contiguous() &&
                  forces.sizes() == coords.sizes(),
              "forces must be a contiguous tensor shaped and typed like coords");
  const bool want[4] = {false, false, false, false};
  Tensor g[4];
  pair_call(PairSpec{TFF_PAIR_NONBONDED, coul.toDouble(), cutoff.toDouble(), -1.0, true, true}, coords, pairs, box,
            charges, sigma, epsilon, Tensor(), Tensor(), want, g, forces);
}

// ------------------------------------------------- rank-0 energy with per-atom potential
// compute_multipolar_energy_and_fields_from_atom_pairs for point charges (reference
// csrc/multipoles/multipoles_field.cu, torchff/multipoles.py:276-291): the energy E (shape (1,)),
// the potential phi = dE/dq and the field, zero for point charges as in the reference's CUDA op.
// E = 1/2 q^T A q with A_ij = pref T(r_ij) (erfc or bare kernel, the exclusion corrections
// included), phi = A q.  For upstream gradients gE (of E) and w (of phi), L = gE E + w . phi:
// dL/dq = A (gE q + w), the potential of the charges gE q + w; and
// dL/dr = grad_r of 1/2 q^T A b with b = gE q + 2 w held fixed, the symmetric bilinear Coulomb
// sum of q and b (TFF_PAIR_COUL_EWALD_SYM).
struct FieldSpec {
  double cutoff, alpha, pref;
};

Tensor coulomb_potential(const Tensor& coords, const Tensor& box, const Tensor& pairs, const Tensor& pe,
                         const Tensor& charges, const FieldSpec& f, Tensor* energy) {
  const bool want[4] = {false, true, false, false};
  Tensor g[4];
  Tensor e = pair_call(PairSpec{TFF_PAIR_COUL_EWALD, f.pref, f.cutoff, f.alpha, false, true}, coords, pairs, box,
                       charges, Tensor(), Tensor(), Tensor(), pe, want, g, Tensor());
  if (energy) *energy = e;
  return g[1];
}

struct MultipoleFields : public torch::autograd::Function<MultipoleFields> {
  static variable_list forward(AutogradContext* ctx, Tensor coords, Tensor q, Tensor box, Tensor pairs,
                               OptT pairs_excl, double cutoff, double alpha, double pref) {
    const FieldSpec f{cutoff, alpha, pref};
    Tensor e;
    const Tensor pot = coulomb_potential(coords, box, pairs, opt(pairs_excl), q, f, &e);
    Tensor field = at::zeros({coords.size(0), 3}, coords.options());
    ctx->mark_non_differentiable({field});
    ctx->save_for_backward({coords.detach(), q.detach(), box, pairs, opt(pairs_excl)});
    ctx->saved_data["cutoff"] = cutoff;
    ctx->saved_data["alpha"] = alpha;
    ctx->saved_data["pref"] = pref;
    return {e, pot, field};
  }
  static variable_list backward(AutogradContext* ctx, variable_list grad_out) {
    const auto sv = ctx->get_saved_variables();
    const Tensor &c = sv[0], &q = sv[1], &box = sv[2], &pairs = sv[3], &pe = sv[4];
    const FieldSpec f{ctx->saved_data["cutoff"].toDouble(), ctx->saved_data["alpha"].toDouble(),
                      ctx->saved_data["pref"].toDouble()};
    const auto dt = c.scalar_type();
    const Tensor gE = grad_out[0].defined() ? grad_out[0].reshape({}).to(dt) : at::zeros({}, c.options());
    const Tensor w = grad_out[1].defined() ? grad_out[1].to(dt).contiguous() : at::zeros_like(q);
    variable_list out(8);
    if (ctx->needs_input_grad(1)) out[1] = coulomb_potential(c, box, pairs, pe, q * gE + w, f, nullptr);
    if (ctx->needs_input_grad(0)) {
      const Tensor b = (q * gE + w * 2.0).contiguous();
      const bool want[4] = {true, false, false, false};
      Tensor g[4];
      pair_call(PairSpec{TFF_PAIR_COUL_EWALD_SYM, f.pref, f.cutoff, f.alpha, false, true}, c, pairs, box, q, b,
                Tensor(), Tensor(), pe, want, g, Tensor());
      out[0] = g[0];
    }
    return out;
  }
};

std::tuple<Tensor, Tensor, Tensor> multipolar_fields_ag(const Tensor& coords, const Tensor& box, const Tensor& pairs,
                                                        const c10::optional<Tensor>& pairs_excl, const Tensor& q,
                                                        const c10::optional<Tensor>& p, const c10::optional<Tensor>& t,
                                                        const c10::Scalar& cutoff, const c10::Scalar& alpha,
                                                        const c10::Scalar& prefactor) {
  rank0_only(p, t);
  auto r = MultipoleFields::apply(coords, q, box, pairs, some(opt(pairs_excl)), cutoff.toDouble(), alpha.toDouble(),
                                  prefactor.toDouble());
  return {r[0], r[1], r[2]};
}
std::tuple<Tensor, Tensor, Tensor> multipolar_fields_cuda(const Tensor& coords, const Tensor& box, const Tensor& pairs,
                                                          const c10::optional<Tensor>& pairs_excl, const Tensor& q,
                                                          const c10::optional<Tensor>& p,
                                                          const c10::optional<Tensor>& t, const c10::Scalar& cutoff,
                                                          const c10::Scalar& alpha, const c10::Scalar& prefactor) {
  rank0_only(p, t);
  Tensor e;
  const Tensor pot = coulomb_potential(coords, box, pairs, opt(pairs_excl), q,
                                       FieldSpec{cutoff.toDouble(), alpha.toDouble(), prefactor.toDouble()}, &e);
  return {e, pot, at::zeros({coords.size(0), 3}, coords.options())};
}

// ---------------------------------------------------------------- PME (point charges)
struct PmeGeometry {
  int K[3];
  Tensor xm, ym, zm;
  int zhalf;
};
PmeGeometry pme_geometry(const Tensor& coords, int64_t K, const Tensor& xm, const Tensor& ym, const Tensor& zm) {
  PmeGeometry g;
  for (int a = 0; a < 3; ++a) g.K[a] = (int)K;
  const auto dt = coords.scalar_type();
  g.xm = xm.detach().to(dt).contiguous();
  g.ym = ym.detach().to(dt).contiguous();
  g.zm = zm.detach().to(dt).contiguous();
  TORCH_CHECK(g.xm.numel() == K && g.ym.numel() == K && (g.zm.numel() == K || g.zm.numel() == K / 2 + 1),
              "pme_long_range: moduli lengths must be (K, K, K//2+1) or (K, K, K)");
  g.zhalf = (g.zm.numel() != K || K == 1) ? 1 : 0;
  return g;
}

Tensor pme_energy(const Tensor& coords_in, const Tensor& box, const Tensor& q_in, int64_t K, double alpha,
                  const Tensor& xm, const Tensor& ym, const Tensor& zm, int64_t order, Tensor* phi_out,
                  Tensor* coords_out, Tensor* q_out) {
  for (const Tensor* t : {&coords_in, &box, &q_in, &xm, &ym, &zm}) require_cuda(*t, "every tensor argument");
  const Tensor c = coords_in.detach().contiguous();
  const Tensor q = q_in.detach().to(c.scalar_type()).contiguous();
  const PmeGeometry g = pme_geometry(c, K, xm, ym, zm);
  const HostBox bh = host_box(box);
  const long long n = c.size(0);
  Tensor energy = at::empty({}, c.options());
  Tensor phi = at::empty({g.K[0], g.K[1], g.K[2]}, c.options());
  Tensor ws = workspace(tff_pme_workspace_bytes(dtype_code(c), n, g.K[0], g.K[1], g.K[2]), c);
  check(tff_pme_energy(dtype_code(c), c.data_ptr(), q.data_ptr(), n, bh.get(), g.K[0], g.K[1], g.K[2], (int)order,
                       alpha, g.xm.data_ptr(), g.ym.data_ptr(), g.zm.data_ptr(), g.zhalf, 1.0, energy.data_ptr(),
                       phi.data_ptr(), ws.data_ptr(), (size_t)ws.numel(), stream_of(c)),
        "pme_long_range");
  if (phi_out) *phi_out = phi;
  if (coords_out) *coords_out = c;
  if (q_out) *q_out = q;
  return energy;
}

struct PmeEnergy : public torch::autograd::Function<PmeEnergy> {
  static Tensor forward(AutogradContext* ctx, Tensor coords, Tensor q, Tensor box, int64_t K, double alpha,
                        Tensor xm, Tensor ym, Tensor zm, int64_t order) {
    Tensor phi, c, qq;
    Tensor e = pme_energy(coords, box, q, K, alpha, xm, ym, zm, order, &phi, &c, &qq);
    ctx->save_for_backward({c, qq, phi, box});
    ctx->saved_data["K"] = K;
    ctx->saved_data["order"] = order;
    ctx->saved_data["alpha"] = alpha;
    ctx->saved_data["need_c"] = coords.requires_grad();
    ctx->saved_data["need_q"] = q.requires_grad();
    return e;
  }
  static variable_list backward(AutogradContext* ctx, variable_list grad_out) {
    const auto sv = ctx->get_saved_variables();
    const Tensor &c = sv[0], &qq = sv[1], &phi = sv[2], &box = sv[3];
    const int K = (int)ctx->saved_data["K"].toInt();
    const int order = (int)ctx->saved_data["order"].toInt();
    const double alpha = ctx->saved_data["alpha"].toDouble();
    Tensor dc = ctx->saved_data["need_c"].toBool() ? at::empty_like(c) : Tensor();
    Tensor dq = ctx->saved_data["need_q"].toBool() ? at::empty_like(qq) : Tensor();
    const HostBox bh = host_box(box);
    check(tff_pme_gather(dtype_code(c), phi.data_ptr(), c.data_ptr(), qq.data_ptr(), c.size(0), bh.get(), K, K, K,
                         order, alpha, 1.0, ptr(dq), ptr(dc), stream_of(c)),
          "pme_long_range backward");
    const Tensor s = grad_out[0].reshape({}).to(c.scalar_type());
    variable_list out(9);
    if (dc.defined()) out[0] = dc * s;
    if (dq.defined()) out[1] = dq * s;
    return out;
  }
};

bool has(const c10::optional<Tensor>& t) { return t.has_value() && t->defined(); }

Tensor pme_multipole_via_python(const Tensor& coords, const Tensor& box, const Tensor& q,
                                const c10::optional<Tensor>& p, const c10::optional<Tensor>& t, const c10::Scalar& K,
                                const c10::Scalar& alpha, const Tensor& xm, const Tensor& ym, const Tensor& zm,
                                int64_t order) {
  auto h = c10::Dispatcher::singleton().findSchema({"torchff::_pme_long_range_multipole", ""});
  TORCH_CHECK(h.has_value(), "pme_long_range: point dipoles and quadrupoles are served by the torchff Python "
              "package (import torchff); this process has only the point-charge operator");
  return h->typed<Tensor(const Tensor&, const Tensor&, const Tensor&, const c10::optional<Tensor>&,
                         const c10::optional<Tensor>&, const c10::Scalar&, const c10::Scalar&, const Tensor&,
                         const Tensor&, const Tensor&, int64_t)>()
      .call(coords, box, q, p, t, K, alpha, xm, ym, zm, order);
}

Tensor pme_long_range_ag(const Tensor& coords, const Tensor& box, const Tensor& q, const c10::optional<Tensor>& p,
                         const c10::optional<Tensor>& t, const c10::Scalar& K, const c10::Scalar& alpha,
                         const Tensor& xm, const Tensor& ym, const Tensor& zm, int64_t order) {
  if (has(p) || has(t)) return pme_multipole_via_python(coords, box, q, p, t, K, alpha, xm, ym, zm, order);
  return PmeEnergy::apply(coords, q, box, K.toLong(), alpha.toDouble(), xm, ym, zm, order);
}
Tensor pme_long_range_cuda(const Tensor& coords, const Tensor& box, const Tensor& q, const c10::optional<Tensor>& p,
                           const c10::optional<Tensor>& t, const c10::Scalar& K, const c10::Scalar& alpha,
                           const Tensor& xm, const Tensor& ym, const Tensor& zm, int64_t order) {
  if (has(p) || has(t)) return pme_multipole_via_python(coords, box, q, p, t, K, alpha, xm, ym, zm, order);
  return pme_energy(coords, box, q, K.toLong(), alpha.toDouble(), xm, ym, zm, order, nullptr, nullptr, nullptr);
}

}  // namespace

TORCH_LIBRARY_FRAGMENT(torchff, m) {
  m.def("build_neighbor_list_nsquared(Tensor coords, Tensor? box, Scalar cutoff, Scalar max_npairs, "
        "Tensor? excl_row_ptr, Tensor? excl_col_indices, bool include_self) -> (Tensor, Tensor)");
  m.def("build_neighbor_list_nsquared_out(Tensor coords, Tensor? box, Scalar cutoff, Tensor(a!) pairs, "
        "Tensor? excl_row_ptr, Tensor? excl_col_indices, bool include_self) -> Tensor");
  m.def("build_neighbor_list_cell_list(Tensor coords, Tensor? box, Scalar cutoff, Scalar max_npairs, "
        "Tensor? excl_row_ptr, Tensor? excl_col_indices, bool include_self) -> (Tensor, Tensor)");
  m.def("build_neighbor_list_cell_list_out(Tensor coords, Tensor? box, Scalar cutoff, Tensor(a!) pairs, "
        "Tensor? excl_row_ptr, Tensor? excl_col_indices, bool include_self) -> (Tensor, Tensor)");
  m.def("compute_nonbonded_energy_from_atom_pairs(Tensor coords, Tensor pairs, Tensor box, Tensor sigma, "
        "Tensor epsilon, Tensor charges, Scalar coul_constant, Scalar cutoff, bool do_shift) -> Tensor");
  m.def("compute_nonbonded_forces_from_atom_pairs(Tensor coords, Tensor pairs, Tensor box, Tensor sigma, "
        "Tensor epsilon, Tensor charges, Scalar coul_constant, Scalar cutoff, Tensor(a!) forces) -> ()");
  m.def("compute_nonbonded_ewald_energy_from_atom_pairs(Tensor coords, Tensor pairs, Tensor box, "
        "Tensor sigma, Tensor epsilon, Tensor charges, Scalar coul_constant, Scalar cutoff, "
        "Scalar ewald_alpha, Tensor? pairs_excl) -> Tensor");
  m.def("compute_multipolar_energy_from_atom_pairs(Tensor coords, Tensor box, Tensor pairs, "
        "Tensor? pairs_excl, Tensor q, Tensor? p, Tensor? t, Scalar cutoff, Scalar ewald_alpha, "
        "Scalar prefactor) -> Tensor");
  m.def("compute_multipolar_energy_and_fields_from_atom_pairs(Tensor coords, Tensor box, Tensor pairs, "
        "Tensor? pairs_excl, Tensor q, Tensor? p, Tensor? t, Scalar cutoff, Scalar ewald_alpha, "
        "Scalar prefactor) -> (Tensor, Tensor, Tensor)");
  m.def("compute_lennard_jones_energy(Tensor coords, Tensor pairs, Tensor box, Tensor sigma, "
        "Tensor epsilon, Scalar cutoff, Tensor? atom_types) -> Tensor");
  m.def("pme_long_range(Tensor coords, Tensor box, Tensor q, Tensor? p, Tensor? t, Scalar K, Scalar alpha, "
        "Tensor xmoduli, Tensor ymoduli, Tensor zmoduli, int order=6) -> Tensor");
}

TORCH_LIBRARY_IMPL(torchff, CUDA, m) {
  m.impl("build_neighbor_list_nsquared", &nsq);
  m.impl("build_neighbor_list_nsquared_out", &nsq_out);
  m.impl("build_neighbor_list_cell_list", &cl);
  m.impl("build_neighbor_list_cell_list_out", &cl_out);
  m.impl("compute_nonbonded_forces_from_atom_pairs", &nonbonded_forces);
  m.impl("compute_nonbonded_energy_from_atom_pairs", &nonbonded_energy_cuda);
  m.impl("compute_nonbonded_ewald_energy_from_atom_pairs", &nonbonded_ewald_energy_cuda);
  m.impl("compute_multipolar_energy_from_atom_pairs", &multipolar_energy_cuda);
  m.impl("compute_multipolar_energy_and_fields_from_atom_pairs", &multipolar_fields_cuda);
  m.impl("compute_lennard_jones_energy", &lj_energy_cuda);
  m.impl("pme_long_range", &pme_long_range_cuda);
}

TORCH_LIBRARY_IMPL(torchff, AutogradCUDA, m) {
  m.impl("compute_nonbonded_energy_from_atom_pairs", &nonbonded_energy_ag);
  m.impl("compute_nonbonded_ewald_energy_from_atom_pairs", &nonbonded_ewald_energy_ag);
  m.impl("compute_multipolar_energy_from_atom_pairs", &multipolar_energy_ag);
  m.impl("compute_multipolar_energy_and_fields_from_atom_pairs", &multipolar_fields_ag);
  m.impl("compute_lennard_jones_energy", &lj_energy_ag);
  m.impl("pme_long_range", &pme_long_range_ag);
}
