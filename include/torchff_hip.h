/*
 * torchff_hip.h — C ABI of the MI355X (gfx950) nonbonded hot path.
 *
 * Plain pointers and sizes only.  All array arguments are DEVICE pointers
 * unless the name ends in _host; every call enqueues work on `stream` and
 * returns without synchronising.  Kernels never allocate: each entry point
 * takes a caller-provided device workspace whose size the matching
 * *_workspace_bytes() query returns (so the calls are safe to capture in a
 * HIP graph).  Return value 0 means success; otherwise tff_last_error()
 * describes the failure (thread-local string).
 *
 * dtype: 0 = float32, 1 = float64 (coordinates, charges, parameters and
 * outputs share it).  Index arrays are int64.  box_host: 9 float64 values,
 * row-major lattice vectors (rows a, b, c) in host memory, or NULL for no
 * periodicity where allowed.
 *
 * Each entry point replaces one operator of the reference library
 * (THGLab/torchff-lib), cited by file:line.
 */
#ifndef TORCHFF_HIP_H
#define TORCHFF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* the HIP stream type (hipStream_t) */
typedef struct ihipStream_t* tff_stream_t;

/* ------------------------------------------------------------------ misc */
const char* tff_last_error(void);
int tff_abi_version(void);

/* Kernel timing hooks (no counterpart in the reference; instrumentation for
 * benchmarks).  When enabled, the main kernel of each entry point is bracketed
 * by HIP events on the caller's stream; tff_prof_collect waits for and sums the
 * recorded launches whose name matches (NULL: all), then forgets them.
 * Names: "nblist_cell_pairs", "nblist_nsquared", "pair_terms", "pme_spread",
 * "pme_fft", "pme_fft_conv", "pme_gather", "fused_cluster_pairs", "fused_tiles",
 * "fused_spread", "fused_fft", "fused_gather". */
void tff_prof_enable(int on);
int tff_prof_collect(const char* name, double* total_ms, long long* count);
/* Schedule of the reciprocal-space chain of tff_fused_energy_forces (no counterpart in the
 * reference; results are identical in every mode).  0: on the caller's stream after the cluster
 * build, in order (a timing aid: its stages can be timed alone); 1 (default): by size, as 3 below
 * 300 000 atoms and as 2 from there on; 2: the charge spread on the caller's stream before the
 * pair search, then the FFT and the gather on an auxiliary stream concurrently with the pair
 * search and pair kernel; 3: the whole chain on the auxiliary stream.  Other values mean 1. */
void tff_fused_set_overlap(int mode);
/* Graph replay of tff_fused_energy_forces (no counterpart in the reference; results are identical
 * either way).  The launches of a call up to its last kernel are captured into a HIP graph per
 * device, input and workspace addresses, box and parameters when two calls in a row on a device
 * have the same such key (up to 8 graphs kept per device, the least recently used dropped after a
 * device synchronisation), and replayed by later calls with that key; the last kernel, which
 * writes forces_out, energies_out and the status words, is launched per call.  mode 1 (default): for float32 arithmetic below 300 000 atoms (where replay
 * measured faster); 2: always; 0: never.  A call whose stream is being captured, or made while
 * tff_prof_enable(1) is on, launches its kernels one by one.  Other values mean 1. */
void tff_fused_set_graphs(int mode);

/* --------------------------------------------------------- neighbor list */
#define TFF_NBLIST_NSQUARED 0
#define TFF_NBLIST_CELL_LIST 1

size_t tff_nblist_workspace_bytes(long long natoms, const double* box_host, double cutoff, int algorithm);

/* Replaces torch.ops.torchff.build_neighbor_list_nsquared(_out) and
 * build_neighbor_list_cell_list(_out)
 * (csrc/nblist/nblist_interface.cpp:5-8; kernels csrc/nblist/nblist_nsquared_cuda.cu:29-198,
 * csrc/nblist/nblist_clist_cuda.cu:44-525; Python caller torchff/nblist.py:142-160).
 * pairs: (max_npairs, 2) int64 output, rows (a, b) with a > b, unused rows -1.
 * npairs: one int32, the number of pairs found (may exceed max_npairs: then
 * only the first max_npairs were stored).  excl_row_ptr / excl_col: CSR
 * exclusions (torchff/nblist.py:50-123) or both NULL.  The cell-list
 * algorithm requires box_host.  Pair predicate: float64 |minimg(dr)|^2 <= cutoff^2. */
int tff_build_neighbor_list(int dtype, const void* coords, long long natoms, const double* box_host,
                            double cutoff, int algorithm, int include_self, const int64_t* excl_row_ptr,
                            const int64_t* excl_col, int64_t* pairs, long long max_npairs, int32_t* npairs,
                            void* workspace, size_t workspace_bytes, tff_stream_t stream);

/* ------------------------------------------------------------ pair terms */
#define TFF_PAIR_NONBONDED 0   /* shifted Coulomb + LB Lennard-Jones */
#define TFF_PAIR_EWALD_LJ 1    /* erfc Coulomb + LB Lennard-Jones + exclusion correction */
#define TFF_PAIR_COUL_EWALD 2  /* erfc Coulomb (rank-0 multipoles) + exclusion correction */
#define TFF_PAIR_LJ_TYPES 3    /* Lennard-Jones from a (T, T) type table */
#define TFF_PAIR_LJ_PAIRS 4    /* Lennard-Jones from per-pair sigma / epsilon */
#define TFF_PAIR_COUL_EWALD_SYM 5 /* COUL_EWALD with q_i q_j replaced by (a_i b_j + a_j b_i) / 2:
                                     a = charges, b = the sigma argument; grad_coords only */

size_t tff_pair_workspace_bytes(int mode, long long natoms, long long npairs, long long ntypes);

/* Replaces torch.ops.torchff.compute_nonbonded_energy_from_atom_pairs and
 * compute_nonbonded_forces_from_atom_pairs (csrc/nonbonded/nonbonded_interface.cpp:9-10,
 * kernel csrc/nonbonded/nonbonded_atom_pairs_cuda.cu:16-305) in mode NONBONDED;
 * compute_multipolar_energy_from_atom_pairs for rank 0
 * (csrc/multipoles/multipoles_interface.cpp:6, csrc/multipoles/multipoles.cu:24-152) in mode
 * COUL_EWALD; compute_lennard_jones_energy (csrc/vdw/vdw_interface.cpp:6, csrc/vdw/vdw.cu:82-195)
 * in modes LJ_TYPES / LJ_PAIRS; and the fused erfc-Coulomb + LJ extension in mode EWALD_LJ.
 *
 * Outputs (any may be NULL): energy_out (1 value) = scale * E;
 * grad_* = scale * dE/d(argument), written (not accumulated), accumulated as exact integer
 * sums (bitwise the same on every run of the same pair list); for LJ_TYPES
 * grad_sigma / grad_epsilon have ntypes^2 entries, for LJ_PAIRS npairs entries,
 * otherwise natoms; forces_accum += -scale * dE/dcoords.
 * pairs_excl are used only by the Ewald modes with ewald_alpha >= 0. */
int tff_pair_energy(int dtype, int mode, const void* coords, long long natoms, const double* box_host,
                    const int64_t* pairs, long long npairs, const int64_t* pairs_excl, long long npairs_excl,
                    const void* charges, const void* sigma, const void* epsilon, const int64_t* atom_types,
                    long long ntypes, double coul_constant, double cutoff, double ewald_alpha, int do_shift,
                    double scale, void* energy_out, void* grad_coords, void* grad_charges, void* grad_sigma,
                    void* grad_epsilon, void* forces_accum, void* workspace, size_t workspace_bytes,
                    tff_stream_t stream);

/* ------------------------------------------------------------------- PME */
size_t tff_pme_workspace_bytes(int dtype, long long natoms, int K0, int K1, int K2);

/* Replaces the forward of torch.ops.torchff.pme_long_range for point charges
 * (csrc/pme/pme_interface.cpp:6-9, csrc/pme/pme.cu:19-168), any spline order 3..8.
 * Writes energy_out = scale * (E_recip + E_self) and the potential grid
 * phi_grid_out (K0*K1*K2 values, z fastest) that tff_pme_gather consumes.
 * xmoduli (K0), ymoduli (K1), zmoduli (K2, or K2/2+1 when zmoduli_half) hold
 * |theta(k)| per axis (torchff/bsplines.py:57-107 for order 6). */
int tff_pme_energy(int dtype, const void* coords, const void* charges, long long natoms, const double* box_host,
                   int K0, int K1, int K2, int order, double alpha, const void* xmoduli, const void* ymoduli,
                   const void* zmoduli, int zmoduli_half, double scale, void* energy_out, void* phi_grid_out,
                   void* workspace, size_t workspace_bytes, tff_stream_t stream);

/* Replaces the backward of pme_long_range (csrc/pme/pme.cu:170-246, interpolate_kernel
 * csrc/pme/kernels.cuh:191-470): epot_out = scale * dE/dq (the potential including the
 * self term), dcoords_out = scale * dE/dcoords.  Either output may be NULL. */
int tff_pme_gather(int dtype, const void* phi_grid, const void* coords, const void* charges, long long natoms,
                   const double* box_host, int K0, int K1, int K2, int order, double alpha, double scale,
                   void* epot_out, void* dcoords_out, tff_stream_t stream);

/* Multipolar form of tff_pme_energy: replaces pme_long_range / pme_long_range_all for ranks 1
 * and 2 (csrc/pme/pme.cu:19-168,251-414; spread_q_kernel<RANK> csrc/pme/kernels.cuh:12-187;
 * Python torchff/pme_ref.py:145-235,425-469).  dipoles (N x 3) and quadrupoles (N x 3 x 3,
 * row-major, taken as given: the self terms are those of traceless quadrupoles) may be NULL,
 * and so may charges when dipoles or quadrupoles are given.  Each atom adds
 * q W + p . grad W + (1/3) theta : grad grad W to the charge grid; energy_out = scale * (E_recip
 * + E_self) with E_self = -a/sqrt(pi) sum (q^2 + (2a^2/3)|p|^2 + (8a^4/45) theta:theta).
 * Quadrupoles need spline order >= 4.  Same workspace as tff_pme_energy. */
int tff_pme_energy_multipole(int dtype, const void* coords, const void* charges, const void* dipoles,
                             const void* quadrupoles, long long natoms, const double* box_host, int K0, int K1, int K2,
                             int order, double alpha, const void* xmoduli, const void* ymoduli, const void* zmoduli,
                             int zmoduli_half, double scale, void* energy_out, void* phi_grid_out, void* workspace,
                             size_t workspace_bytes, tff_stream_t stream);

/* Interpolation for multipoles (csrc/pme/kernels.cuh:191-470 interpolate_kernel<RANK>,
 * csrc/pme/kernels_with_field.cuh interpolate_kernel_with_field), all outputs scaled by scale
 * and each may be NULL:
 *   epot_out (N)          = dE/dq  = phi - 2a q/sqrt(pi)
 *   field_out (N x 3)     = -dE/dp = -grad phi + 4a^3/(3 sqrt(pi)) p
 *   field_grad_out (N x 9)= -3 dE/dtheta = -grad grad phi + 16a^5/(15 sqrt(pi)) theta
 *   dcoords_out (N x 3)   = dE/dr  = q grad phi + (grad grad phi) p + (1/3) theta : grad^3 phi
 * for the potential grid phi_grid of tff_pme_energy(_multipole) and the multipoles given here
 * (which may differ from those that made the grid: the backward of pme_long_range_all contracts
 * one grid with two multipole sets). */
int tff_pme_gather_multipole(int dtype, const void* phi_grid, const void* coords, const void* charges,
                             const void* dipoles, const void* quadrupoles, long long natoms, const double* box_host,
                             int K0, int K1, int K2, int order, double alpha, double scale, void* epot_out,
                             void* field_out, void* field_grad_out, void* dcoords_out, tff_stream_t stream);

/* ------------------------------------------------------- direct Ewald */
size_t tff_ewald_workspace_bytes(int dtype, long long natoms, int kmax);

/* Replaces torch.ops.torchff.ewald_long_range and ewald_long_range_all for point charges
 * (rank 0; csrc/ewald/ewald_interface.cpp:6-11, csrc/ewald/ewald.cu:9-171,304-420; Python
 * torchff/ewald.py:60-110): the FFT-free reciprocal sum over the half-space integer vectors
 * |m_a| <= kmax (1 <= kmax <= 31), k = B^-1 m, G = exp(-pi^2 k^2/alpha^2)/k^2.  energy_out (device,
 * 1 double) = 1/(pi V) sum G |S|^2 - alpha/sqrt(pi) sum q^2 (Gaussian units); potential_out
 * (dtype, N) = dE/dq including the self term; forces_out (dtype, N x 3) = -dE/dr.  Either of the
 * last two may be NULL.  k = B^-1 m is the reciprocal lattice of row-major lattice vectors; the
 * reference forms m B^-1, the same for orthorhombic boxes. */
int tff_ewald(int dtype, const void* coords, const void* charges, long long natoms, const double* box_host, int kmax,
              double alpha, double* energy_out, void* potential_out, void* forces_out, void* workspace,
              size_t workspace_bytes, tff_stream_t stream);

/* Multipolar form of tff_ewald (ranks 1, 2: csrc/ewald/ewald.cu:9-610, torchff/ewald.py:60-110):
 * L(k) = q - (4 pi^2/3) k.theta.k + i 2 pi k.p, S = sum_j L_j exp(2 pi i k.r_j).  Source multipoles
 * charges (N), dipoles (N x 3), quadrupoles (N x 3 x 3 row-major), each may be NULL (zero).
 * energy_out (device double) = 1/(pi V) sum G |S|^2 - a/sqrt(pi) sum (q^2 + (2a^2/3)|p|^2
 * + (8a^4/45) theta:theta); potential_out = dE/dq, field_out = -dE/dp, field_grad_out (N x 9) =
 * -3 dE/dtheta, all with the self terms; forces_out (N x 3) = -dE/dr, or, when any of c_charges /
 * c_dipoles / c_quadrupoles is given, the force of the set X = (c_*) in the potential of the source
 * set: 4/V sum_k k Im(L^X_i A_i), A_i = G S* exp(2 pi i k.r_i).  Each output may be NULL. */
int tff_ewald_multipole(int dtype, const void* coords, long long natoms, const double* box_host, int kmax,
                        double alpha, const void* charges, const void* dipoles, const void* quadrupoles,
                        const void* c_charges, const void* c_dipoles, const void* c_quadrupoles, double* energy_out,
                        void* potential_out, void* field_out, void* field_grad_out, void* forces_out, void* workspace,
                        size_t workspace_bytes, tff_stream_t stream);

/* ------------------------------------------------------- cluster pairs */
/* Cluster-pair neighbour list in the reference's format (torchff/nblist.py:260-344
 * build_cluster_pairs / decode_cluster_pairs; csrc/nblist/nblist_cluster_pairs_cuda.cu:18-720) and
 * the nonbonded terms evaluated on it (the schemas commented at
 * csrc/nonbonded/nonbonded_interface.cpp:5-8).  Clusters are 32 consecutive slots of
 * sorted_atom_indices; cluster_exclusions (2 x E int32, row 0 = x, row 1 = y, x <= y) lists every
 * diagonal block and every block holding an excluded pair, with bitmask_exclusions (E x 32 uint32:
 * word [e][l] bit k set = skip slots (32x + l, 32y + k)); interacting_clusters (M int32) and
 * interacting_atoms (M x 32 int32, -1 padded) list, for each cluster x, the atoms of the clusters
 * y > x outside its exclusion set that lie within the cutoff of its bounding sphere.  Decoding
 * visits every non-excluded pair within the cutoff once.  Exclusions are a symmetric CSR over
 * atoms (int64, at most 16 partners per atom); NULL means none.  cutoff must not exceed half the
 * smallest box height.  Two calls: _build writes (E, M, flags) to counts_out (device, 3 int64;
 * flags: 1 = an atom has more than 16 exclusion partners, 2 = a cluster has more than 32
 * exclusion-partner clusters; either makes the list unusable), _emit writes the tensors with the
 * given capacities (-1 padding past E and M) from the same workspace. */
size_t tff_cluster_pairs_workspace_bytes(long long natoms, const double* box_host, double cutoff, double cell_size);
int tff_cluster_pairs_build(int dtype, const void* coords, long long natoms, const double* box_host, double cutoff,
                            double cell_size, const int64_t* excl_row_ptr, const int64_t* excl_col,
                            long long* counts_out, void* workspace, size_t workspace_bytes, tff_stream_t stream);
int tff_cluster_pairs_emit(int dtype, const void* coords, long long natoms, const double* box_host, double cutoff,
                           double cell_size, const int64_t* excl_row_ptr, const int64_t* excl_col,
                           int32_t* sorted_atom_indices, int32_t* cluster_exclusions, uint32_t* bitmask_exclusions,
                           long long excl_capacity, int32_t* interacting_clusters, int32_t* interacting_atoms,
                           long long row_capacity, void* workspace, size_t workspace_bytes, tff_stream_t stream);

/* Workspace of _decode and _nonbonded for E exclusion blocks (columns of cluster_exclusions,
 * padding included) and M rows. */
size_t tff_cluster_pairs_traverse_workspace_bytes(long long natoms, long long num_exclusion_blocks,
                                                  long long num_rows);
/* decode_cluster_pairs: the pairs (i, j) within the cutoff (float64 minimum-image test) into
 * pairs_out (capacity x 2 int32, -1 after the last), in entry, lane, column order; npairs_out
 * (device, 1 int32) = the number found, which may exceed the capacity (then only capacity are
 * written). */
int tff_cluster_pairs_decode(int dtype, const void* coords, long long natoms, const double* box_host, double cutoff,
                             const int32_t* sorted_atom_indices, const int32_t* cluster_exclusions,
                             const uint32_t* bitmask_exclusions, long long num_exclusion_blocks,
                             const int32_t* interacting_clusters, const int32_t* interacting_atoms, long long num_rows,
                             int32_t* pairs_out, long long pair_capacity, int32_t* npairs_out, void* workspace,
                             size_t workspace_bytes, tff_stream_t stream);
/* compute_nonbonded_energy / forces_from_cluster_pairs: the pair term of
 * compute_nonbonded_energy_from_atom_pairs (k q_i q_j (1/r - [1/rc]) + 4 eps s^6 (s^6 - 1), LB
 * mixing) over the cluster pairs; outputs as tff_pair_energy (energy_out = scale * E, gradients
 * scale * dE/dx, forces_accum += -scale * dE/dr), each may be NULL. */
int tff_cluster_pairs_nonbonded(int dtype, const void* coords, long long natoms, const double* box_host,
                                const void* charges, const void* sigma, const void* epsilon, double coul_constant,
                                double cutoff, int do_shift, const int32_t* sorted_atom_indices,
                                const int32_t* cluster_exclusions, const uint32_t* bitmask_exclusions,
                                long long num_exclusion_blocks, const int32_t* interacting_clusters,
                                const int32_t* interacting_atoms, long long num_rows, double scale, void* energy_out,
                                void* grad_coords, void* grad_charges, void* grad_sigma, void* grad_epsilon,
                                void* forces_accum, void* workspace, size_t workspace_bytes, tff_stream_t stream);

/* ------------------------------------------------- fused energy + forces */
#define TFF_FUSED_PAIR_OVERFLOW 1  /* cluster-pair capacity exceeded: results incomplete */
#define TFF_FUSED_EXCL_OVERFLOW 2  /* an atom has more than 8 exclusion partners */
#define TFF_FUSED_COLUMN_OVERFLOW 4 /* an x-y column holds more than 8192 atoms: left unsorted in
                                       z, which the charge spreading's cluster search relies on
                                       (reciprocal-space results incomplete) */
#define TFF_FUSED_BRICK_OVERFLOW 8  /* a charge-spreading brick is reached by more than 1024
                                       clusters (float64 spread), or holds the stencil bases of
                                       more than 4 x the mean atoms per brick + 64 (float spread):
                                       strongly inhomogeneous system or a coarse grid;
                                       reciprocal-space results incomplete */
#define TFF_FUSED_BOX_STALE 16      /* box_device no longer holds the values of box_host that the
                                       call was planned from: results invalid */

size_t tff_fused_workspace_bytes(int compute_dtype, long long natoms, const double* box_host, double list_cutoff,
                                 int K0, int K1, int K2);

/* Energy and forces of the TIP3P-style composition in one call: the sum of
 * compute_multipolar_energy_from_atom_pairs (rank 0, with the erfc - 1 correction on every
 * excluded pair), compute_lennard_jones_energy with Lorentz-Berthelot mixing of per-atom
 * sigma / epsilon, and prefactor * pme_long_range (any spline order 3..8), as composed by
 * examples/tip3p/model.py:100-123, without materialising a pair list.  Orthorhombic box,
 * each edge at least 2 * list_cutoff.  compute_dtype 0 = float32 arithmetic with int64
 * fixed-point force / energy accumulation and float64 k-space sums, 1 = float64.
 * energies_out (device, 5 doubles): Lennard-Jones, real-space Coulomb (incl. exclusion
 * correction), reciprocal, self, total.  forces_out (io_dtype, N x 3) = -dE/dr.  status_out
 * (device, 2 int32): [0] TFF_FUSED_* flags, 0 when complete; [1] number of 32 x 32
 * cluster-pair tiles within the list cutoff (their 8 x 8 sub-tiles are pruned at the
 * cutoff on every call).  box_device (optional, io_dtype, 9 values): the caller's device copy
 * of the box; the call flags TFF_FUSED_BOX_STALE when it differs from box_host.  status_host
 * (optional, 2 int32 in pinned host memory mapped for the device): the last kernel of the call
 * copies status_out there, word 1 first and word 0 after a system-scope fence, so the caller can
 * read the flags without a device-to-host copy (they are final when the last kernel starts: a
 * caller that stores -1 in word 0 before the call reads valid flags once word 0 is not -1).  Moduli as for tff_pme_energy with zmoduli of length
 * K2/2 + 1, in compute_dtype. */
int tff_fused_energy_forces(int io_dtype, int compute_dtype, const void* coords, const void* charges,
                            const void* sigma, const void* epsilon, long long natoms, const double* box_host,
                            const int64_t* excl_row_ptr, const int64_t* excl_col, double cutoff, double list_cutoff,
                            double ewald_alpha, double coul_constant, int K0, int K1, int K2, int order,
                            const void* xmoduli, const void* ymoduli, const void* zmoduli, double* energies_out,
                            void* forces_out, int32_t* status_out, const void* box_device, int32_t* status_host,
                            void* workspace, size_t workspace_bytes, tff_stream_t stream);

/* Allocation-free variant of tff_fused_energy_forces (no counterpart in the reference): for
 * callers that cannot allocate device memory (code under HIP graph capture with allocation
 * forbidden, tight serving or MD loops).  Same arguments and results, without the workspace: the
 * call uses a static device array of TFF_FUSED_SCRATCH_BYTES per device, held from the
 * library's first use on that device until the process ends, and calls no allocating routine
 * (the first call on a device looks up the array's address and creates the auxiliary stream).
 * A system whose tff_fused_workspace_bytes exceeds that size is refused with TFF_ERR_WORKSPACE
 * (64 MiB: the 96,000-atom water box takes 35 MiB in float32 and 44 MiB in float64 arithmetic).
 * All calls on one device share the array: calls that overlap on one device race. */
#define TFF_FUSED_SCRATCH_BYTES (64ull << 20)
int tff_fused_energy_forces_noalloc(int io_dtype, int compute_dtype, const void* coords, const void* charges,
                                    const void* sigma, const void* epsilon, long long natoms,
                                    const double* box_host, const int64_t* excl_row_ptr, const int64_t* excl_col,
                                    double cutoff, double list_cutoff, double ewald_alpha, double coul_constant,
                                    int K0, int K1, int K2, int order, const void* xmoduli, const void* ymoduli,
                                    const void* zmoduli, double* energies_out, void* forces_out,
                                    int32_t* status_out, const void* box_device, int32_t* status_host,
                                    tff_stream_t stream);

/* Vector-Jacobian product of tff_fused_energy_forces, for force matching: given upstream
 * gradients grad_energy (device, 1 double: dL/dE; NULL = 0) and grad_forces (device, io_dtype
 * N x 3: dL/dF; NULL = 0), returns dL/dcoords, dL/dcharges, dL/dsigma, dL/depsilon (device,
 * io_dtype; each may be NULL), i.e. gE dE/dtheta - d(sum_i v_i . dE/dr_i)/dtheta with
 * v = grad_forces: the second derivatives of the energy that the reference obtains only from
 * its pure-PyTorch path by double backward (torchff/multipoles.py:244-341, torchff/vdw.py:86-111,
 * torchff/pme_ref.py:11-469 with create_graph; its CUDA ops' backward only scales stored first
 * derivatives, e.g. csrc/nonbonded/nonbonded_atom_pairs_cuda.cu:143-260).  The cluster pair list
 * is rebuilt here (tff_fused_energy_forces_backward_reuse takes the forward call's).
 * dL/depsilon_i is 0 for pairs with eps_i eps_j = 0 (as the forward's first derivatives).  Same
 * arguments and constraints as tff_fused_energy_forces otherwise; workspace sized by
 * tff_fused_backward_workspace_bytes. */
size_t tff_fused_backward_workspace_bytes(int compute_dtype, long long natoms, const double* box_host,
                                          double list_cutoff, int K0, int K1, int K2);
int tff_fused_energy_forces_backward(int io_dtype, int compute_dtype, const void* coords, const void* charges,
                                     const void* sigma, const void* epsilon, long long natoms,
                                     const double* box_host, const int64_t* excl_row_ptr,
                                     const int64_t* excl_col, double cutoff, double list_cutoff,
                                     double ewald_alpha, double coul_constant, int K0, int K1, int K2, int order,
                                     const void* xmoduli, const void* ymoduli, const void* zmoduli,
                                     const double* grad_energy, const void* grad_forces, void* grad_coords,
                                     void* grad_charges, void* grad_sigma, void* grad_epsilon,
                                     int32_t* status_out, const void* box_device, int32_t* status_host,
                                     void* workspace, size_t workspace_bytes, tff_stream_t stream);
/* As tff_fused_energy_forces_backward, but the cluster arrays, the exclusion table, the pair list
 * and the charges' potential grid are taken from the workspace as the last
 * tff_fused_energy_forces call on it left them, instead of being made again (about 130 us of a
 * force-matching step at 96 000 atoms); the results are bitwise the same.  The caller
 * guarantees that this call was made with the same (backward-sized) workspace, the same inputs
 * (coordinates, charges, sigma, epsilon, box, exclusions, cutoffs, grid) and on the same stream, and
 * that no other call has written the workspace since; the forward call's status words (capacity
 * flags, tile count) start this call's.  No counterpart in the reference, whose autograd recomputes
 * the forward's graph; it serves autograd callers that keep the forward's workspace. */
int tff_fused_energy_forces_backward_reuse(int io_dtype, int compute_dtype, const void* coords,
                                           const void* charges, const void* sigma, const void* epsilon,
                                           long long natoms, const double* box_host,
                                           const int64_t* excl_row_ptr, const int64_t* excl_col, double cutoff,
                                           double list_cutoff, double ewald_alpha, double coul_constant, int K0,
                                           int K1, int K2, int order, const void* xmoduli, const void* ymoduli,
                                           const void* zmoduli, const double* grad_energy,
                                           const void* grad_forces, void* grad_coords, void* grad_charges,
                                           void* grad_sigma, void* grad_epsilon, int32_t* status_out,
                                           const void* box_device, int32_t* status_host, void* workspace,
                                           size_t workspace_bytes, tff_stream_t stream);

/* ----------------------------------------------------- batched serving */
/* Energy and forces of many small orthorhombic water-like systems in one fixed sequence of
 * launches: each system is the composition of tff_fused_energy_forces (per-atom Lorentz-Berthelot
 * LJ, rank-0 Ewald real space, smooth PME, Coulomb constant coul_constant) with no
 * restriction on the box size: the real-space sum runs over every periodic image within the
 * cutoff, and the pairs of one molecule are excluded on their nearest image (the correction
 * k q_i q_j (erfc(alpha r) - 1) / r, examples/tip3p/model.py:110-112 and
 * torchff/multipoles.py:256-257; no LJ there).  The reference evaluates one system per operator
 * call (examples/tip3p/model.py:100-123); these entry points have no single counterpart there:
 * they serve that per-system evaluation to many systems at once.
 *
 * Capacities (the same four on every call for one workspace): max_systems, max_atoms (atom
 * records), max_k (largest grid edge; float32: <= 62, float64: <= 32, one x-slab transform in
 * 64 KiB of LDS), max_grid_points (summed K0 K1 K2).  Launch grids follow them, never
 * the batch, so the tff_serve_energy_forces call can be captured once into a HIP graph and
 * replayed after each tff_serve_upload.
 *
 * A batch is a host table, tff_serve_table_bytes long (pinned memory for asynchronous uploads):
 * a head, one record per system (tff_serve_table_add), and from byte tff_serve_atoms_offset on
 * the atom records, 8 float64 each: x, y, z (nm), charge (e), sigma (nm), epsilon (kJ/mol),
 * molecule id (equal within a molecule: the exclusion group), 0.  tff_serve_upload copies the
 * head, the records and the atom records up to the highest one used into the workspace: a
 * 16-byte aligned table is read by one kernel, which takes the size from the head when it runs
 * (so the upload can be captured into one graph with the evaluation and serve batches of any
 * size), another by one copy; tff_serve_energy_forces evaluates it into out (float64, device memory or pinned host
 * memory mapped for the device, written with plain stores only): 5 values per system in
 * table order [real space (LJ + Coulomb incl. the exclusion correction), 0, reciprocal, self,
 * total] at out[5 s], then from out[5 max_systems] the forces -dE/dr (3 per atom record, rows of
 * records in no system not written).  A new workspace is zeroed once by
 * tff_serve_workspace_init; the charge grid is cleaned by each evaluation for the next.  The
 * workspace begins with the table's layout: after tff_serve_upload the atom records sit at byte
 * tff_serve_atoms_offset(max_systems) of the workspace, where the caller may rewrite their
 * coordinates on the device between evaluations (e.g. each MD step) without another upload. */
#define TFF_FULL 5 /* tff_serve_table_add: the system does not fit the capacities (table unchanged) */

size_t tff_serve_table_bytes(int max_systems, int max_atoms);
size_t tff_serve_atoms_offset(int max_systems);
void tff_serve_table_reset(void* table_host);
/* totals_host (8 int64): systems, atoms, 64-atom tasks, FFT tiles of the z, y and x passes, grid
 * points, one past the highest atom record used */
int tff_serve_table_totals(const void* table_host, long long* totals_host);
/* Appends a system of atom records atom_off .. atom_off + natoms - 1; box_lengths_host (3
 * doubles, nm), K_host (3 grid sizes 2^a 3^b 5^c, each <= max_k), moduli_off_host (3 offsets
 * into the moduli table where |theta(m)|, m = 0..K-1, of each axis start, as tff_pme_energy's
 * xmoduli, in compute_dtype).  Returns 0, TFF_FULL, or an error. */
int tff_serve_table_add(void* table_host, int compute_dtype, int max_systems, int max_atoms, int max_k,
                        long long max_grid_points, int atom_off, int natoms, const double* box_lengths_host,
                        const int* K_host, const int* moduli_off_host, double cutoff, double ewald_alpha);
/* 0 when the capacities are rejected */
size_t tff_serve_workspace_bytes(int compute_dtype, int max_systems, int max_atoms, int max_k,
                                 long long max_grid_points);
int tff_serve_workspace_init(int compute_dtype, int max_systems, int max_atoms, int max_k, long long max_grid_points,
                             void* workspace, size_t workspace_bytes, tff_stream_t stream);
int tff_serve_upload(const void* table_host, int compute_dtype, int max_systems, int max_atoms, int max_k,
                     long long max_grid_points, void* workspace, size_t workspace_bytes, tff_stream_t stream);
/* Spline order 3..8.  Calls that overlap on one workspace race. */
int tff_serve_energy_forces(int compute_dtype, double coul_constant, int order, const void* moduli, int max_systems,
                            int max_atoms, int max_k, long long max_grid_points, double* out, void* workspace,
                            size_t workspace_bytes, tff_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TORCHFF_HIP_H */
