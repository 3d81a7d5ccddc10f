/*
 * qs3x.h -- C ABI of libqs3x, an MI355X (gfx950) implementation of QS3's
 * symmetry-adapted matrix-free Lanczos engine (U(1) x translation basis).
 *
 * Every function returns an int status: 0 on success, non-zero on error, with
 * a message available from qs3x_last_error() (thread-local).  Plain C types
 * only: int32_t / int64_t / double, complex numbers as interleaved
 * (re, im) double pairs, opaque handles.  Host pointers unless a name says
 * "_dev"; device pointers are plain HIP device addresses.
 *
 * Each entry point names the reference routine it replaces (file:line, paths
 * relative to the QS-Cube/ED source tree).  The ISO_C_BINDING interfaces a
 * Fortran driver binds are in ed_amd/fortran/qs3x_iso_c.f90.
 *
 * Integer kinds: ranks, dimensions and row indices are int64_t (C(1000,4)
 * exceeds 2^31); sites, counts and lattice sizes are int32_t.  Ranks are the
 * reference's 1-based colex ranks (state_lists.f90:131-139); list indices are
 * 0-based in this ABI and 1-based in the Fortran shims.
 */
#ifndef QS3X_H
#define QS3X_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qs3x_model qs3x_model;
typedef struct qs3x_sector qs3x_sector;

/* Precision of the vectors an operation works on. */
enum {
  QS3X_FP64 = 0, /* complex double vectors, FP64 arithmetic                     */
  QS3X_FP32 = 1  /* complex float vectors; H applied as H - e_zz (e_zz = sum Jz/4,
                    added back in FP64), dots and norms accumulated in FP64      */
};

/* ---------------------------------------------------------------- errors */
const char *qs3x_last_error(void);
/* Device the library runs on (HIP ordinal); default 0. */
int qs3x_set_device(int32_t device);

/* Tunables (no counterpart in the reference):
 *   "tr_reorth"     0 (default): the thick restart reorthogonalises every new
 *                   Lanczos vector against all previous basis vectors, as
 *                   eigen_solver.f90:233-239 does; 1: against the locked Ritz
 *                   vectors and the last two Lanczos vectors only; 2 (with
 *                   tr_check_every > 0): the three-term recurrence in the
 *                   first cycle, then as 1.
 *   "tr_cgs_passes" classical Gram-Schmidt passes per step (default 2; steps
 *                   of restart cycles take at least 2).
 *   "fp32_floor"    Ritz-value change (absolute) that ends the FP32 phase of
 *                   the mixed thick restart (default 1e-6).
 *   "hv_kernel"     0 (default): the fastest applicable H.v kernel (nearest-
 *                   neighbour chain, batched-load chain, per-hop chain, general
 *                   lattice, in that order); 1, 2, 3: start the list at the
 *                   batched-load chain, per-hop chain or general kernel.
 *   "tr_check_every" thick restart (the FP64 solve and the FP32 stage of the
 *                   mixed one; not its FP64 fallback): 0 (default) checks
 *                   convergence at restarts as the reference does; n > 0
 *                   checks the projection every n steps inside a cycle and
 *                   ends the cycle there, once the change is below conv and
 *                   so is the remaining change extrapolated from the last two
 *                   (with tr_reorth >= 1 its first cycle is the plain
 *                   three-term recurrence).  NOM is at most 1023, and the
 *                   rotation stages NOM x NOK FP64 coefficients (NOM x 16
 *                   ceil(NOK / 16) with element loads) in at most 156 KB.
 *   "fp64_refine"   FP32-mixed thick restart: after the FP64 Rayleigh-Ritz
 *                   values are accepted, Rayleigh-Ritz on the block Krylov
 *                   space {Y, HY, ..., H^s Y} of the NOE accepted vectors, in
 *                   FP64 (s = value, default 4; 0 skips this refinement).
 *                   It costs s NOE + (s + 1) NOE FP64 H.v and needs
 *                   NOE (2 s + 3) <= max(min(NOM, 64), 2 NOK + NOE) + 1
 *                   (s is lowered to fit).
 *   "rot_vec"       Ritz rotation with 16-byte basis loads where the slots are
 *                   16-byte aligned (1, default) or element loads (0).
 *   "hv_xcd"        L2 domains (XCDs) for the H.v tile schedule (8: MI355X's
 *                   XCDs; 1: plain round robin over workgroups).  0, the
 *                   default: 8.
 *   "hv_chunk"      Nearest-neighbour chain H.v with hv_xcd > 1: 0 gives each
 *                   domain one contiguous range of rows; c > 0 deals chunks
 *                   of up to 2^c tiles to the domains in turn (smaller for
 *                   small sectors, so that every domain owns at least 8).
 *                   -1, the default: 5 (the fastest at L = 1000, N_down = 4
 *                   in both precisions).
 *   "hv_stencil"    1 (default): nearest-neighbour chains with uniform bonds
 *                   and N_down = 4 use the gap-lattice stencil H.v (tiles of
 *                   (n2, n3) swept over n4 through LDS, every hop of a clear
 *                   row read from there, the other rows in a second launch
 *                   through the key search); 0: the nearest-neighbour chain
 *                   kernel.  "hv_stencil_kernel" picks how a block's threads
 *                   take a tile's rows: 0 slab window (the rows of each slab
 *                   dealt out in turn), 1 column (a thread keeps its window
 *                   positions for the whole sweep), 2 (default) pair (column
 *                   form with two adjacent positions per lane and 16-byte
 *                   accesses, FP32 sectors under 2^28 rows; else column).
 *                   All three give the same bits.  "hv_stencil_lines"
 *                   (even, 2..14; 0, the default: 10; the slab-window form
 *                   takes at most 8, the pair form 10) and
 *                   "hv_stencil_slabs" (0, the default: 16 in FP32, 8 in
 *                   FP64) set its tile height and the n4-slabs per task;
 *                   "hv_stencil_span" (column form, FP32: 1 or 2 window
 *                   positions 64 apart per lane); "hv_stencil_depth" (2, 3
 *                   slab-window form, 4 pair form: slab windows in flight);
 *                   "hv_stencil_order" 1 (default) takes the tasks n4-major,
 *                   0 tile-major; "hv_stencil_side" 1 runs its key-search
 *                   rows on a side stream overlapping the stencil (default 0:
 *                   after it; the pair form always after it).
 *   "dense_solver"  eigensolver of qs3x_full_diag: 0 (default) QR iteration
 *                   (rocSOLVER zheev, the algorithm of LAPACK zheev that the
 *                   reference calls); 1 divide and conquer (zheevd).
 *   "verbose"       0 (default); 1: a stderr line per thick restart and per
 *                   100 Lanczos steps; 2: per convergence check.
 * The environment variable QS3X_OPTIONS="name=value,name=value" applies the
 * same settings once, before the first model is created (for drivers that
 * cannot call this function); a bad entry makes qs3x_model_create fail. */
int qs3x_set_option(const char *name, double value);

/* ----------------------------------------------------------------- model */
/* Lattice and XXZ bond list.  Replaces read_ip's bond handling
 * (input_param.f90:58-103: p1 < p2 swap, sJint = (Jxy/2, Jz/4)) and the
 * translation tables mk_shift_{x,y,z}_SQ (state_lists.f90:36-64, main.f90:49-51).
 * p_xxz: 2*NOxxz site indices (1-based), pairs in bond-file order;
 * jxy, jz: NOxxz raw couplings (Jint(:,2), Jint(:,3)).  NOS = LX*LY*LZ. */
int qs3x_model_create(int32_t NOS, int32_t LX, int32_t LY, int32_t LZ, int32_t NOxxz,
                      const int32_t *p_xxz, const double *jxy, const double *jz,
                      qs3x_model **out);
int qs3x_model_destroy(qs3x_model *m);
/* sum_b Jz_b/4: the energy of the N_down = 0 state.  Replaces calcu_FM_energy
 * (eigen_solver.f90:341-347). */
int qs3x_fm_energy(const qs3x_model *m, double *e);

/* ---------------------------------------------------------------- sectors */
/* Wave vector exactly as read_ip computes it: rk = 2*acos(-1)*K/L
 * (input_param.f90:76-78). */
double qs3x_wavevector(int32_t K, int32_t L);

/* Builds the (N_down, rk) sector on the device: translation representatives
 * (colex-minimal orbit members), their norms and a rank-addressed index.
 * Replaces allocate_lists_omp_SQ (state_lists.f90:225-292) and, with
 * rk = rk + q, allocate_lists_omp_SQ_new (state_lists.f90:418-486).
 * THS (the sector dimension) is returned through dim.  Every rk_a L_a / (2 pi)
 * must be an integer (to 1e-8): an incommensurate wave vector, which the
 * reference accepts for rk + q without a check, is an error here. */
int qs3x_sector_build(qs3x_model *m, int32_t nod, double rkx, double rky, double rkz,
                      qs3x_sector **out, int64_t *dim);
/* U(1)-only block: every configuration with nod down spins, rows in colex
 * rank order (row i = rank i + 1), no translation symmetry.  Replaces the
 * list setup of the U(1)-only code (source_only_u1/main.f90:23-29, THS =
 * combination(NOS, NOD)); H . v follows source_only_u1/ham2vec_v3.f90:35-65.
 * Such a sector supports H . v, the Lanczos solvers and the DSF recurrence;
 * qs3x_sector_lists returns list_s(i) = i and unit norms.  Any NOS below
 * 65536 (up to 64 sites a row's occupation is one 64-bit word; beyond, the
 * kernel works from the sorted down-spin sites). */
int qs3x_sector_build_u1(qs3x_model *m, int32_t nod, qs3x_sector **out, int64_t *dim);
/* Same sector from caller-supplied lists (e.g. the reference's own list_s /
 * list_r / explist), for drop-in use behind ham_to_vec_wave_vector: list_s
 * ascending 1-based ranks, list_r norms, explist (LX+1)(LY+1)(LZ+1) complex
 * column-major (input_param.f90:88-89). */
int qs3x_sector_from_lists(qs3x_model *m, int32_t nod, int64_t dim, const int64_t *list_s,
                           const double *list_r, const double *explist, qs3x_sector **out);
int qs3x_sector_destroy(qs3x_sector *s);
int qs3x_sector_dim(const qs3x_sector *s, int64_t *dim);
/* Copies list_s (ranks) and list_r (norms) to host arrays of length dim
 * (input_param.f90:33-34 globals list_s, list_r). */
int qs3x_sector_lists(const qs3x_sector *s, int64_t *list_s, double *list_r);
/* Number of off-diagonal hop terms with a target inside the sector (nnz of
 * the matrix-free operator), counted on the device. */
int qs3x_sector_nnz(qs3x_sector *s, int64_t *nnz);

/* ------------------------------------------------------------------ H . v */
/* v0 = H v1 with host vectors (complex double, interleaved, length dim).
 * Replaces ham_to_vec_wave_vector (ham2vec_v3.f90:37-72).  Stages both
 * vectors over PCIe; solvers below keep everything on the device. */
int qs3x_ham_to_vec_wave_vector(qs3x_sector *s, double *v0, const double *v1);
/* The same H . v as qs3x_ham_to_vec_wave_vector (host vectors, FP64), with no
 * device allocation: no counterpart in the reference; for callers that cannot
 * allocate device memory (loops that forbid hipMalloc, harnesses that
 * pre-size all device memory).  Both vectors are staged through one static
 * per-device array of 64 MiB, so dim <= 2,097,152 rows (larger sectors return
 * an error).  Calls that overlap on one GPU share that array and race. */
int qs3x_ham_to_vec_wave_vector_noalloc(qs3x_sector *s, double *v0, const double *v1);
/* Device-pointer H . v on the given stream (hipStream_t passed as void*;
 * NULL is the HIP null stream).  prec: QS3X_FP64 -> v0, v1 complex double;
 * QS3X_FP32 -> complex float with the e_zz shift (v0 = (H - e_zz) v1). */
int qs3x_hv_dev(qs3x_sector *s, int32_t prec, void *v0_dev, const void *v1_dev, void *stream);

/* --------------------------------------------------------------- solvers */
/* Plain Lanczos, device resident.  Replaces lanczos_routines /
 * lancz_eigen_val (lanczos.f90:14-152): three-term recurrence without
 * reorthogonalisation, lowest Ritz value every itrint steps past minitr,
 * stop when |E_prev/E - 1| < conv; optional second pass for the eigenvector.
 * FP64 stops at dim steps as the reference does; the FP32-mixed recurrence
 * may run to 2 dim + 8 steps (lost orthogonality needs them).
 * start: host start vector (complex double, length dim) or NULL for the
 * counter-based generator keyed by seed.  ene[0] = E0; psi (host, length
 * dim, may be NULL) receives the normalised ground state; iters the number of
 * Lanczos steps of the first pass. */
int qs3x_lanczos(qs3x_sector *s, int32_t prec, int64_t maxitr, int64_t minitr, int64_t itrint,
                 double conv, const double *start, uint64_t seed, double *ene, double *psi,
                 int64_t *iters);

/* Thick-restart Lanczos with full reorthogonalisation, device resident.
 * Replaces my_trlanczos_routines_complex / lancz_eigen_val_complex
 * (eigen_solver.f90:10-261).  NOM is raised to 2*NOK as in the reference.
 * prec QS3X_FP32 runs the restarts on complex-float Krylov vectors (FP64
 * dots, reorthogonalisation coefficients and Ritz rotation) and then refines
 * in FP64 until the FP64 criterion holds.  ene (length NOE) receives the
 * lowest Ritz values; psi (host, dim x NOE column-major, may be NULL) the
 * Ritz vectors; stats (may be NULL, 8 doubles): [restarts, hv_fp32, hv_fp64,
 * seconds_total, seconds_fp32, max_residual, rr_accepted (1 certified by the
 * residual bound, 2 accepted by agreement, 0 continued in FP64), seconds_alloc];
 * seconds_total, seconds_fp32 and the solver trace exclude seconds_alloc, the
 * time spent allocating the Krylov bases, and the time freeing them
 * (qs3x_solver_certificate).  In FP32 mode the FP32 Ritz vectors
 * go through one FP64 Rayleigh-Ritz step (NOK FP64 H.v) and, with option
 * fp64_refine > 0, the FP64 block-Krylov refinement; the result is accepted
 * (rr_accepted = 1) when, for every wanted pair, r^2/gap <= conv |theta| before
 * or after the refinement, or the FP32 restarts met the convergence test and
 * the FP64 Rayleigh quotient equals the FP32 Ritz value to conv |theta|;
 * otherwise an FP64 thick restart (full reorthogonalisation) continues from
 * the refined vectors. */
int qs3x_trlanczos(qs3x_sector *s, int32_t prec, int32_t NOE, int32_t NOK, int32_t NOM,
                   int64_t maxitr, double conv, const double *start, uint64_t seed,
                   double *ene, double *psi, double *stats);

/* ------------------------------------------------ dense path (ALG = 3) */
/* The dense H of the sector into host ham (complex double, column-major,
 * leading dimension ld >= dim): ham(i, j) = the term of row i on configuration
 * j.  Replaces make_full_hamiltonian (ham2vec_v3.f90:5-35).  Column j is one
 * FP64 H . v launch on e_j, so the matrix is the matrix-free operator itself;
 * needs 16 dim^2 bytes of device memory. */
int qs3x_make_full_hamiltonian(qs3x_sector *s, double *ham, int64_t ld);
/* Full diagonalisation on the device: ene (host, dim) the eigenvalues in
 * ascending order; psi (host, column-major, leading dimension ld >= dim, may be
 * NULL for eigenvalues only) the orthonormal eigenvectors.  Replaces
 * Full_diag_routines / fulldia_routines_with_momentum
 * (eigen_solver.f90:51-67,263-339): the dense H above, diagonalised from its
 * upper triangle (as zheev('V', 'U') reads it) by rocSOLVER zheev (option
 * "dense_solver" 1: zheevd).  A non-finite eigenvalue or eigenvector
 * component is an error.  dim < 2^31. */
int qs3x_full_diag(qs3x_sector *s, double *ene, double *psi, int64_t ld);

/* ------------------------------------------------------ batched wave vectors */
/* A chain sector holding every translation representative of N_down (the
 * union of all k sectors, rows in ascending rank).  No counterpart in the
 * reference; its norms are per wave vector, so it serves the batched entry
 * point below (its list_r is 1 and it is not a sector for qs3x_hv_dev). */
int qs3x_sector_build_all_k(qs3x_model *m, int32_t nod, qs3x_sector **out);

/* The DSF recurrence for nq <= 64 wave vectors in one block.  Replaces nq
 * calls of calcu_DSF_wavevector for S^- (get_expectation_values.f90:379-605,
 * with calcu_Sq_a at :281-377 and allocate_lists_omp_SQ_new at
 * state_lists.f90:418-486): for each j, f_j = S^-_q |psi> with
 * q = 2 pi K[j] / L - k_gs is built into the k = 2 pi K[j] / L sector of
 * N_down + 1, and itr_dsf steps of the recurrence run on all f_j together
 * (interleaved vectors; hop structure shared by all of them).  gs: the
 * ground-state chain sector (nearest-neighbour chains); psi: its vector (host,
 * complex double); pos: 3 x NOS site positions.  Outputs as nq calls of
 * qs3x_dsf_recurrence: alpha, beta2 (nq x itr_dsf, row-major), norm1 (nq).
 * stats (8 doubles, may be NULL): [seconds setup, seconds S^-, seconds
 * recurrence, ms per block step, rows of the all-k sector, steps, 1 if the
 * clear rows ran through the block gap-lattice stencil (N_down + 1 = 4,
 * uniform bonds; option block_stencil), 2 if it ran the fused step (option
 * block_fused: H.v and the three-term update in one pass, alpha and beta
 * from four per-q sums; four vector buffers instead of three) else 0, rows
 * through the per-row block kernel in those cases]. */
int qs3x_dsf_sminus_batch(qs3x_sector *gs, int32_t nq, const int32_t *K, const double *pos,
                          const double *psi, int32_t itr_dsf, double rfield, int32_t prec,
                          double *alpha, double *beta2, double *norm1, double *stats);
/* The tile shapes the last qs3x_dsf_sminus_batch timed with option
 * block_autotune 1 (no counterpart in the reference): rows of 4 doubles
 * (lines per tile, threads per block, slabs per task, ms per block stencil
 * launch), the chosen shape repeated as the last row; none without the
 * option or where the block stencil does not apply. */
int qs3x_block_tune_last(int32_t max_rows, int32_t *n_rows, double *rows);

/* Plain Lanczos for nsec sectors of one chain model at once.  Replaces nsec
 * calls of lanczos_routines (lanczos.f90:14-152; energies only, no second
 * pass): every sector runs the reference's recurrence from its own
 * counter-based start vector (seed), is checked every itrint steps past
 * minitr, and stops at |E_prev/E - 1| < conv, at breakdown, at its dimension
 * or at maxitr.  Each step is one H.v launch per N_down (all k sectors of it)
 * plus one update and two segmented reductions for all sectors;
 * use_graph != 0 replays the step as a captured hipGraph.  ene (nsec):
 * lowest Ritz values; iters (nsec, may be NULL): steps per sector; stats (8,
 * may be NULL): [seconds total, seconds setup, group steps, step launches,
 * padded rows, N_down groups, bytes of the hop cache, 0].  Option
 * group_hop_cache (1, the default; no counterpart in the reference): step 0
 * records every row's hops (target index and a code of bond type,
 * stabiliser order and translation; 8 bytes per hop slot) and the later
 * steps read them instead of searching for representatives again, when the
 * tables of all N_down groups fit in 32 GiB; the arithmetic, and so every
 * result, is the same to the bit. */
int qs3x_lanczos_group(qs3x_sector *const *secs, int32_t nsec, int32_t prec, int64_t maxitr,
                       int64_t minitr, int64_t itrint, double conv, uint64_t seed, int32_t use_graph,
                       double *ene, int64_t *iters, double *stats);

/* Certificate of the last FP32-mixed qs3x_trlanczos (no counterpart in the
 * reference): out (6 doubles) = [spectral gap of the lowest pair, estimated
 * from the FP32 projection's next distinct Ritz value; that pair's FP64
 * residual |H y - theta y|; the bound r^2 / gap on |theta - E0|; acceptance:
 * 1 certified by the bound (<= conv |theta|), 2 accepted because the FP32
 * and FP64 values agree (not certified), 0 not accepted (an FP64 thick
 * restart continued), -1 no mixed solve yet; then, for the last thick
 * restart of either precision, seconds allocating Krylov bases (stats[7])
 * and seconds freeing them (not in stats[3] or the trace: the free of the
 * FP32 basis of a mixed solve comes after its energies are known)]. */
int qs3x_solver_certificate(double *out);

/* Convergence trace of the last qs3x_lanczos / qs3x_trlanczos call (no
 * counterpart in the reference, whose solvers print progress instead):
 * *n rows of [seconds since the solve started, H.v applications so far,
 * lowest Ritz value, 32 or 64 = precision of the Krylov vectors]; the first
 * min(cap, *n) rows are written to out (4 doubles each; out may be NULL). */
int qs3x_solver_trace(double *out, int64_t cap, int64_t *n);

/* ------------------------------------------------------------ observables */
/* <S^z_i S^z_j> and <S^+_i S^-_j> for npairs ordered pairs (1-based sites,
 * pairs[2q], pairs[2q+1]) in the state psi (host, complex double, length dim).
 * Replaces get_cf_2_wave_vector / calcu_cf_trans_2 / mk_list_ij
 * (get_expectation_values.f90:97-131,234-279; ham2vec_v3.f90:111-174):
 * values are computed once per translation orbit and replicated.  The pair
 * list is read as FILECF is (each line "a b" stored as list_expe_ss(1)=b,
 * list_expe_ss(2)=a, get_expectation_values.f90:25-27) and must hold all NOS^2
 * pairs, as the reference assumes. */
int qs3x_correlations(qs3x_sector *s, const double *psi, int64_t npairs, const int32_t *pairs,
                      double *szsz, double *spsm);
/* <S^z_j> for all sites.  Replaces get_lm_2_wave_vector / calcu_lm_trans_2
 * (get_expectation_values.f90:65-95,197-232; ham2vec_v3.f90:74-109). */
int qs3x_local_mag(qs3x_sector *s, const double *psi, double *expe_sz);

/* f = S^a_q |psi> into the target sector tgt.  Replaces calcu_Sq_a and
 * search_representative_S{p,m,z}i_state_SQ (get_expectation_values.f90:281-377;
 * state_lists.f90:488-583), as a pull-form gather (no atomics).
 * spsmsz 1: S+, 2: S-, 3: Sz; q[3] the transferred wave vector; pos 3*NOS
 * site positions (list_site_position_*). psi, f: host complex double. */
int qs3x_apply_sq(qs3x_sector *gs, qs3x_sector *tgt, int32_t spsmsz, const double *q,
                  const double *pos, const double *psi, double *f);

/* DSF continued-fraction recurrence, device resident.  Replaces
 * calcu_DSF_wavevector (get_expectation_values.f90:379-605) after the S_q step:
 * f (host, target-sector vector, not normalised); H - rfield*sztot_new on the
 * target sector; itr_dsf steps.  alpha[itr_dsf], beta2[itr_dsf] (beta2[0] = 0,
 * stored like beta_dsf), norm1 = <f|f>.  prec FP32 keeps the three Lanczos
 * vectors in complex float with FP64 coefficients.  nvec > 1 runs nvec
 * independent recurrences in one batch: f is dim x nvec, alpha/beta2
 * itr_dsf x nvec, norm1 nvec. */
int qs3x_dsf_recurrence(qs3x_sector *tgt, int32_t prec, int32_t nvec, const double *f,
                        int32_t itr_dsf, double rfield, double *alpha, double *beta2,
                        double *norm1);

/* S(q, w) = -Im G(w)/pi from alpha / beta2 (beta2[0] = 0).  Replaces
 * DSF_analysis (source_DSFLan/main.f90:40-94).  z = w + E0 + i eta.
 * reference_quirk = 1 reproduces DSFLan's deepest level (imaginary part not
 * divided by the denominator, main.f90:64-65); 0 is the exact fraction. */
int qs3x_continued_fraction(int32_t nlev, const double *alpha, const double *beta2, double norm1,
                            double E0, double eta, int64_t nomega, const double *omega,
                            int32_t reference_quirk, double *dsf);

/* ---------------------------------------------------------- measurement */
/* The kernel family of the sector's last H . v launch (no counterpart in the
 * reference): 0 nearest-neighbour chain, 1 batched-load chain, 2 per-hop
 * chain, 3 general lattice, 4 U(1)-only, 5 gap-lattice stencil, 6 the fused
 * Lanczos step on the stencil (qs3x_lanczos on such a sector, option
 * lanczos_fused); -1 before the first launch.
 * lds_bytes: its dynamic LDS. */
int qs3x_hv_last_launch(const qs3x_sector *s, int32_t *family, int64_t *lds_bytes);
/* The stencil H . v's plan for this sector and precision, after its first
 * launch (no counterpart in the reference): info (8 int64): rows left to the
 * key search, non-empty (n3, n4) lines, tasks, tile height in lines, n4-slabs
 * per task, workgroups, kernel form (0 slab window, 1 column, 2 pair), window
 * positions per lane of the column form. */
int qs3x_hv_stencil_info(const qs3x_sector *s, int32_t prec, int64_t *info);
/* Times n_rep device H . v applications (events on the library stream) and
 * returns the mean milliseconds per application. */
int qs3x_time_hv(qs3x_sector *s, int32_t prec, int32_t n_rep, double *ms);
/* Launch autotuner for the sector's H . v (no counterpart in the reference).
 * Times n_rep applications on a random vector of, first, every kernel family
 * that applies in its default configuration (stencil, nearest-neighbour
 * chain, batched-load chain, per-hop chain, general lattice), then the
 * winning family's own parameters: for the stencil its kernel form (the
 * precision's default -- pair in FP32, column in FP64 -- and the slab-window
 * form), tile height (4 .. 14 lines, as each form allows), task length (8,
 * 16, 32 slabs) and slab windows in flight (2, 3 slab window; 2, 4 pair); for
 * the nearest-neighbour kernel rows per lane (FP32: two where they
 * apply, or one), interleaved XCD chunk (2^0, 2^3, 2^5, 2^7 tiles), tile
 * schedule (per-XCD or round robin) and grid (4 or 2 quarters of the
 * resident grid); for the others the schedule and grid (4..1 quarters).  It
 * keeps the fastest (a later candidate must beat the best so far by 2 %) for
 * this sector and precision; every later H . v of the sector (solvers
 * included) uses it.  best (11 ints, may be NULL): family (as
 * qs3x_hv_last_launch), kernel, stencil, lines, slabs, depth, pairs, chunk,
 * xcd, grid quarters, stencil form (-1: the option's); best_ms its time;
 * results (12 doubles per candidate, up to max_results): the same 11 fields
 * and ms, where ms < 0 is the single
 * timed launch of a candidate over 3x slower than the best so far.  Every
 * candidate computes the same H . v (each kernel and parameter is checked
 * against the oracle by the test suite). */
int qs3x_hv_autotune(qs3x_sector *s, int32_t prec, int32_t n_rep, int32_t *best, double *best_ms,
                     int32_t max_results, int32_t *n_results, double *results);
/* Forgets the autotuned configurations of the sector (both precisions). */
int qs3x_hv_tune_clear(qs3x_sector *s);

#ifdef __cplusplus
}
#endif
#endif /* QS3X_H */
