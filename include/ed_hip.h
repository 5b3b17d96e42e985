/*
 * ed_hip.h -- C-ABI of the MI355X exact-diagonalisation engine (libed_hip.so).
 *
 * Plain C types only (no HIP or torch types): Fortran binds it through
 * ISO_C_BINDING, Python through ctypes.  Every entry point returns an int status
 * (ED_OK = 0); ed_last_error() gives the message of the last failure on the
 * calling thread.  The reference STOPs where this library returns an error.
 *
 * Each declaration names the reference interface it replaces (paths relative to
 * the reference repository, directory ed_solver/).
 *
 * Conventions shared with the reference:
 *   per-spin bit layout     bath site k -> bit k, impurity orbital c -> bit nb+c
 *                           (quantum_algebra_Anderson_Impurity_class.f90:73-86)
 *   spin-sector ordering    states with popcount N in increasing integer order
 *                           (quantum_fermion_hilbert_class.f90:385-404)
 *   vector index            istate = iup + ido * d_up, 0-based (the reference's
 *                           iup + (ido-1)*d_up with 1-based iup, ido)
 *   table ranks             rankoff is 0-based (reference value minus one)
 *   tridiagonal storage     diag[0..n-1]; subdiag[0] = ||v1||, subdiag[1..n-1]
 *                           the off-diagonal (Dagotto storage, matrices_tridiag.f90:55-63)
 */
#ifndef ED_HIP_H
#define ED_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ED_OK 0
#define ED_ERR_ARG (-1)       /* invalid argument (size mismatch, bad sector ...) */
#define ED_ERR_HIP (-2)       /* a HIP runtime call failed */
#define ED_ERR_NOMEM (-3)     /* device allocation failed */
#define ED_ERR_BREAKDOWN (-4) /* zero-norm Lanczos vector: Lanczos_plain.h:70-92 STOPs */
#define ED_ERR_STATE (-5)     /* call out of order (e.g. no Krylov basis kept) */
#define ED_ERR_CF (-6)        /* continued-fraction denominator < eps (matrices_tridiag.f90:174-186) */

#define ED_PREC_F64 0         /* vectors and H.v in FP64 */
#define ED_PREC_F32 1         /* vectors and H.v in FP32, dot products / recurrence scalars in FP64 */
#define ED_PREC_MIXED_RQ 2    /* ed_lanczos_diagonalize only: ED_PREC_F32, then E0 from the FP64
                                 Rayleigh quotient of the FP32 Ritz vector when its bound allows */

/* Anderson impurity model, one real (FP64) parameter set.  All arrays row-major.
 * Replaces the fields of AIM_type / AIM2_type / impurity_type the Hamiltonian
 * reads (quantum_algebra_Anderson_Impurity_class2.f90:62-135, HAIMupdo.f90:53-121). */
typedef struct ed_model {
    int32_t nc;          /* impurity orbitals per spin (Nc)                           */
    int32_t nb;          /* bath sites per spin (Nb); Ns = nc + nb <= 15              */
    const double *ec;    /* [2][nc][nc]  Ec(spin)                                     */
    const double *eb;    /* [2][nb][nb]  Eb(spin)                                     */
    const double *vbc;   /* [2][nb][nc]  Vbc(spin)(k, c)                              */
    const double *u;     /* [nc][nc]     U_ii, and U_ij already reduced by 2J (i != j)*/
    const double *jj;    /* [nc][nc]     JJmatrix                                     */
    double jhund;        /* Jhund: spin flip + pair hopping are on when |Jhund| > 1e-5 */
    double cutoff;       /* cutoff_hamilt_param                                       */
} ed_model;

/* One (N_up, N_dn) sector resident on one GPU: integer tables, H.v kernels and
 * Lanczos work space.  Opaque. */
typedef struct ed_sector ed_sector;

/* S_z (Nambu) sector for a superconducting bath (new_HAIMsz / tab_HAIMsz, HAIMsz.f90:51-177;
 * the Hmult__ branch of H_class.f90:84-93,123-132).  The down spin is particle-hole
 * transformed: a state is 2 Ns bits (up: bath k -> bit k, impurity c -> bit nb + c;
 * down holes at the same positions + Ns) with npart = N_up + (Ns - N_dn) = Ns + 2 S_z set
 * bits, in increasing order.  pb[nb][nb] = Pb (bath pairing b+_k,up b+_l,dn + h.c.; NULL =
 * zero), pvbc[2][nb][nc] = PVbc (impurity-bath pairing, used when pairing_imp_to_bath,
 * PAIRING_IMP_TO_BATH).  Quadratic part: tab_HAIM2 over the Nambu matrices (Nambu_Ec,
 * Nambu_Eb, Nambu_Vbc); diagonal interaction QUART_INT_SZ (with -J for parallel spins
 * when |Jhund| > 1e-5).  With |Jhund| > 1e-5 and some U_cc' != 0 the spin-flip and
 * pair-hopping terms of HAIMsz_multc (HAIMsz.f90:300-406) are added as well, so a real
 * model with J runs the algebra the reference only runs through its COMPLEX(8) path
 * (its real HAIMsz_multr stops on J, :473-476).  Ns <= 13.  The returned sector works with every
 * H.v, Lanczos, block Lanczos, ARPACK-mirror and dense-ED entry point; eigenvalues are
 * those of the Nambu H, the physical energies are E + ed_sector_energy_shift(). */
int ed_sector_create_sz(const ed_model *model, const double *pb, const double *pvbc, int pairing_imp_to_bath,
                        int npart, ed_sector **out);

/* new_HAIMsz of a complex model (the COMPLEX(8) build: tab_HAIM2 with complex entries over
 * the Nambu orbitals, HAIMsz_multc, HAIMsz.f90:188-452; the _complex branches of Nambu_Ec /
 * Nambu_Eb / Nambu_Vbc, impurity_class.f90:1018-1037, bath_class.f90:956-1035: -Ec_dn^T,
 * -Eb_dn^T, conj(Pb)^T below the diagonal, -conj(Vbc_dn)).  model holds the real parts;
 * ec_im[2][nc][nc] (required), eb_im[2][nb][nb], vbc_im[2][nb][nc], pb_im[nb][nb],
 * pvbc_im[2][nb][nc] the imaginary parts (NULL = 0).  Vectors are complex as for
 * ed_sector_create_c (planar in the real-typed entry points, COMPLEX(8) in
 * ed_hmult_c128); FP64 only; ed_sector_set_fly selects HAIMsz_mult_fly_c. */
int ed_sector_create_sz_c(const ed_model *model, const double *ec_im, const double *eb_im, const double *vbc_im,
                          const double *pb, const double *pb_im, const double *pvbc, const double *pvbc_im,
                          int pairing_imp_to_bath, int npart, ed_sector **out);

/* ON_FLY (H_class.f90:84-93 -> HAIMsz_mult_fly, HAIMsz.f90:615-1027) for an S_z sector:
 * on = 1 makes every later H.v generate each row's tab_HAIM2 entries from the Nambu
 * matrices instead of reading the stored table (same entries, same order, same bits);
 * on = 0 returns to the table. */
int ed_sector_set_fly(ed_sector *sec, int on);

/* Hund transitions of an S_z sector as an ELL: *kh per state (0 without J), rk[kh][dim]
 * target ranks and val[kh][dim] = -J sign in the order of HAIMsz_multc (HAIMsz.f90:300-406),
 * rows padded with the row itself and 0.  rk / val may be NULL (query *kh). */
int ed_sector_hund_sz(const ed_sector *sec, int *kh, int32_t *rk, double *val);

/* QUART_INT_SZ of an S_z sector (HAIMsz.f90:129-177), one value per state [dimen]. */
int ed_sector_quart_sz(const ed_sector *sec, double *quart);

/* energy_global_shift + energy_global_shift2 (impurity_class.f90:1037, bath_class.f90:976):
 * tr(Ec_dn) + tr(Eb_dn) for an S_z sector, 0 for an (N_up, N_dn) sector. */
int ed_sector_energy_shift(const ed_sector *sec, double *shift);

/* Error text of the last failing call on this thread (never NULL). */
const char *ed_last_error(void);

/* Select the device for subsequent calls on this thread.
 * Replaces init_gpu_device / choose_gpu (core_functions_initsim_closesim.f90:131). */
int ed_init(int device);

/* Library version string, e.g. "ed_hip 0.1 gfx950". */
const char *ed_version(void);

/* ------------------------------------------------------------------ sectors */

/* new_H(AIM, sector) -> new_HAIMupdo (H_class.f90:28-37, HAIMupdo.f90:53-121):
 * enumerates both spin sectors, builds rank tables and tab_HAIM2 tables
 * (HAIM_offdiag_arrays.f90:89-148, 210-606) with integer kernels on the device. */
int ed_sector_create(const ed_model *model, int nup, int ndn, ed_sector **out);

/* new_H with the Slater 4-index impurity interaction (flag_slater_int, HAIMupdo.f90:
 * 905-999; Slater_Coulomb_r): u4[nc][nc][nc][nc] row-major, U(s1, s2, s3, s4) at
 * ((s1 nc + s2) nc + s3) nc + s4, added to the model's U / J terms.  The H.v gathers
 * w_i += U(s1, s2, s3, s4) sign v_j over the states j reached from i by destroying site4
 * (spin is2) and site3 (spin is1), then creating site2 (spin is1) and site1 (spin is2),
 * for both spins is1, is2, with the reference's signs (orbitals occupied below site4 /
 * site1 in the is2 string, strictly between site2 and site3 in the is1 string).  The term
 * acts on impurity bits only and is tabulated per impurity configuration pair: its
 * diagonal joins the density terms, its transitions the Hund list, so every H.v and
 * Lanczos kernel takes it.  ed_hf_gradient rejects such sectors (its U derivative is of
 * the Kanamori form). */
int ed_sector_create_slater(const ed_model *model, const double *u4, int nup, int ndn, ed_sector **out);

/* new_H for a complex AIM (the COMPLEX(8) branch of tab_HAIM2 and HAIMupdo_multc,
 * HAIMupdo.f90:150-656): the model's ec/eb/vbc are the real parts and ec_im[2][nc][nc],
 * eb_im[2][nb][nb], vbc_im[2][nb][nc] the imaginary parts (eb_im, vbc_im may be NULL =
 * 0).  Ec and Eb must be Hermitian per spin.  Entries are kept when nonzero and
 * |t| > cutoff (complex modulus); hybridisation entries are Vbc (impurity -> bath) and
 * conj(Vbc) (bath -> impurity).  ed_hmult_c128 applies H to COMPLEX(8) vectors.  Every
 * real-typed vector of the other entry points (Lanczos start and ground state, apply_C,
 * dynamic, ed_hmult_f64) has length 2 dimen_H() and holds the planar form
 * (Re[dimen_H], Im[dimen_H]): H acts there as the symmetric [[A, -B], [B, A]] of
 * H = A + iB, whose plain Lanczos has the complex Lanczos' (real) alpha and beta.  FP64
 * only; ed_full_ed gives the complex eigenvectors (planar) up to 512 amplitudes; block
 * Lanczos, ed_dynamic_from_state and the Hellmann-Feynman gradients take real sectors
 * only. */
int ed_sector_create_c(const ed_model *model, const double *ec_im, const double *eb_im, const double *vbc_im,
                       int nup, int ndn, ed_sector **out);

/* delete_H (H_class.f90:48-58).  NULL is accepted. */
int ed_sector_destroy(ed_sector *sec);

/* dimen_H (H_class.f90:149-152) = d_up * d_dn (the number of complex amplitudes of a
 * complex sector). */
int64_t ed_sector_dim(const ed_sector *sec);

/* Spin-sector dimensions d_up, d_dn (fermion_sector2 up%dimen, down%dimen). */
int ed_sector_spin_dims(const ed_sector *sec, int32_t *d_up, int32_t *d_dn);

/* Number of tabulated off-diagonal entries of one spin (noffdiagup / noffdiagdo,
 * HAIM_offdiag_arrays.f90:37). */
int ed_sector_nnz(const ed_sector *sec, int spin, int64_t *nnz);

/* Copy one spin's tables to host memory (any pointer may be NULL):
 * states[d], diag[d], noff[d], rankoff[nnz] (0-based), offdiag[nnz]
 * (sector%state and diagup/noffup/rankoffup/offdiagup, HAIM_offdiag_arrays.f90:20-44). */
int ed_sector_tables(const ed_sector *sec, int spin, int32_t *states, double *diag,
                     int32_t *noff, int32_t *rankoff, double *offdiag);

/* Imaginary parts of one spin's offdiag[nnz] (complex sectors; offdiagup is COMPLEX(8)
 * there, HAIM_offdiag_arrays.f90:20-44). */
int ed_sector_tables_im(const ed_sector *sec, int spin, double *offdiag_im);

/* ---------------------------------------------------------------------- H.v */

/* Hmultr (H_class.f90:109-136) -> HAIMupdo_multr (HAIMupdo.f90:677-1195):
 * vec_out = H vec_in on host arrays of length n = dimen_H().  Synchronous; copies
 * through device buffers owned by the sector. */
int ed_hmult_f64(ed_sector *sec, int64_t n, double *vec_out, const double *vec_in);

/* Device-resident H.v (the fast path): d_out / d_in are device pointers of length
 * n; `stream` is a hipStream_t or NULL for the sector's own stream.  Asynchronous. */
int ed_hmult_f64_dev(ed_sector *sec, int64_t n, double *d_out, const double *d_in, void *stream);

/* FP32 H.v on device vectors (tables rounded to FP32). */
int ed_hmult_f32_dev(ed_sector *sec, int64_t n, float *d_out, const float *d_in, void *stream);

/* Hmultc (H_class.f90:69-98) -> HAIMupdo_multc (HAIMupdo.f90:150-656) on a complex
 * sector: vec_out = H vec_in for COMPLEX(8) arrays (re, im interleaved) of n = dimen_H()
 * amplitudes; w_i = sum_j conj(<j|H|i>) v_j.  Host arrays, synchronous. */
int ed_hmult_c128(ed_sector *sec, int64_t n, double *vec_out, const double *vec_in);

/* Hmultc on device COMPLEX(8) vectors (2 n doubles each; d_out may equal d_in);
 * `stream` as for ed_hmult_f64_dev.  Asynchronous. */
int ed_hmult_c128_dev(ed_sector *sec, int64_t n, double *d_out, const double *d_in, void *stream);

/* Waits for all work queued on the sector's own stream (the asynchronous device
 * entry points above use it when called with stream == NULL). */
int ed_sector_synchronize(ed_sector *sec);

/* Block H.v on nvec device vectors stored column by column with leading dimension
 * ld (>= n): V(:, 1:nvec) = H U(:, 1:nvec), the LFLAG=1 callback of
 * Block_Lanczos_diagonalize (Block_Lanczos.f90:189-259).  FP64. */
int ed_hmult_block_f64_dev(ed_sector *sec, int64_t n, int nvec, int64_t ld, double *d_out,
                           const double *d_in, void *stream);

/* Autotuner of the H.v launch schedule (no reference counterpart): times the FP64
 * H.v on this sector under each applicable schedule -- the row kernel with 1, 2 or 4
 * columns per workgroup, the fast kernel (one column per workgroup, up to 1024
 * threads, packed tables in LDS) and the fast kernel with its workgroups mapped to
 * columns in an XCD-grouped order (neighbouring columns of the down-hop graph run on
 * the same XCD and share its L2), and that kernel with 2, 3 or 4 rows per thread
 * (smaller workgroups with every row's loads in flight) -- and keeps the fastest for
 * every later H.v and Lanczos call on the sector.  *choice = 1, 2, 4 (row kernel), 8
 * (fast kernel), 16 (fast kernel, XCD-grouped order) or 18, 19, 20 (XCD-grouped, 2, 3, 4
 * rows per thread); ms_per_choice[8] (may be NULL) = measured times in that order (0
 * where a schedule does not apply).  Lanczos calls run it on first use. */
int ed_hmult_tune(ed_sector *sec, int *choice, double *ms_per_choice);

/* Forces one H.v schedule (1, 2, 4, 8, 16, 18, 19 or 20 as above) on this sector, e.g. to test or time
 * a schedule the autotuner did not pick.  No reference counterpart. */
int ed_hmult_set_schedule(ed_sector *sec, int choice);

/* Times `reps` fixed-length runs of `niter` Lanczos steps (after one warm-up run), each
 * from the default start vector: a HIP event pair on the sector's stream around the
 * steps of a run (the persistent kernel's one launch, or the captured two-kernel chunks).
 * *avg_ms = average per run; *bytes_per_run = its algorithmic HBM bytes 5 * sizeof(T) *
 * dimen_H() * niter (per step: read u_j and write H u_j in the H.v, read u_{j-1}, u_j and
 * write u_{j+1} in the recurrence); *persistent = 1 when the persistent kernel ran.  No
 * reference counterpart: a measurement hook. */
int ed_lanczos_benchmark(ed_sector *sec, int precision, int niter, int reps, double *avg_ms, double *bytes_per_run,
                         int *persistent);

/* Persistent Lanczos kernel (no reference counterpart; an engine option).  Where the
 * sector's H.v schedule is a fast-kernel one (8, 16, 18-20), d_up <= 1024 and every
 * workgroup of the grid fits on the device at once, fixed-length runs of at least 16
 * steps (ed_lanczos_run, the excitation runs of ed_dynamic) and the ground-state solves
 * of ed_lanczos_diagonalize execute in one cooperative launch: H.v and update separated
 * by grid barriers, the static row data in registers and LDS, the vectors exchanged with
 * device-coherent loads and stores.  Tables of width <= 12 with at most 2 Hund
 * transitions per row use the multi-column form (up to 4 columns per workgroup of 1024
 * threads, the row tables shared in LDS); others the one-column form.  Where each thread
 * owns one row (d_up <= 256) the multi-column form runs ONE grid barrier per step by
 * default: it exchanges the vector before the alpha_j u_j subtraction, rebuilds u_{j+1}
 * bit for bit wherever it is gathered, and forms beta_{j+1} from the expansion of its norm
 * (the norm itself where the expansion cancels), so alpha and beta agree with the
 * two-kernel step to rounding (1e-14 over the first 80 steps at d_up = 252; later steps
 * drift apart as those of any two Lanczos runs whose rounding differs, with the same Ritz
 * values).  Its two-barrier form and the one-column form keep the two-kernel step's arithmetic and
 * order of every sum, so alpha and beta are the same numbers.  enable = 1 (default): use
 * it where it applies, one-barrier form; 0: never; 2 (a test hook): as 1, but every
 * cooperative launch takes the path of a launch the runtime refuses; 3: the one-column
 * form only; 4: the two-barrier multi-column form; 5 (a test hook): the one-barrier form
 * with every beta from the norm (one more barrier per step).  A barrier wait
 * that exceeds a few seconds stops the run with ED_ERR_HIP and turns the kernel off for
 * the sector.  A cooperative launch the runtime refuses (hipErrorCooperativeLaunchTooLarge,
 * hipErrorIllegalState: nothing ran) turns it off for that sector and precision, is
 * counted, and the run continues with the two-kernel step; any other launch error
 * returns ED_ERR_HIP. */
int ed_lanczos_set_persistent(ed_sector *sec, int enable);

/* Launch shape of the persistent kernel for `precision` on this sector (no reference
 * counterpart; for reports): columns per workgroup (0: the kernel does not apply; 1 with
 * the one-column kernel), workgroups, rows per thread, compiled table width.  Any output
 * may be NULL. */
int ed_lanczos_persistent_shape(ed_sector *sec, int precision, int *cols_per_wg, int *workgroups, int *rows,
                                int *width);

/* Cooperative launches of the persistent kernel the runtime refused on this sector (no
 * reference counterpart; lets a caller tell a persistent run from its fallback). */
int ed_lanczos_persistent_refusals(ed_sector *sec, int *refused);

/* Steps of the last launch of the one-barrier persistent kernel on this sector whose
 * beta_{j+1} came from the reduced norm of u_{j+1} (one more grid barrier) rather than
 * from the expansion of that norm, which had cancelled (no reference counterpart; a
 * measurement hook).  0 when the sector never ran that kernel. */
int ed_lanczos_exact_beta_steps(ed_sector *sec, int *steps);

/* *active = 1 when a fixed-length Lanczos run of at least 16 steps in `precision` on
 * this sector uses the persistent kernel (tunes the sector first if needed), else 0. */
int ed_lanczos_persistent_active(ed_sector *sec, int precision, int *active);

/* ------------------------------------------------------------------ Lanczos */

/* Deterministic start vector (period 10000, values k/1000 in (0, 1]) used when a
 * Lanczos entry point gets start == NULL.  It takes the place of
 * create_fix_initial_vector (rcvector_class.f90:43-60), whose gfortran random
 * stream cannot be reproduced; Fortran callers may pass their own vector. */
int ed_start_vector(int64_t n, double *out);

typedef struct ed_lanczos_opts {
    int nitermax;        /* Nitermax (default 128)                                    */
    int neigen;          /* Neigen (default 1)                                        */
    double tolerance;    /* convergence of |lambda_k(T_j) - lambda_k(T_j-1)| (1e-12)  */
    int precision;       /* ED_PREC_F64 or ED_PREC_F32                                */
    int check_every;     /* iterations between host convergence checks (default 8)    */
    int reorth;          /* 0: plain (the reference); 1: full reorthogonalisation (CGS2) */
} ed_lanczos_opts;

/* Fills *opts with the reference defaults (dmft_ed_solver_variables.h:9-20). */
int ed_lanczos_default_opts(ed_lanczos_opts *opts);

/* Lanczos_fast_diagonalize (Lanczos_plain.f90:33-134): plain Lanczos from `start`
 * (host, length dimen_H(), or NULL for ed_start_vector) with the reference's
 * convergence test.  Outputs (host): *iters; evals[neigen'] (neigen' = min(neigen,
 * dim, niter)); vecp[neigen'][*iters] Lanczos coefficients (row k = eigenvector
 * k of T); diag/subdiag[*iters] (may be NULL, capacity nitermax).  The Krylov basis
 * stays on the device for ed_lanczos_gs_vector.  precision ED_PREC_F32 selects the
 * mixed form: FP32-vector Lanczos to |dlambda| < max(tolerance, 1e-6), then an FP64
 * Lanczos restarted from that Ritz vector to `tolerance`; the outputs describe the
 * FP64 phase (ed_lanczos_info gives the FP32 phase's iteration count). */
int ed_lanczos_diagonalize(ed_sector *sec, const ed_lanczos_opts *opts, const double *start,
                           int *iters, int *neig_out, double *evals, double *vecp,
                           double *diag, double *subdiag);

/* Iterations of the FP32 phase of the last mixed-precision ed_lanczos_diagonalize
 * on this sector (0 after an FP64 call).  No reference counterpart. */
int ed_lanczos_info(ed_sector *sec, int *iters_f32_phase);

/* The last ED_PREC_MIXED_RQ solve on this sector (no reference counterpart): with one
 * eigenvalue asked for, the FP32 phase runs until the Lanczos estimate of its Ritz
 * residual meets the target below, then its Ritz vector psi gets one FP64 H.v; the
 * Rayleigh quotient E replaces the FP64 restart when res2 = ||H psi - E psi||^2 /
 * ||psi||^2 <= tolerance * gap / 2, gap = l_2 - E with l_2 = theta_2 - rho_2 (the second
 * Ritz value of the FP32 T less its residual estimate), the lower estimate of lambda_2 in
 * the Kato-Temple bound E - E0 <= res2 / (l_2 - E).  The FP32 phase stops waiting for
 * the target, and the FP64 restart runs, when the target lies below what FP32 vectors
 * reach ((4 eps32 ||T||)^2 > tolerance * gap / 2), when theta_2 collapses onto theta_1, or
 * after max(16, n) steps beyond its eigenvalue convergence at step n.  *accepted = 1
 * when E was taken, and the outputs of ed_lanczos_diagonalize are those of the FP32 phase
 * with evals[0] = E (ed_lanczos_gs_vector gives psi from the FP32 basis); otherwise the
 * call continued as ED_PREC_F32 (res2 and gap are 0 when the test did not run). */
int ed_lanczos_rq_info(ed_sector *sec, int *accepted, double *res2, double *gap);

/* Lanczos steps the device ran in the last ed_lanczos_diagonalize on this sector (both
 * phases of the mixed form).  Where the persistent kernel applies, one launch runs while
 * the host tests convergence on the scalars it publishes and is stopped a few steps
 * after the iteration that ends the loop, so this can exceed the iterations returned;
 * those extra steps change none of the outputs.  No reference counterpart. */
int ed_lanczos_steps_run(ed_sector *sec, int *steps);

/* ED_diago (ED_ARPACK.f90:590-687), the path of sectors of dimension <= FULL_ED_SWITCH
 * (solver.h:156-180): dense H from H e_i on the device, all eigenpairs on the host
 * (cyclic Jacobi), ascending.  evals[neig]; evecs[neig][dim] (may be NULL).  dim <= 1024.
 * A complex sector (dimen_H <= 512): the real embedding of H is diagonalised, each
 * doubled eigenvalue taken once, and evecs rows are planar [neig][2 dimen_H]. */
int ed_full_ed(ed_sector *sec, int neig, double *evals, double *evecs);

/* Lanczos_GPU_updo (Lanczos_gpu.f90:106): exactly `niter` Lanczos steps (no
 * convergence test) from `start` (host or NULL); writes diag[niter], subdiag[niter].
 * Returns ED_ERR_BREAKDOWN when a zero-norm vector appears before niter. */
int ed_lanczos_run(ed_sector *sec, int niter, const double *start, int precision,
                   double *diag, double *subdiag);

/* Lanczos_GPU_get_GS_updo (Lanczos_gpu.f90:48) / Lanczos_get_GS_sector
 * (Lanczos_plain.f90:156-229): psi = normalised sum_j vecp[j] v_j over the Krylov
 * basis kept by the last Lanczos call on this sector.  A sector that keeps no basis of
 * `iters` vectors (a fresh handle: the reference calls new_H again before this step,
 * solver.h:43-49) runs the reference's second pass: the FP64 recurrence again from
 * ed_start_vector (create_fix_initial_vector, Lanczos_plain.f90:196), so vecp must come
 * from an FP64 ed_lanczos_diagonalize / ed_lanczos_run started from the default vector.
 * psi is a host array of length dimen_H() (FP64). */
int ed_lanczos_gs_vector(ed_sector *sec, int iters, const double *vecp, double *psi);

/* Same, result left on the device (d_psi, FP64). */
int ed_lanczos_gs_vector_dev(ed_sector *sec, int iters, const double *vecp, double *d_psi);

/* Rayleigh quotient <psi|H|psi>/<psi|psi> in FP64 on a host vector (the FP64
 * polish of the mixed-precision path). */
int ed_rayleigh_f64(ed_sector *sec, int64_t n, const double *psi, double *energy);

/* Eigenpairs of the symmetric tridiagonal T (diag[n], subdiag[1..n-1]; subdiag[0]
 * unused), ascending: w[n], z[k*n + r] = component r of eigenvector k (sign: largest
 * component positive).  The DSTEDC call of diagonalize_tridiag
 * (matrices_tridiag.f90:210-259), host FP64. */
int ed_tridiag_eig(int n, const double *diag, const double *subdiag, double *w, double *z);

/* The k lowest eigenvalues of T_n (diag[n], subdiag[1..n-1]), ascending, as the
 * Lanczos convergence test computes them every step: the DSTEDC 'N' call of
 * Lanczos_fast_diagonalize (Lanczos_plain.f90:33-134, matrices_tridiag.f90:210-259).
 * prev[k] (may be NULL) holds those of T_{n-1} (1e300 where unknown); by interlacing they
 * bound the new ones from above and seed Newton's method for the lowest.  Each value to
 * about 2 eps |lambda|; entries k >= n are 1e300. */
int ed_tridiag_lowest(int n, const double *diag, const double *subdiag, int k, const double *prev, double *w);

/* nev lowest eigenpairs of a symmetric band matrix a (n x n, row-major, both
 * triangles, half-bandwidth b): w[nev] ascending, z (n x nev, column-major, may be
 * NULL; sign: largest component positive).  The Ritz step of the block Lanczos driver,
 * standing in for the one inside BLZPACK's BLZDRD (Block_Lanczos.f90:132,195), host
 * FP64: band reduction by Givens rotations, implicit QL, inverse iteration. */
int ed_band_eig(int n, int b, const double *a, int nev, double *w, double *z);

/* ----------------------------------------------------------- block Lanczos */

/* Block_Lanczos_diagonalize (Block_Lanczos.f90:37-355): the Neigen lowest eigenpairs of
 * the sector by a block Lanczos of block size NVB (Block_size, :64-68), at most
 * max_steps block steps (ISTOR(6) = min(Nitermax, dimen_H), :63).  BLZPACK (BLZDRD) is
 * replaced by a device driver: block H.v, full reorthogonalisation (classical
 * Gram-Schmidt twice) and Cholesky-QR2 from FP64 MFMA contractions, Ritz values of the
 * band matrix T on the host.  Converged when every wanted Ritz value moved by less than
 * `tolerance` (RSTOR(3), :82) between two tests `check_every` block steps apart and its
 * residual bound is below sqrt(tolerance) * max(1, |E|). */
typedef struct ed_block_opts {
    int block_size;      /* NVB, 1..16                                           */
    int nev;             /* Neigen, 1..16                                        */
    int max_steps;       /* block steps                                          */
    int check_every;     /* Ritz test every this many block steps                */
    double tolerance;
    int reorth;          /* 1: partial reorthogonalisation (block omega-recurrence;
                            the new block is projected against all earlier blocks only
                            when the estimated loss of orthogonality exceeds sqrt(eps),
                            and at the step after), as BLZPACK; 0: against every
                            earlier block at every step                              */
} ed_block_opts;

int ed_block_default_opts(ed_block_opts *opts);   /* 8, 8, 128, 2, 1e-12, 1 */

/* evals[nev] ascending (VALP(:,1)), resid[nev] = ||H x_k - E_k x_k|| (VALP(:,2), may be
 * NULL), evecs = host D x nev column-major with leading dimension ldv (VECP, may be
 * NULL), *nconv = eigenpairs whose residual meets the test (ISTOR(67)), *steps = block
 * steps (ISTOR(19)). */
int ed_block_lanczos(ed_sector *sec, const ed_block_opts *opts, double *evals, double *resid, double *evecs,
                     int64_t ldv, int *nconv, int *steps);
/* Same, with the selection of Block_Lanczos_diagonalize (Block_Lanczos.f90:289-300): only
 * the eigenpairs within dEmax of the lowest one are kept, and only their eigenvectors are
 * copied to the host (the leading *nkept rows of evecs, the values being ascending; the
 * other rows are not written).  evals / resid still receive all nev values.  dEmax >= 0. */
int ed_block_lanczos_window(ed_sector *sec, const ed_block_opts *opts, double dEmax, double *evals, double *resid,
                            double *evecs, int64_t ldv, int *nconv, int *steps, int *nkept);
/* Same, eigenvectors written to device memory d_evecs (leading dimension ldv). */
int ed_block_lanczos_dev(ed_sector *sec, const ed_block_opts *opts, double *evals, double *resid, double *d_evecs,
                         int64_t ldv, int *nconv, int *steps);

/* ARPACK_diago (ED_ARPACK.f90:475-570) through arpack_eigenvector_sym_matrix
 * (ED_ARPACK.f90:35-250): the implicitly restarted Lanczos method of dsaupd / dseupd
 * (bmat = 'I', mode 1, exact shifts) with the basis, the reorthogonalisation and the
 * restart rotation on the device; only the ncv x ncv tridiagonal work runs on the host.
 * The reference passes which = 'BE', nev = Neigen_arpack - 1, ncv = Neigen_arpack,
 * tol = tolerance, maxitr = 300 (ED_ARPACK_BE keeps nev / 2 Ritz values from the low end
 * of the spectrum and the rest from the high end, as dsgets; _SA the nev smallest; _LA the
 * nev largest).  Outputs: evals[nconv] ascending (the converged wanted Ritz values, d(:,1)
 * of dseupd), resid[nconv] = ||H x - E x|| / |E| (d(:,2), may be NULL), evecs = host
 * nconv x dimen_H() with leading dimension ldv (VECP, may be NULL), *nconv = iparam(5),
 * *iters = implicit restarts + 1 (iparam(3)), *matvecs = H.v applications (iparam(9)).
 * Fewer than nev converged within maxitr is not an error (info = 1 of dsaupd): *nconv
 * says how many.  Real sectors ((N_up, N_dn) and S_z); ED_ERR_ARG unless
 * 1 <= nev < ncv <= min(dimen_H(), 1024). */
#define ED_ARPACK_BE 0
#define ED_ARPACK_SA 1
#define ED_ARPACK_LA 2
typedef struct ed_arpack_opts {
    int nev;          /* wanted Ritz values (NEV)                  */
    int ncv;          /* basis length (NCV)                         */
    int which;        /* ED_ARPACK_BE / _SA / _LA                   */
    double tol;       /* relative accuracy of the Ritz values (TOL) */
    int maxitr;       /* implicit restarts allowed (iparam(3))      */
} ed_arpack_opts;

int ed_arpack_default_opts(ed_arpack_opts *opts);   /* 8, 9, BE, 1e-12, 300 */

int ed_arpack(ed_sector *sec, const ed_arpack_opts *opts, const double *start, double *evals, double *resid,
              double *evecs, int64_t ldv, int *nconv, int *iters, int *matvecs);

/* The last ed_block_lanczos call: block steps, the steps that projected the new block
 * against every earlier block (all of them with reorth = 0), the call's host wall time and
 * the part of it spent in the Ritz tests (the band solver on T).  No reference
 * counterpart: it tells a caller what partial reorthogonalisation saved and where the
 * time goes.  Any output may be NULL. */
int ed_block_last_stats(int *steps, int *full_passes, double *ms, double *ritz_ms);

/* The contractions of the block orthogonalisation, on device matrices (column-major):
 *   ed_gram_f64_dev:  g (host, row-major p x q) = X^T Y, X n x p, Y n x q, q <= 16
 *                     (16-byte aligned pointers, even leading dimensions);
 *   ed_xc_f64_dev:    Y = beta Y + alpha X C, C host row-major p x q, q <= 16
 *                     (Y may be X when p == q).
 * FP64 MFMA; synchronous.  No reference counterpart (they are the dense
 * products inside BLZDRD). */
int ed_gram_f64_dev(int64_t n, int p, int q, const double *d_x, int64_t ldx, const double *d_y, int64_t ldy, double *g,
                    void *stream);
int ed_xc_f64_dev(int64_t n, int p, int q, const double *d_x, int64_t ldx, const double *c, double alpha, double beta,
                  double *d_y, int64_t ldy, void *stream);

/* ---------------------------------------------------------- derivatives */

/* Hellmann-Feynman gradient dE/dtheta = <psi|dH/dtheta|psi> / <psi|psi> for an
 * eigenvector psi of the sector (FP64, host vector of length n = dimen_H(), or device
 * vector for the _dev form).  d_level[ns] = derivatives in the site levels (bit a:
 * bath site a < nb -> Eb(a,a), impurity orbital c = a - nb -> Ec(c,c); both spins),
 * d_v[nb*nc] (row-major) = derivatives in the real hybridisations Vbc(k,c) (both spins;
 * the term Vbc b+_k c_c + h.c. of HAIM_offdiag_arrays.f90:483-527), *d_u = derivative in
 * the interaction scale U of U_ii = U, U_ij = U - 2J at fixed J (HAIMupdo.f90:785-829).
 * Any output may be NULL.  No reference counterpart: the reference has no derivative
 * of the energy; it stands beside the finite differences of the bath fit. */
int ed_hf_gradient(ed_sector *sec, int64_t n, const double *psi, double *d_level, double *d_v, double *d_u);
int ed_hf_gradient_dev(ed_sector *sec, const double *d_psi, double *d_level, double *d_v, double *d_u);

/* distance_func / diff_correl (bath_class_hybrid.f90:1151-1204, correl_class.f90:
 * 1061-1160) for a diagonal bath eps[nb] with real couplings vbc[nb*nc] (row-major,
 * Vbc(k,c)), shared by both spins: Delta_ij(i w_n) = sum_k Vbc(k,i) Vbc(k,j)/(i w_n - eps_k)
 * (bath2hybrid, bath_class_hybrid.f90:587-903, normal state), target T[nw][2][nc][nc]
 * complex (re, im interleaved; spin up, spin down), wn[nw] the Matsubara frequencies:
 *   S = sum_n sum_{sigma,i,j} |Delta_ij - T^sigma_ij|^weight_expo / (|w_n - w_1| + shift)^fit_weight_power
 *   *dist = |S| / (nw (2 nc)^2),  shift = fit_shift, or |w_1| when |fit_shift| < 1e-10.
 * grad_eps[nb], grad_v[nb*nc] (may be NULL) = the analytic gradient of *dist (the
 * reference takes finite differences, conj_grad_civelli.f90:282-303).  Device kernel. */
int ed_bath_distance(int nc, int nb, const double *eps, const double *vbc, int nw, const double *wn,
                     const double *target, double weight_expo, double fit_weight_power, double fit_shift, double *dist,
                     double *grad_eps, double *grad_v);

/* --------------------------------------------------- Green's functions */

/* apply_Cupdo (apply_C.f90:497-598): phi = c+ (pm = +1) or c (pm = -1) of impurity
 * orbital `site` (0-based), spin 0 (up) / 1 (down), applied to psi of sector `src`;
 * phi lives in sector `dst` ((nup+-1, ndn) or (nup, ndn+-1)).  Host arrays.
 * S_z sectors (apply_Csz / apply_Csz_updo, apply_C.f90:277-388): creates (pm = +1) or
 * destroys the Nambu orbital (site, spin), bit spin*Ns + nb + site, with the sign of the
 * Nambu orbitals occupied above it; dst is the npart + pm sector.  The Nambu orbital of
 * spin 1 is the particle-hole transformed down orbital: pm = +1 there is the physical
 * c_dn (the reference's apply_Cdo with '-'), pm = -1 the physical c+_dn. */
int ed_apply_c(const ed_sector *src, const ed_sector *dst, int site, int spin, int pm,
               const double *psi, double *phi);

/* apply_Csz_N_updo (apply_C.f90:404-483; apply_N_Cup / apply_N_Cdo :134-145,168-179),
 * the correlated-hopping excitation of the FLAG_NCUP Green's functions: ed_apply_c on an
 * S_z sector restricted to the states whose other Nambu orbital of the same site holds
 * the reference's condition.  spin 0 acts on the Nambu up orbital where the Nambu down
 * orbital is empty (NDN == 0, a physical down electron: c+_up n_dn and n_dn c_up); spin 1
 * acts on the Nambu down orbital where the Nambu up orbital is occupied (NUP == 1).  pm
 * creates or destroys the Nambu orbital as in ed_apply_c (the reference's spin-2 '+' is
 * pm = -1 here).  (N_up, N_dn) sectors give ED_ERR_ARG, where the reference STOPs.
 * Host arrays. */
int ed_apply_c_n(const ed_sector *src, const ed_sector *dst, int site, int spin, int pm,
                 const double *psi, double *phi);

/* invert_zmtridiag (matrices_tridiag.f90:163-198) for nw points z = i*wn[k]:
 * out[2k], out[2k+1] = Re, Im of <0|(z - T)^-1|0>; T = (a[0..n-1], b[1..n-1]).
 * Evaluated on the device, one thread per frequency. */
int ed_continued_fraction(int n, const double *a, const double *b, int nw, const double *wn,
                          double *out);

/* compute_dynamic_cpu (green_class_compute_dynamic.f90:191-407), fermionic, real:
 * niter = min(nitergreen, dim) Lanczos steps from phi/||phi|| in sector `sec`,
 * then dyn(i wn) = ||phi||^2 CF(i wn) of T' = ph (T - energy), ph = +1 for
 * iph = 1 (particle), -1 for iph = 2 (hole).  dyn is complex interleaved [2*nw].
 * ||phi||^2 < cutoff_dynamic gives dyn = 0.  a/b (may be NULL, capacity niter)
 * receive the raw tridiagonal. */
int ed_dynamic(ed_sector *sec, int iph, double energy, const double *phi, int nitergreen,
               double cutoff_dynamic, int precision, int nw, const double *wn, double *dyn,
               double *a, double *b);

/* ed_dynamic for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same arguments, meaning, outputs and errors as ed_dynamic; the excitation
 * Lanczos runs as plain two-kernel steps in one static device array of about 34 MB per
 * GPU (held until the process ends) instead of the sector's own work space, which it
 * leaves untouched.  It calls no allocating routine.  Limits: dimen of `sec` (2 dimen_H
 * for a complex sector) <= ED_NOALLOC_MAX_DIM, min(nitergreen, dimen) <=
 * ED_NOALLOC_MAX_ITER, nw <= ED_NOALLOC_MAX_NW; beyond them it returns ED_ERR_ARG.  Every
 * call on one GPU shares that GPU's array (the last call's contents stay there): calls
 * that overlap on one GPU race. */
#define ED_NOALLOC_MAX_DIM 1048576
#define ED_NOALLOC_MAX_ITER 1024
#define ED_NOALLOC_MAX_NW 16384
int ed_dynamic_noalloc(ed_sector *sec, int iph, double energy, const double *phi, int nitergreen,
                       double cutoff_dynamic, int precision, int nw, const double *wn, double *dyn,
                       double *a, double *b);

/* compute_dynamic_full_ed (green_class_compute_dynamic.f90:77-165), the dynamic
 * correlation of a target sector of dimension dim <= FULL_ED_SWITCH (40,
 * Global_Variables.f90:23) from the nst eigenpairs stored for that sector (evals[nst],
 * evecs[nst][dim] row per state; the reference uses the states its eigen list holds,
 * which are all of them only with FLAG_FULL_ED_GREEN):
 *   dyn(i wn) = sum_n |<n|phi>|^2 / (i wn - ph (E_n - energy)),  ph = +1 (iph = 1) / -1.
 * ||phi||^2 < cutoff_dynamic or nst = 0 gives dyn = 0.  Overlaps and the Lehmann sum on
 * the device; dyn complex interleaved [2*nw]. */
int ed_dynamic_full_ed(int dim, int iph, double energy, const double *phi, int nst, const double *evals,
                       const double *evecs, double cutoff_dynamic, int nw, const double *wn, double *dyn);

/* The same for a complex sector: phi[2 dim] and the rows of evecs[nst][2 dim] planar
 * (Re[dim], Im[dim]); the weights are |<n|phi>|^2 with the complex inner product. */
int ed_dynamic_full_ed_c(int dim, int iph, double energy, const double *phi, int nst, const double *evals,
                         const double *evecs, double cutoff_dynamic, int nw, const double *wn, double *dyn);

/* Batched excitation Lanczos (no reference counterpart as one call: it fuses the
 * per-(state, orbital, iph) calls of compute_dynamic made by
 * green_class_computeAA.h:101-302): njobs independent runs, job k in sector
 * secs[k] from the host vector phis[k] (length dimen of that sector), all advanced
 * together.  Runs in real sectors with d_up <= 1024 and table width <= 12 go through
 * ONE cooperative persistent launch for all of them (a table gives each run workgroups
 * in proportion to its columns); the others, or all when the persistent kernel is
 * switched off on a sector (ed_lanczos_set_persistent 0) or the runtime refuses the
 * launch, take one batched kernel per recurrence phase.  Per job: iph[k], energy[k];
 * output dyn[k][2*nw], a[k][nitergreen], b[k][nitergreen]. */
int ed_dynamic_batched(int njobs, ed_sector *const *secs, const double *const *phis,
                       const int *iph, const double *energy, int nitergreen,
                       double cutoff_dynamic, int precision, int nw, const double *wn,
                       double *dyn, double *a, double *b);

/* The excitation runs of one eigenstate in compute_greenAA (loop_over_states /
 * apply_creat / compute_dynamic, green_class_computeAA.h:101-302) with every vector on
 * the device: psi (host, dimen of gs) is copied once; job k builds
 * phi_k = sum_t coefs[t] c^(pm)_{sites[t], spin} psi in sector dst[k] (pm = +1 for
 * iph[k] = 1, -1 for iph[k] = 2; nterm[k] consecutive (site, coef) terms, the vectors of
 * symmetric_combineAA, green_class_compute_symmetric_combination.f90:39-137), and all
 * runs advance together as in ed_dynamic_batched from energy = E_psi.  Output
 * dyn[k][2*nw].  Real (N_up, N_dn) and S_z sectors; for S_z sectors sites[t] is the
 * Nambu orbital spin * nc + site of apply_Csz (see ed_apply_c) and `spin` is unused;
 * ED_SITE_NCOND or-ed into an S_z entry takes the correlated excitation of ed_apply_c_n
 * for that term. */
#define ED_SITE_NCOND 0x10000
int ed_dynamic_from_state(ed_sector *gs, const double *psi, double energy, int njobs,
                          ed_sector *const *dst, const int *iph, const int *nterm,
                          const int *sites, const double *coefs, int spin, int nitergreen,
                          double cutoff_dynamic, int precision, int nw, const double *wn,
                          double *dyn);

/* ed_dynamic_from_state for several eigenstates at once (the loop_over_states of
 * compute_greenAA, green_class_computeAA.h:101-302, with every state's runs in one call):
 * state q is psi[q] (host, dimen of gs[q]) with energy[q]; job k runs from state
 * state_of[k] as ed_dynamic_from_state's job k.  All runs advance together: the ones the
 * multi-run persistent kernel takes in as few cooperative launches as fill the GPU, the
 * others in the two-launch batch.  Output dyn[k][2*nw] (unweighted; the caller applies
 * the Boltzmann weights). */
int ed_dynamic_from_states(int nstates, ed_sector *const *gs, const double *const *psi,
                           const double *energy, int njobs, const int *state_of, ed_sector *const *dst,
                           const int *iph, const int *nterm, const int *sites, const double *coefs,
                           int spin, int nitergreen, double cutoff_dynamic, int precision, int nw,
                           const double *wn, double *dyn);

/* Lanczos_GPU_updo (Lanczos_gpu.f90:106) for many sectors at once: niter fixed steps of
 * the recurrence from each sector's create_fix_initial_vector start (rcvector_class.f90:
 * 43-60, as ed_start_vector), all sectors advancing together (the multi-run persistent
 * kernel where the sectors fit it, cut into launches that fill the GPU; the two-launch
 * batch otherwise; a sector of dimension <= niter, or one neither batch kernel takes,
 * runs alone).  alpha[k][niter], beta[k][niter] as ed_lanczos_run (beta[k][0] = ||v_1||);
 * steps[k] = niter, or the step at which sector k's Krylov space closed (beta below the
 * breakdown threshold; the coefficients after it are zero).  FP64 or FP32 vectors. */
int ed_lanczos_batched(int nsec, ed_sector *const *secs, int niter, int precision, double *alpha,
                       double *beta, int *steps);

/* The convergence test of Lanczos_fast_diagonalize (Lanczos_plain.f90:33-134) replayed on
 * recorded coefficients alpha[n], beta[n] (as ed_lanczos_run returns them): the first
 * iteration j <= min(n, nitermax) at which every one of the neigen lowest eigenvalues of T_j
 * moved by less than tolerance since T_{j-1} (j > 2), or min(n, nitermax); *iters = j and
 * evals[neigen] = those eigenvalues of T_j (ascending).  Host only. */
int ed_lanczos_converged(int n, const double *alpha, const double *beta, int neigen, double tolerance,
                         int nitermax, int *iters, double *evals);

/* ----------------------------------------------------------- diagnostics */

/* Launches of the multi-run persistent kernel of ed_dynamic_batched /
 * ed_dynamic_from_state since the library was loaded, launches of it the runtime refused
 * (those batches ran the two-launch path), and for the last launch its duration (a HIP
 * event pair around it on the sector's stream), its job-steps (runs x steps) and its
 * algorithmic HBM bytes (5 * sizeof(T) * dimension per job-step).  No reference
 * counterpart: it lets a caller tell which path ran and measure it.  Any output may be
 * NULL. */
int ed_dynamic_batched_stats(int *persistent_launches, int *refusals, double *last_ms, double *last_job_steps,
                             double *last_bytes);

/* Starts an FP64 (ED_PREC_F64) or FP32 (ED_PREC_F32) Lanczos recurrence from the
 * default start vector and launches the fused H.v of its first step (z = H u_1 and
 * the partial sums of <z, u_1>) `reps` times back to back, after one warm-up launch,
 * between one HIP event pair on the sector's stream.  *avg_ms = elapsed / reps (the
 * kernel's duration plus the dispatch gap); *bytes_per_launch = its algorithmic HBM
 * bytes 2 * sizeof(T) * dimen_H() (read u_j, write H u_j).  No reference
 * counterpart: it is the measurement hook of the engine. */
int ed_hmult_benchmark(ed_sector *sec, int precision, int reps, double *avg_ms, double *bytes_per_launch);

#ifdef __cplusplus
}
#endif

#endif /* ED_HIP_H */
