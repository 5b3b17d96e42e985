! ISO_C_BINDING interface of libed_hip.so (include/ed_hip.h) for Fortran hosts.
!
! Only plain C types cross the boundary: the model is a bind(C) struct of scalars
! and pointers, vectors are assumed-size REAL(8) arrays passed by reference, the
! sector is an opaque type(c_ptr).  Every function returns an integer status
! (0 = ED_OK); ed_check turns a non-zero status into an error stop with the
! library's message, which is how the reference reports the same conditions (STOP).
module ed_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: ED_OK = 0, ED_PREC_F64 = 0, ED_PREC_F32 = 1, ED_PREC_MIXED_RQ = 2
  integer(c_int), parameter :: ED_ERR_ARG = -1, ED_ERR_HIP = -2, ED_ERR_NOMEM = -3, ED_ERR_BREAKDOWN = -4, &
                               ED_ERR_STATE = -5, ED_ERR_CF = -6

  ! ed_lanczos_opts (Lanczos_fast_diagonalize parameters, dmft_ed_solver_variables.h:9-20)
  type, bind(C) :: ed_lanczos_opts_t
    integer(c_int) :: nitermax, neigen
    real(c_double) :: tolerance
    integer(c_int) :: precision, check_every, reorth
  end type

  type, bind(C) :: ed_model
    integer(c_int32_t) :: nc, nb
    type(c_ptr)        :: ec, eb, vbc, u, jj
    real(c_double)     :: jhund, cutoff
  end type

  ! ed_arpack_opts (arpack_eigenvector_sym_matrix parameters, ED_ARPACK.f90:35-250)
  integer(c_int), parameter :: ED_ARPACK_BE = 0, ED_ARPACK_SA = 1, ED_ARPACK_LA = 2
  type, bind(C) :: ed_arpack_opts_t
    integer(c_int) :: nev, ncv, which
    real(c_double) :: tol
    integer(c_int) :: maxitr
  end type

  ! ed_block_opts (Block_Lanczos_diagonalize parameters)
  type, bind(C) :: ed_block_opts_t
    integer(c_int) :: block_size, nev, max_steps, check_every
    real(c_double) :: tolerance
    integer(c_int) :: reorth
  end type

  interface
    function ed_last_error() bind(C, name="ed_last_error")
      import :: c_ptr
      type(c_ptr) :: ed_last_error
    end function
    integer(c_int) function ed_init(device) bind(C, name="ed_init")
      import :: c_int
      integer(c_int), value :: device
    end function
    integer(c_int) function ed_sector_create(model, nup, ndn, sec) bind(C, name="ed_sector_create")
      import :: c_int, c_ptr, ed_model
      type(ed_model), intent(in) :: model
      integer(c_int), value      :: nup, ndn
      type(c_ptr), intent(out)   :: sec
    end function
    ! complex AIM (COMPLEX(8) branch of new_H / HAIMupdo_multc): imaginary parts of Ec, Eb,
    ! Vbc in the layout of the model's arrays; eb_im / vbc_im may be c_null_ptr (= 0)
    integer(c_int) function ed_sector_create_c(model, ec_im, eb_im, vbc_im, nup, ndn, sec) &
        bind(C, name="ed_sector_create_c")
      import :: c_int, c_ptr, ed_model
      type(ed_model), intent(in) :: model
      type(c_ptr), value         :: ec_im, eb_im, vbc_im
      integer(c_int), value      :: nup, ndn
      type(c_ptr), intent(out)   :: sec
    end function
    integer(c_int) function ed_sector_create_slater(model, u4, nup, ndn, sec) &
        bind(C, name="ed_sector_create_slater")
      import :: c_int, c_ptr, ed_model
      type(ed_model), intent(in) :: model
      type(c_ptr), value         :: u4
      integer(c_int), value      :: nup, ndn
      type(c_ptr), intent(out)   :: sec
    end function
    integer(c_int) function ed_hmult_c128(sec, n, vec_out, vec_in) bind(C, name="ed_hmult_c128")
      import :: c_int, c_int64_t, c_ptr, c_double_complex
      type(c_ptr), value                   :: sec
      integer(c_int64_t), value            :: n
      complex(c_double_complex)            :: vec_out(*)
      complex(c_double_complex), intent(in):: vec_in(*)
    end function
    ! S_z (Nambu) sector of a superconducting bath (new_HAIMsz); pb / pvbc may be c_null_ptr
    integer(c_int) function ed_sector_create_sz(model, pb, pvbc, pairing_imp_to_bath, npart, sec) &
        bind(C, name="ed_sector_create_sz")
      import :: c_int, c_ptr, ed_model
      type(ed_model), intent(in) :: model
      type(c_ptr), value         :: pb, pvbc
      integer(c_int), value      :: pairing_imp_to_bath, npart
      type(c_ptr), intent(out)   :: sec
    end function
    ! complex S_z sector (COMPLEX(8) new_HAIMsz); every imaginary part but ec_im may be c_null_ptr
    integer(c_int) function ed_sector_create_sz_c(model, ec_im, eb_im, vbc_im, pb, pb_im, pvbc, pvbc_im, &
        pairing_imp_to_bath, npart, sec) bind(C, name="ed_sector_create_sz_c")
      import :: c_int, c_ptr, ed_model
      type(ed_model), intent(in) :: model
      type(c_ptr), value         :: ec_im, eb_im, vbc_im, pb, pb_im, pvbc, pvbc_im
      integer(c_int), value      :: pairing_imp_to_bath, npart
      type(c_ptr), intent(out)   :: sec
    end function
    integer(c_int) function ed_sector_energy_shift(sec, shift) bind(C, name="ed_sector_energy_shift")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value          :: sec
      real(c_double), intent(out) :: shift
    end function
    integer(c_int) function ed_sector_destroy(sec) bind(C, name="ed_sector_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
    end function
    integer(c_int64_t) function ed_sector_dim(sec) bind(C, name="ed_sector_dim")
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: sec
    end function
    integer(c_int) function ed_hmult_f64(sec, n, vec_out, vec_in) bind(C, name="ed_hmult_f64")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value        :: sec
      integer(c_int64_t), value :: n
      real(c_double)            :: vec_out(*)
      real(c_double), intent(in):: vec_in(*)
    end function
    integer(c_int) function ed_lanczos_run(sec, niter, start, precision, diag, subdiag) bind(C, name="ed_lanczos_run")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value     :: sec
      integer(c_int), value  :: niter, precision
      type(c_ptr), value     :: start
      real(c_double)         :: diag(*), subdiag(*)
    end function
    integer(c_int) function ed_lanczos_gs_vector(sec, iters, vecp, psi) bind(C, name="ed_lanczos_gs_vector")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value     :: sec
      integer(c_int), value  :: iters
      real(c_double)         :: vecp(*), psi(*)
    end function
    integer(c_int) function ed_tridiag_eig(n, diag, subdiag, w, z) bind(C, name="ed_tridiag_eig")
      import :: c_int, c_double
      integer(c_int), value :: n
      real(c_double)        :: diag(*), subdiag(*), w(*), z(*)
    end function
    integer(c_int) function ed_start_vector(n, out) bind(C, name="ed_start_vector")
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: n
      real(c_double)            :: out(*)
    end function
    integer(c_int) function ed_continued_fraction(n, a, b, nw, wn, out) bind(C, name="ed_continued_fraction")
      import :: c_int, c_double
      integer(c_int), value :: n, nw
      real(c_double)        :: a(*), b(*), wn(*), out(*)
    end function
    integer(c_int) function ed_block_default_opts(opts) bind(C, name="ed_block_default_opts")
      import :: c_int, ed_block_opts_t
      type(ed_block_opts_t), intent(out) :: opts
    end function
    integer(c_int) function ed_block_lanczos(sec, opts, evals, resid, evecs, ldv, nconv, steps) &
        bind(C, name="ed_block_lanczos")
      import :: c_int, c_int64_t, c_ptr, c_double, ed_block_opts_t
      type(c_ptr), value                :: sec
      type(ed_block_opts_t), intent(in) :: opts
      real(c_double)                    :: evals(*), resid(*), evecs(*)
      integer(c_int64_t), value         :: ldv
      integer(c_int), intent(out)       :: nconv, steps
    end function
    ! Same, keeping the eigenpairs within dEmax of the lowest (Block_Lanczos.f90:289-300): only
    ! their vectors (the leading nkept columns of evecs) are copied
    integer(c_int) function ed_block_lanczos_window(sec, opts, demax, evals, resid, evecs, ldv, nconv, steps, &
        nkept) bind(C, name="ed_block_lanczos_window")
      import :: c_int, c_int64_t, c_ptr, c_double, ed_block_opts_t
      type(c_ptr), value                :: sec
      type(ed_block_opts_t), intent(in) :: opts
      real(c_double), value             :: demax
      real(c_double)                    :: evals(*), resid(*), evecs(*)
      integer(c_int64_t), value         :: ldv
      integer(c_int), intent(out)       :: nconv, steps, nkept
    end function
    ! Lanczos_GPU_updo for many sectors at once; secs(k) = sector handles
    integer(c_int) function ed_lanczos_batched(nsec, secs, niter, precision, alpha, beta, steps) &
        bind(C, name="ed_lanczos_batched")
      import :: c_int, c_ptr, c_double
      integer(c_int), value :: nsec, niter, precision
      type(c_ptr)           :: secs(*)
      real(c_double)        :: alpha(*), beta(*)
      integer(c_int)        :: steps(*)
    end function
    ! Lanczos_fast_diagonalize's convergence test on recorded coefficients
    integer(c_int) function ed_lanczos_converged(n, alpha, beta, neigen, tolerance, nitermax, iters, evals) &
        bind(C, name="ed_lanczos_converged")
      import :: c_int, c_double
      integer(c_int), value       :: n, neigen, nitermax
      real(c_double)              :: alpha(*), beta(*), evals(*)
      real(c_double), value       :: tolerance
      integer(c_int), intent(out) :: iters
    end function
    integer(c_int) function ed_block_last_stats(steps, full_passes, ms, ritz_ms) bind(C, name="ed_block_last_stats")
      import :: c_int, c_double
      integer(c_int), intent(out) :: steps, full_passes
      real(c_double), intent(out) :: ms, ritz_ms
    end function
    integer(c_int) function ed_arpack_default_opts(opts) bind(C, name="ed_arpack_default_opts")
      import :: c_int, ed_arpack_opts_t
      type(ed_arpack_opts_t), intent(out) :: opts
    end function
    ! ARPACK_diago on the device; start may be c_null_ptr (create_fix_initial_vector's vector)
    integer(c_int) function ed_arpack(sec, opts, start, evals, resid, evecs, ldv, nconv, iters, matvecs) &
        bind(C, name="ed_arpack")
      import :: c_int, c_int64_t, c_ptr, c_double, ed_arpack_opts_t
      type(c_ptr), value                 :: sec, start
      type(ed_arpack_opts_t), intent(in) :: opts
      real(c_double)                     :: evals(*), resid(*), evecs(*)
      integer(c_int64_t), value          :: ldv
      integer(c_int), intent(out)        :: nconv, iters, matvecs
    end function
    integer(c_int) function ed_hf_gradient(sec, n, psi, d_level, d_v, d_u) bind(C, name="ed_hf_gradient")
      import :: c_int, c_int64_t, c_ptr, c_double
      type(c_ptr), value          :: sec
      integer(c_int64_t), value   :: n
      real(c_double), intent(in)  :: psi(*)
      real(c_double)              :: d_level(*), d_v(*)
      real(c_double), intent(out) :: d_u
    end function
    integer(c_int) function ed_bath_distance(nc, nb, eps, vbc, nw, wn, target, weight_expo, fit_weight_power, &
        fit_shift, dist, grad_eps, grad_v) bind(C, name="ed_bath_distance")
      import :: c_int, c_double
      integer(c_int), value       :: nc, nb, nw
      real(c_double), intent(in)  :: eps(*), vbc(*), wn(*), target(*)
      real(c_double), value       :: weight_expo, fit_weight_power, fit_shift
      real(c_double), intent(out) :: dist
      real(c_double)              :: grad_eps(*), grad_v(*)
    end function
    ! ---- the remaining entry points of ed_hip.h, one interface each (device pointers and
    ! the optional start / prev vectors as type(c_ptr), value; c_null_ptr for NULL)
    integer(c_int) function ed_sector_set_fly(sec, on) bind(C, name="ed_sector_set_fly")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: on
    end function
    integer(c_int) function ed_sector_hund_sz(sec, kh, rk, val) bind(C, name="ed_sector_hund_sz")
      import :: c_double, c_int, c_int32_t, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), intent(out) :: kh
      integer(c_int32_t) :: rk(*)
      real(c_double) :: val(*)
    end function
    integer(c_int) function ed_sector_quart_sz(sec, quart) bind(C, name="ed_sector_quart_sz")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: sec
      real(c_double) :: quart(*)
    end function
    function ed_version() bind(C, name="ed_version")
      import :: c_ptr
      type(c_ptr) :: ed_version
    end function
    integer(c_int) function ed_sector_spin_dims(sec, d_up, d_dn) bind(C, name="ed_sector_spin_dims")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: sec
      integer(c_int32_t) :: d_up(*)
      integer(c_int32_t) :: d_dn(*)
    end function
    integer(c_int) function ed_sector_nnz(sec, spin, nnz) bind(C, name="ed_sector_nnz")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: spin
      integer(c_int64_t), intent(out) :: nnz
    end function
    integer(c_int) function ed_sector_tables(sec, spin, states, diag, noff, rankoff, offdiag) bind(C, name="ed_sector_tables")
      import :: c_double, c_int, c_int32_t, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: spin
      integer(c_int32_t) :: states(*)
      real(c_double) :: diag(*)
      integer(c_int32_t) :: noff(*)
      integer(c_int32_t) :: rankoff(*)
      real(c_double) :: offdiag(*)
    end function
    integer(c_int) function ed_sector_tables_im(sec, spin, offdiag_im) bind(C, name="ed_sector_tables_im")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: spin
      real(c_double) :: offdiag_im(*)
    end function
    integer(c_int) function ed_hmult_f64_dev(sec, n, d_out, d_in, stream) bind(C, name="ed_hmult_f64_dev")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: sec
      integer(c_int64_t), value :: n
      type(c_ptr), value :: d_out
      type(c_ptr), value :: d_in
      type(c_ptr), value :: stream
    end function
    integer(c_int) function ed_hmult_f32_dev(sec, n, d_out, d_in, stream) bind(C, name="ed_hmult_f32_dev")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: sec
      integer(c_int64_t), value :: n
      type(c_ptr), value :: d_out
      type(c_ptr), value :: d_in
      type(c_ptr), value :: stream
    end function
    integer(c_int) function ed_hmult_c128_dev(sec, n, d_out, d_in, stream) bind(C, name="ed_hmult_c128_dev")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: sec
      integer(c_int64_t), value :: n
      type(c_ptr), value :: d_out
      type(c_ptr), value :: d_in
      type(c_ptr), value :: stream
    end function
    integer(c_int) function ed_sector_synchronize(sec) bind(C, name="ed_sector_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
    end function
    integer(c_int) function ed_hmult_block_f64_dev(sec, n, nvec, ld, d_out, d_in, stream) bind(C, name="ed_hmult_block_f64_dev")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value :: sec
      integer(c_int64_t), value :: n
      integer(c_int), value :: nvec
      integer(c_int64_t), value :: ld
      type(c_ptr), value :: d_out
      type(c_ptr), value :: d_in
      type(c_ptr), value :: stream
    end function
    integer(c_int) function ed_hmult_tune(sec, choice, ms_per_choice) bind(C, name="ed_hmult_tune")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), intent(out) :: choice
      real(c_double) :: ms_per_choice(*)
    end function
    integer(c_int) function ed_hmult_set_schedule(sec, choice) bind(C, name="ed_hmult_set_schedule")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: choice
    end function
    integer(c_int) function ed_lanczos_benchmark(sec, precision, niter, reps, avg_ms, bytes_per_run, persistent) &
        bind(C, name="ed_lanczos_benchmark")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: precision
      integer(c_int), value :: niter
      integer(c_int), value :: reps
      real(c_double), intent(out) :: avg_ms
      real(c_double), intent(out) :: bytes_per_run
      integer(c_int), intent(out) :: persistent
    end function
    integer(c_int) function ed_lanczos_set_persistent(sec, enable) bind(C, name="ed_lanczos_set_persistent")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: enable
    end function
    integer(c_int) function ed_lanczos_persistent_active(sec, precision, active) bind(C, name="ed_lanczos_persistent_active")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: precision
      integer(c_int), intent(out) :: active
    end function
    integer(c_int) function ed_lanczos_default_opts(opts) bind(C, name="ed_lanczos_default_opts")
      import :: c_int, ed_lanczos_opts_t
      type(ed_lanczos_opts_t) :: opts
    end function
    integer(c_int) function ed_lanczos_diagonalize(sec, opts, start, iters, neig_out, evals, vecp, diag, subdiag) &
        bind(C, name="ed_lanczos_diagonalize")
      import :: c_double, c_int, c_ptr, ed_lanczos_opts_t
      type(c_ptr), value :: sec
      type(ed_lanczos_opts_t) :: opts
      type(c_ptr), value :: start
      integer(c_int), intent(out) :: iters
      integer(c_int), intent(out) :: neig_out
      real(c_double) :: evals(*)
      real(c_double) :: vecp(*)
      real(c_double) :: diag(*)
      real(c_double) :: subdiag(*)
    end function
    integer(c_int) function ed_lanczos_info(sec, iters_f32_phase) bind(C, name="ed_lanczos_info")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), intent(out) :: iters_f32_phase
    end function
    integer(c_int) function ed_lanczos_rq_info(sec, accepted, res2, gap) bind(C, name="ed_lanczos_rq_info")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), intent(out) :: accepted
      real(c_double), intent(out) :: res2
      real(c_double), intent(out) :: gap
    end function
    integer(c_int) function ed_lanczos_steps_run(sec, steps) bind(C, name="ed_lanczos_steps_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), intent(out) :: steps
    end function
    integer(c_int) function ed_full_ed(sec, neig, evals, evecs) bind(C, name="ed_full_ed")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: neig
      real(c_double) :: evals(*)
      real(c_double) :: evecs(*)
    end function
    integer(c_int) function ed_lanczos_gs_vector_dev(sec, iters, vecp, d_psi) bind(C, name="ed_lanczos_gs_vector_dev")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: iters
      real(c_double) :: vecp(*)
      type(c_ptr), value :: d_psi
    end function
    integer(c_int) function ed_rayleigh_f64(sec, n, psi, energy) bind(C, name="ed_rayleigh_f64")
      import :: c_double, c_int, c_int64_t, c_ptr
      type(c_ptr), value :: sec
      integer(c_int64_t), value :: n
      real(c_double) :: psi(*)
      real(c_double), intent(out) :: energy
    end function
    integer(c_int) function ed_tridiag_lowest(n, diag, subdiag, k, prev, w) bind(C, name="ed_tridiag_lowest")
      import :: c_double, c_int, c_ptr
      integer(c_int), value :: n
      real(c_double) :: diag(*)
      real(c_double) :: subdiag(*)
      integer(c_int), value :: k
      type(c_ptr), value :: prev
      real(c_double) :: w(*)
    end function
    integer(c_int) function ed_band_eig(n, b, a, nev, w, z) bind(C, name="ed_band_eig")
      import :: c_double, c_int
      integer(c_int), value :: n
      integer(c_int), value :: b
      real(c_double) :: a(*)
      integer(c_int), value :: nev
      real(c_double) :: w(*)
      real(c_double) :: z(*)
    end function
    integer(c_int) function ed_block_lanczos_dev(sec, opts, evals, resid, d_evecs, ldv, nconv, steps) &
        bind(C, name="ed_block_lanczos_dev")
      import :: c_double, c_int, c_int64_t, c_ptr, ed_block_opts_t
      type(c_ptr), value :: sec
      type(ed_block_opts_t), intent(in) :: opts
      real(c_double) :: evals(*)
      real(c_double) :: resid(*)
      type(c_ptr), value :: d_evecs
      integer(c_int64_t), value :: ldv
      integer(c_int), intent(out) :: nconv
      integer(c_int), intent(out) :: steps
    end function
    integer(c_int) function ed_gram_f64_dev(n, p, q, d_x, ldx, d_y, ldy, g, stream) bind(C, name="ed_gram_f64_dev")
      import :: c_double, c_int, c_int64_t, c_ptr
      integer(c_int64_t), value :: n
      integer(c_int), value :: p
      integer(c_int), value :: q
      type(c_ptr), value :: d_x
      integer(c_int64_t), value :: ldx
      type(c_ptr), value :: d_y
      integer(c_int64_t), value :: ldy
      real(c_double) :: g(*)
      type(c_ptr), value :: stream
    end function
    integer(c_int) function ed_xc_f64_dev(n, p, q, d_x, ldx, c, alpha, beta, d_y, ldy, stream) bind(C, name="ed_xc_f64_dev")
      import :: c_double, c_int, c_int64_t, c_ptr
      integer(c_int64_t), value :: n
      integer(c_int), value :: p
      integer(c_int), value :: q
      type(c_ptr), value :: d_x
      integer(c_int64_t), value :: ldx
      real(c_double) :: c(*)
      real(c_double), value :: alpha
      real(c_double), value :: beta
      type(c_ptr), value :: d_y
      integer(c_int64_t), value :: ldy
      type(c_ptr), value :: stream
    end function
    integer(c_int) function ed_hf_gradient_dev(sec, d_psi, d_level, d_v, d_u) bind(C, name="ed_hf_gradient_dev")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: sec
      type(c_ptr), value :: d_psi
      real(c_double) :: d_level(*)
      real(c_double) :: d_v(*)
      real(c_double), intent(out) :: d_u
    end function
    integer(c_int) function ed_apply_c(src, dst, site, spin, pm, psi, phi) bind(C, name="ed_apply_c")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: src
      type(c_ptr), value :: dst
      integer(c_int), value :: site
      integer(c_int), value :: spin
      integer(c_int), value :: pm
      real(c_double) :: psi(*)
      real(c_double) :: phi(*)
    end function
    ! apply_Csz_N_updo (apply_C.f90:404-483): the correlated excitation on S_z sectors
    integer(c_int) function ed_apply_c_n(src, dst, site, spin, pm, psi, phi) bind(C, name="ed_apply_c_n")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: src
      type(c_ptr), value :: dst
      integer(c_int), value :: site
      integer(c_int), value :: spin
      integer(c_int), value :: pm
      real(c_double) :: psi(*)
      real(c_double) :: phi(*)
    end function
    integer(c_int) function ed_dynamic(sec, iph, energy, phi, nitergreen, cutoff_dynamic, precision, nw, wn, dyn, a, b) &
        bind(C, name="ed_dynamic")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: iph
      real(c_double), value :: energy
      real(c_double) :: phi(*)
      integer(c_int), value :: nitergreen
      real(c_double), value :: cutoff_dynamic
      integer(c_int), value :: precision
      integer(c_int), value :: nw
      real(c_double) :: wn(*)
      real(c_double) :: dyn(*)
      real(c_double) :: a(*)
      real(c_double) :: b(*)
    end function
    integer(c_int) function ed_dynamic_full_ed(dim, iph, energy, phi, nst, evals, evecs, cutoff_dynamic, nw, wn, dyn) &
        bind(C, name="ed_dynamic_full_ed")
      import :: c_double, c_int
      integer(c_int), value :: dim
      integer(c_int), value :: iph
      real(c_double), value :: energy
      real(c_double) :: phi(*)
      integer(c_int), value :: nst
      real(c_double) :: evals(*)
      real(c_double) :: evecs(*)
      real(c_double), value :: cutoff_dynamic
      integer(c_int), value :: nw
      real(c_double) :: wn(*)
      real(c_double) :: dyn(*)
    end function
    integer(c_int) function ed_dynamic_full_ed_c(dim, iph, energy, phi, nst, evals, evecs, cutoff_dynamic, nw, wn, dyn) &
        bind(C, name="ed_dynamic_full_ed_c")
      import :: c_double, c_int
      integer(c_int), value :: dim
      integer(c_int), value :: iph
      real(c_double), value :: energy
      real(c_double) :: phi(*)
      integer(c_int), value :: nst
      real(c_double) :: evals(*)
      real(c_double) :: evecs(*)
      real(c_double), value :: cutoff_dynamic
      integer(c_int), value :: nw
      real(c_double) :: wn(*)
      real(c_double) :: dyn(*)
    end function
    integer(c_int) function ed_dynamic_batched(njobs, secs, phis, iph, energy, nitergreen, cutoff_dynamic, precision, nw, wn, dyn, a, b) &
        bind(C, name="ed_dynamic_batched")
      import :: c_double, c_int, c_ptr
      integer(c_int), value :: njobs
      type(c_ptr) :: secs(*)
      type(c_ptr) :: phis(*)
      integer(c_int) :: iph(*)
      real(c_double), intent(out) :: energy
      integer(c_int), value :: nitergreen
      real(c_double), value :: cutoff_dynamic
      integer(c_int), value :: precision
      integer(c_int), value :: nw
      real(c_double) :: wn(*)
      real(c_double) :: dyn(*)
      real(c_double) :: a(*)
      real(c_double) :: b(*)
    end function
    integer(c_int) function ed_dynamic_from_state(gs, psi, energy, njobs, dst, iph, nterm, sites, coefs, spin, nitergreen, cutoff_dynamic, precision, nw, wn, dyn) &
        bind(C, name="ed_dynamic_from_state")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: gs
      real(c_double) :: psi(*)
      real(c_double), value :: energy
      integer(c_int), value :: njobs
      type(c_ptr) :: dst(*)
      integer(c_int) :: iph(*)
      integer(c_int) :: nterm(*)
      integer(c_int) :: sites(*)
      real(c_double) :: coefs(*)
      integer(c_int), value :: spin
      integer(c_int), value :: nitergreen
      real(c_double), value :: cutoff_dynamic
      integer(c_int), value :: precision
      integer(c_int), value :: nw
      real(c_double) :: wn(*)
      real(c_double) :: dyn(*)
    end function
    ! all excitation runs of several states in one call; psi(q) = c_loc of state q's vector
    integer(c_int) function ed_dynamic_from_states(nstates, gs, psi, energy, njobs, state_of, dst, iph, nterm, &
        sites, coefs, spin, nitergreen, cutoff_dynamic, precision, nw, wn, dyn) bind(C, name="ed_dynamic_from_states")
      import :: c_double, c_int, c_ptr
      integer(c_int), value :: nstates
      type(c_ptr) :: gs(*)
      type(c_ptr) :: psi(*)
      real(c_double) :: energy(*)
      integer(c_int), value :: njobs
      integer(c_int) :: state_of(*)
      type(c_ptr) :: dst(*)
      integer(c_int) :: iph(*)
      integer(c_int) :: nterm(*)
      integer(c_int) :: sites(*)
      real(c_double) :: coefs(*)
      integer(c_int), value :: spin
      integer(c_int), value :: nitergreen
      real(c_double), value :: cutoff_dynamic
      integer(c_int), value :: precision
      integer(c_int), value :: nw
      real(c_double) :: wn(*)
      real(c_double) :: dyn(*)
    end function
    integer(c_int) function ed_hmult_benchmark(sec, precision, reps, avg_ms, bytes_per_launch) bind(C, name="ed_hmult_benchmark")
      import :: c_double, c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: precision
      integer(c_int), value :: reps
      real(c_double), intent(out) :: avg_ms
      real(c_double), intent(out) :: bytes_per_launch
    end function

    integer(c_int) function ed_lanczos_persistent_shape(sec, precision, cols_per_wg, workgroups, rows, width) &
        bind(C, name="ed_lanczos_persistent_shape")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), value :: precision
      integer(c_int), intent(out) :: cols_per_wg, workgroups, rows, width
    end function
    integer(c_int) function ed_lanczos_persistent_refusals(sec, refused) &
        bind(C, name="ed_lanczos_persistent_refusals")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), intent(out) :: refused
    end function
    integer(c_int) function ed_lanczos_exact_beta_steps(sec, steps) bind(C, name="ed_lanczos_exact_beta_steps")
      import :: c_int, c_ptr
      type(c_ptr), value :: sec
      integer(c_int), intent(out) :: steps
    end function
    integer(c_int) function ed_dynamic_batched_stats(persistent_launches, refusals, last_ms, last_job_steps, &
        last_bytes) bind(C, name="ed_dynamic_batched_stats")
      import :: c_int, c_double
      integer(c_int), intent(out) :: persistent_launches, refusals
      real(c_double), intent(out) :: last_ms, last_job_steps, last_bytes
    end function
    ! ed_dynamic without device allocations (static scratch; limits ED_NOALLOC_MAX_* of ed_hip.h)
    integer(c_int) function ed_dynamic_noalloc(sec, iph, energy, phi, nitergreen, cutoff_dynamic, precision, nw, wn, &
        dyn, a, b) bind(C, name="ed_dynamic_noalloc")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value          :: sec
      integer(c_int), value       :: iph, nitergreen, precision, nw
      real(c_double), value       :: energy, cutoff_dynamic
      real(c_double), intent(in)  :: phi(*), wn(*)
      real(c_double)              :: dyn(*), a(*), b(*)
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  subroutine ed_check(rc, where)
    integer(c_int), intent(in)   :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n
    if (rc == ED_OK) return
    p = ed_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    write(*, '(a,a,i0,a,*(a))') where, ': ed_hip status ', rc, ': ', msg(1:n)
    error stop 'ed_hip call failed, critical'
  end subroutine

end module ed_hip
