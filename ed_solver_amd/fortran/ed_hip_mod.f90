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

  integer(c_int), parameter :: ED_OK = 0, ED_PREC_F64 = 0, ED_PREC_F32 = 1

  type, bind(C) :: ed_model
    integer(c_int32_t) :: nc, nb
    type(c_ptr)        :: ec, eb, vbc, u, jj
    real(c_double)     :: jhund, cutoff
  end type

  ! ed_block_opts (Block_Lanczos_diagonalize parameters)
  type, bind(C) :: ed_block_opts_t
    integer(c_int) :: block_size, nev, max_steps, check_every
    real(c_double) :: tolerance
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
