! H_class surface over the MI355X engine.
!
! Keeps the interface the reference's solvers call (H_class.f90:13-152):
!   new_H / delete_H / Hmult__ (= Hmultr) / dimen_H
! and the coarse GPU entry points the reference sketches for which_lanczos='GPU'
! (Lanczos_gpu.f90:48,106; green_class_compute_dynamic.f90:468):
!   Lanczos_GPU_updo / Lanczos_GPU_get_GS_updo / Lanczos_dynamic_GPU_updo
! so Lanczos_plain, ED_ARPACK, Block_Lanczos and green_class_compute_dynamic keep
! working unchanged on top of it.
!
! new_H takes the model arrays the reference reads from AIM_type
! (quantum_algebra_Anderson_Impurity_class2.f90:62-135: Ec, Eb, Vbc per spin; the
! impurity U and JJmatrix; Jhund; cutoff_hamilt_param) as plain arrays, in C row-major
! order, so the caller's new_H(AIM, sector) wrapper only gathers them.
module H_class_hip
  use iso_c_binding
  use ed_hip
  implicit none
  private

  type(c_ptr), save :: sector_h = c_null_ptr        ! H_class.f90:11
  real(c_double), allocatable, target, save :: m_ec(:), m_eb(:), m_vbc(:), m_u(:), m_jj(:)
  real(c_double), allocatable, target, save :: m_ec_im(:), m_eb_im(:), m_vbc_im(:)
  real(c_double), allocatable, target, save :: m_pb(:), m_pvbc(:), m_u4(:)

  interface Hmult__                                   ! H_class.f90:13-15
    module procedure Hmultr, Hmultc
  end interface

  public :: new_H, new_H_complex, new_H_sz, new_H_slater, delete_H, Hmult__, Hmultr, Hmultc, dimen_H, &
            sector_handle, energy_shift_H
  public :: Lanczos_GPU_updo, Lanczos_GPU_get_GS_updo, Lanczos_dynamic_GPU_updo, ARPACK_diago_hip

contains

  subroutine new_H(nc, nb, ec, eb, vbc, u, jj, jhund, cutoff, nup, ndn)
    integer, intent(in)        :: nc, nb, nup, ndn
    real(8), intent(in)        :: ec(2*nc*nc), eb(max(1, 2*nb*nb)), vbc(max(1, 2*nb*nc)), u(nc*nc), jj(nc*nc)
    real(8), intent(in)        :: jhund, cutoff
    type(ed_model) :: m
    call delete_H()
    m_ec = ec; m_eb = eb; m_vbc = vbc; m_u = u; m_jj = jj
    m%nc = nc; m%nb = nb
    m%ec = c_loc(m_ec); m%eb = c_loc(m_eb); m%vbc = c_loc(m_vbc); m%u = c_loc(m_u); m%jj = c_loc(m_jj)
    m%jhund = jhund; m%cutoff = cutoff
    call ed_check(ed_sector_create(m, int(nup, c_int), int(ndn, c_int), sector_h), 'new_H')
  end subroutine

  ! new_H for a complex AIM: the imaginary parts of Ec, Eb, Vbc beside the real ones;
  ! the sector's vectors are then COMPLEX(8) (Hmultc).
  subroutine new_H_complex(nc, nb, ec, eb, vbc, ec_im, eb_im, vbc_im, u, jj, jhund, cutoff, nup, ndn)
    integer, intent(in)        :: nc, nb, nup, ndn
    real(8), intent(in)        :: ec(2*nc*nc), eb(max(1, 2*nb*nb)), vbc(max(1, 2*nb*nc)), u(nc*nc), jj(nc*nc)
    real(8), intent(in)        :: ec_im(2*nc*nc), eb_im(max(1, 2*nb*nb)), vbc_im(max(1, 2*nb*nc))
    real(8), intent(in)        :: jhund, cutoff
    type(ed_model) :: m
    call delete_H()
    m_ec = ec; m_eb = eb; m_vbc = vbc; m_u = u; m_jj = jj
    m_ec_im = ec_im; m_eb_im = eb_im; m_vbc_im = vbc_im
    m%nc = nc; m%nb = nb
    m%ec = c_loc(m_ec); m%eb = c_loc(m_eb); m%vbc = c_loc(m_vbc); m%u = c_loc(m_u); m%jj = c_loc(m_jj)
    m%jhund = jhund; m%cutoff = cutoff
    call ed_check(ed_sector_create_c(m, c_loc(m_ec_im), c_loc(m_eb_im), c_loc(m_vbc_im), int(nup, c_int), &
                                     int(ndn, c_int), sector_h), 'new_H_complex')
  end subroutine

  ! The S_z branch of new_H (H_class.f90:28-37 -> new_HAIMsz, HAIMsz.f90:51-177): the Nambu
  ! sector with npart = N_up + (Ns - N_dn) set bits, bath pairing pb(nb*nb) (Pb) and
  ! impurity-bath pairing pvbc(2*nb*nc) (PVbc, used when pairing_imp_to_bath /= 0), both in
  ! C row-major order.  Hmultr then applies HAIMsz_mult; the reference's reported energies
  ! add energy_shift_H() to the eigenvalues of this H (eigen_sector_class.f90:329).
  subroutine new_H_sz(nc, nb, ec, eb, vbc, u, jj, jhund, cutoff, pb, pvbc, pairing_imp_to_bath, npart)
    integer, intent(in)        :: nc, nb, npart, pairing_imp_to_bath
    real(8), intent(in)        :: ec(2*nc*nc), eb(max(1, 2*nb*nb)), vbc(max(1, 2*nb*nc)), u(nc*nc), jj(nc*nc)
    real(8), intent(in)        :: pb(max(1, nb*nb)), pvbc(max(1, 2*nb*nc))
    real(8), intent(in)        :: jhund, cutoff
    type(ed_model) :: m
    call delete_H()
    m_ec = ec; m_eb = eb; m_vbc = vbc; m_u = u; m_jj = jj; m_pb = pb; m_pvbc = pvbc
    m%nc = nc; m%nb = nb
    m%ec = c_loc(m_ec); m%eb = c_loc(m_eb); m%vbc = c_loc(m_vbc); m%u = c_loc(m_u); m%jj = c_loc(m_jj)
    m%jhund = jhund; m%cutoff = cutoff
    call ed_check(ed_sector_create_sz(m, c_loc(m_pb), c_loc(m_pvbc), int(pairing_imp_to_bath, c_int), &
                                      int(npart, c_int), sector_h), 'new_H_sz')
  end subroutine

  ! new_H with the Slater 4-index impurity interaction of flag_slater_int (HAIMupdo.f90:905-999):
  ! u4(nc**4) = Slater_Coulomb_r(site1, site2, site3, site4) in C order (site4 fastest).
  subroutine new_H_slater(nc, nb, ec, eb, vbc, u, jj, jhund, cutoff, u4, nup, ndn)
    integer, intent(in)        :: nc, nb, nup, ndn
    real(8), intent(in)        :: ec(2*nc*nc), eb(max(1, 2*nb*nb)), vbc(max(1, 2*nb*nc)), u(nc*nc), jj(nc*nc)
    real(8), intent(in)        :: u4(nc*nc*nc*nc)
    real(8), intent(in)        :: jhund, cutoff
    type(ed_model) :: m
    call delete_H()
    m_ec = ec; m_eb = eb; m_vbc = vbc; m_u = u; m_jj = jj; m_u4 = u4
    m%nc = nc; m%nb = nb
    m%ec = c_loc(m_ec); m%eb = c_loc(m_eb); m%vbc = c_loc(m_vbc); m%u = c_loc(m_u); m%jj = c_loc(m_jj)
    m%jhund = jhund; m%cutoff = cutoff
    call ed_check(ed_sector_create_slater(m, c_loc(m_u4), int(nup, c_int), int(ndn, c_int), sector_h), &
                  'new_H_slater')
  end subroutine

  ! The constant the S_z basis moves out of H (tr Ec_dn + tr Eb_dn); 0 for other sectors.
  real(8) function energy_shift_H()
    real(c_double) :: sh
    if (.not. c_associated(sector_h)) stop 'energy_shift_H: no sector'
    call ed_check(ed_sector_energy_shift(sector_h, sh), 'energy_shift_H')
    energy_shift_H = sh
  end function

  subroutine delete_H()
    integer(c_int) :: rc
    if (c_associated(sector_h)) then           ! no-op if H was never created (H_class.f90:48-58)
      rc = ed_sector_destroy(sector_h)
      sector_h = c_null_ptr
    end if
  end subroutine

  function sector_handle() result(p)
    type(c_ptr) :: p
    p = sector_h
  end function

  subroutine Hmultr(n, vec_out, vec_in)
    integer                :: n
    real(8), intent(inout) :: vec_out(n)
    real(8), intent(in)    :: vec_in(n)
    if (.not. c_associated(sector_h)) stop 'Hmultr, no array are associated!'
    call ed_check(ed_hmult_f64(sector_h, int(n, c_int64_t), vec_out, vec_in), 'Hmultr')
  end subroutine

  ! Hmultc (H_class.f90:69-98) -> HAIMupdo_multc on a sector made by new_H_complex
  subroutine Hmultc(n, vec_out, vec_in)
    integer                   :: n
    complex(8), intent(inout) :: vec_out(n)
    complex(8), intent(in)    :: vec_in(n)
    if (.not. c_associated(sector_h)) stop 'Hmultc, no array are associated!'
    call ed_check(ed_hmult_c128(sector_h, int(n, c_int64_t), vec_out, vec_in), 'Hmultc')
  end subroutine

  pure integer function dimen_H()
    interface
      pure integer(c_int64_t) function dim_pure(sec) bind(C, name="ed_sector_dim")
        import :: c_int64_t, c_ptr
        type(c_ptr), value :: sec
      end function
    end interface
    dimen_H = int(dim_pure(sector_h))
  end function

  ! Lanczos_GPU_updo(Niter, sizevec, diag, subdiag) (Lanczos_gpu.f90:106): Niter steps
  ! from the default start vector; subdiag(1) = ||v1|| as in the reference's storage.
  subroutine Lanczos_GPU_updo(Niter, sizevec, diag, subdiag)
    integer, intent(in)  :: Niter, sizevec
    real(8), intent(out) :: diag(Niter), subdiag(Niter)
    if (sizevec /= dimen_H()) stop 'Lanczos_GPU_updo: sizevec /= dimen_H()'
    call ed_check(ed_lanczos_run(sector_h, int(Niter, c_int), c_null_ptr, ED_PREC_F64, diag, subdiag), &
                  'Lanczos_GPU_updo')
  end subroutine

  ! Lanczos_GPU_get_GS_updo(iter, sizevec, vecp, vec) (Lanczos_gpu.f90:48).
  subroutine Lanczos_GPU_get_GS_updo(iter, sizevec, vecp, vec)
    integer, intent(in)  :: iter, sizevec
    real(8), intent(in)  :: vecp(iter)
    real(8), intent(out) :: vec(sizevec)
    real(8), allocatable :: tmp(:)
    tmp = vecp
    call ed_check(ed_lanczos_gs_vector(sector_h, int(iter, c_int), tmp, vec), 'Lanczos_GPU_get_GS_updo')
  end subroutine

  ! Lanczos_dynamic_GPU_updo(Niter, sizevec, diag, subdiag, vec)
  ! (green_class_compute_dynamic.f90:468): Niter steps from the excitation vector.
  subroutine Lanczos_dynamic_GPU_updo(Niter, sizevec, diag, subdiag, vec)
    integer, intent(in)         :: Niter, sizevec
    real(8), intent(out)        :: diag(Niter), subdiag(Niter)
    real(8), intent(in), target :: vec(sizevec)
    if (sizevec /= dimen_H()) stop 'Lanczos_dynamic_GPU_updo: sizevec /= dimen_H()'
    call ed_check(ed_lanczos_run(sector_h, int(Niter, c_int), c_loc(vec), ED_PREC_F64, diag, subdiag), &
                  'Lanczos_dynamic_GPU_updo')
  end subroutine

  ! The arpack_eigenvector_sym_matrix call of ARPACK_diago (ED_ARPACK.f90:498-525) on the
  ! device for the current H: nev = Neigen_ - 1 Ritz pairs from both ends ('BE'),
  ! ncv = Neigen_ (or ncv when given), tol = tolerance, maxitr 300, start vector
  ! create_fix_initial_vector's.  VALP(1:nconv) ascending, VECP(:, i) the vectors,
  ! rel_resid(i) = ||H x - VALP(i) x|| / |VALP(i)|; the dEmax window stays with the caller
  ! (:527-545).
  subroutine ARPACK_diago_hip(Neigen_, tolerance, VALP, VECP, rel_resid, nconv, ncv)
    integer, intent(in)           :: Neigen_
    real(8), intent(in)           :: tolerance
    real(8), intent(out)          :: VALP(:), rel_resid(:)
    real(8), intent(out)          :: VECP(:, :)
    integer, intent(out)          :: nconv
    integer, intent(in), optional :: ncv
    type(ed_arpack_opts_t) :: ao
    integer(c_int) :: nc, iters, mv
    if (Neigen_ < 2) stop 'ARPACK_diago_hip: Neigen_ < 2'
    if (size(VALP) < Neigen_ - 1 .or. size(rel_resid) < Neigen_ - 1 .or. size(VECP, 2) < Neigen_ - 1) &
      stop 'ARPACK_diago_hip: VALP, VECP or rel_resid shorter than Neigen_ - 1'
    if (size(VECP, 1) /= dimen_H()) stop 'ARPACK_diago_hip: size(VECP, 1) /= dimen_H()'
    call ed_check(ed_arpack_default_opts(ao), 'ed_arpack_default_opts')
    ao%nev = Neigen_ - 1
    ao%ncv = Neigen_
    if (present(ncv)) ao%ncv = max(Neigen_, ncv)
    ao%which = ED_ARPACK_BE
    ao%tol = tolerance
    ao%maxitr = 300
    call ed_check(ed_arpack(sector_h, ao, c_null_ptr, VALP, rel_resid, VECP, int(size(VECP, 1), c_int64_t), &
                            nc, iters, mv), 'ARPACK_diago_hip')
    nconv = nc
  end subroutine

end module H_class_hip
