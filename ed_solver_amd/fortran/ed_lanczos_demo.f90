! Fortran host driver over H_class_hip.
!
! Reads a model file (first line: nc nb nup ndn nitermax; then Ec(2*nc*nc),
! Eb(2*nb*nb), Vbc(2*nb*nc), U(nc*nc), JJ(nc*nc) in C row-major order, then Jhund
! and cutoff) and runs:
!   1. the plain Lanczos loop of Lanczos_fast_diagonalize (Lanczos_plain.f90:33-134
!      with lanc_opt, Lanczos_plain.h:66-173) on the host, calling Hmult__ for every
!      H.v -- the reference's own solver shape on top of the drop-in Hmult__;
!   2. the coarse device path Lanczos_GPU_updo + Lanczos_GPU_get_GS_updo;
!   3. Block_Lanczos_diagonalize through ed_block_lanczos (block size 4, 4 states) when
!      the sector has at least 16 states;
!   4. the reference's get_GS order (solver.h:43-60): delete_H, new_H again, then
!      Lanczos_GPU_get_GS_updo on the fresh handle, which runs the second pass of
!      Lanczos_get_GS_sector (Lanczos_plain.f90:156-229); the vector is written to the
!      file named by the second argument, when given (raw REAL(8) stream);
!   5. the S_z branch of new_H (new_H_sz, no pairing) with the host Lanczos loop over
!      Hmult__, when that sector has at most 10^6 states;
!   6. ARPACK_diago's eigensolver on the device (ARPACK_diago_hip: Neigen_ = 5, so 4 Ritz
!      pairs from both ends of the spectrum, ncv 24) on the (N_up, N_dn) sector.
! Prints E0 from each, the iteration count and the Rayleigh quotient of the device
! ground state.
program ed_lanczos_demo
  use iso_c_binding
  use ed_hip
  use H_class_hip
  implicit none
  integer :: nc, nb, nup, ndn, nitermax, n, niter, k, npart, nsz, u_out
  real(8), allocatable :: ec(:), eb(:), vbc(:), u(:), jj(:), pb(:), pvbc(:)
  real(8) :: jhund, cutoff
  real(8), allocatable :: diag(:), subdiag(:), w(:), z(:), psi(:), hpsi(:), psi2(:)
  real(8) :: e0_host, e0_dev, rq, e0_sz
  type(ed_block_opts_t) :: bo
  real(8) :: valp(4, 2), valp_a(4), res_a(4)
  integer :: nconv_a
  real(8), allocatable :: vecp(:, :)
  integer(c_int) :: nconv, nsteps
  character(len=512) :: fname, fout
  integer :: u_in

  call get_command_argument(1, fname)
  open(newunit=u_in, file=trim(fname), status='old', action='read')
  read(u_in, *) nc, nb, nup, ndn, nitermax
  allocate(ec(2*nc*nc), eb(max(1, 2*nb*nb)), vbc(max(1, 2*nb*nc)), u(nc*nc), jj(nc*nc))
  eb = 0; vbc = 0
  read(u_in, *) ec
  if (nb > 0) read(u_in, *) eb(1:2*nb*nb)
  if (nb > 0) read(u_in, *) vbc(1:2*nb*nc)
  read(u_in, *) u
  read(u_in, *) jj
  read(u_in, *) jhund, cutoff
  close(u_in)

  call ed_check(ed_init(0_c_int), 'ed_init')
  call new_H(nc, nb, ec, eb, vbc, u, jj, jhund, cutoff, nup, ndn)
  n = dimen_H()
  niter = min(n - 1, nitermax)
  allocate(diag(niter), subdiag(niter), w(niter), z(niter*niter))

  ! ---- 1. host Lanczos loop over Hmult__ (one_step_Lanczos_fast + lanc_opt)
  call host_lanczos(n, niter, diag, subdiag, e0_host, k)
  write(*, '(a,f22.14,a,i0)') 'E0_host_loop ', e0_host, ' iters ', k

  ! ---- 2. coarse device path: Lanczos_GPU_updo + Lanczos_GPU_get_GS_updo
  call Lanczos_GPU_updo(k, n, diag(1:k), subdiag(1:k))
  call ed_check(ed_tridiag_eig(int(k, c_int), diag, subdiag, w, z), 'ed_tridiag_eig')
  e0_dev = w(1)
  allocate(psi(n), hpsi(n))
  call Lanczos_GPU_get_GS_updo(k, n, z(1:k), psi)
  call Hmult__(n, hpsi, psi)
  rq = sum(psi * hpsi) / sum(psi * psi)
  write(*, '(a,f22.14)') 'E0_gpu_updo ', e0_dev
  write(*, '(a,f22.14)') 'E0_rayleigh ', rq

  ! ---- 3. Block_Lanczos_diagonalize (Block_Lanczos.f90:37-355) on the device
  if (n >= 16) then
    call ed_check(ed_block_default_opts(bo), 'ed_block_default_opts')
    bo%block_size = 4; bo%nev = 4; bo%max_steps = min(nitermax, n / 4); bo%tolerance = 1d-12
    allocate(vecp(n, 4))
    call ed_check(ed_block_lanczos(sector_handle(), bo, valp(:, 1), valp(:, 2), vecp, int(n, c_int64_t), &
                                   nconv, nsteps), 'ed_block_lanczos')
    write(*, '(a,4f22.14,a,i0,a,i0)') 'E_block ', valp(1:4, 1), ' nconv ', nconv, ' steps ', nsteps
  end if

  ! ---- 6. ARPACK_diago on the device (before the handle is rebuilt: any H will do)
  if (n >= 30) then
    block
      real(8), allocatable :: vecp_a(:, :)
      allocate(vecp_a(n, 4))
      call ARPACK_diago_hip(5, 1d-12, valp_a, vecp_a, res_a, nconv_a, ncv=24)
      write(*, '(a,4f22.14,a,i0,a,es10.2)') 'E_arpack ', valp_a, ' nconv ', nconv_a, ' maxres ', &
        maxval(res_a(1:max(1, nconv_a)))
    end block
  end if

  ! ---- 4. get_GS order: the handle is rebuilt before the ground-state vector is formed
  call delete_H()
  call new_H(nc, nb, ec, eb, vbc, u, jj, jhund, cutoff, nup, ndn)
  allocate(psi2(n))
  call Lanczos_GPU_get_GS_updo(k, n, z(1:k), psi2)
  write(*, '(a,es12.4)') 'GS_second_pass_overlap ', abs(sum(psi * psi2))
  if (command_argument_count() >= 2) then
    call get_command_argument(2, fout)
    open(newunit=u_out, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
    write(u_out) psi2
    close(u_out)
  end if
  call delete_H()

  ! ---- 5. the S_z branch of new_H (no pairing): the Nambu sector npart = N_up + (Ns - N_dn)
  npart = nup + (nc + nb - ndn)
  nsz = 1
  do k = 1, npart
    nsz = nsz * (2 * (nc + nb) - k + 1) / k
  end do
  if (nsz <= 1000000) then
    allocate(pb(max(1, nb * nb)), pvbc(max(1, 2 * nb * nc)))
    pb = 0; pvbc = 0
    call new_H_sz(nc, nb, ec, eb, vbc, u, jj, jhund, cutoff, pb, pvbc, 0, npart)
    n = dimen_H()
    niter = min(n - 1, nitermax)
    deallocate(diag, subdiag)
    allocate(diag(niter), subdiag(niter))
    call host_lanczos(n, niter, diag, subdiag, e0_sz, k)
    write(*, '(a,f22.14,a,i0,a,i0)') 'E0_sz_host_loop ', e0_sz + energy_shift_H(), ' iters ', k, ' dim ', n
    call delete_H()
  end if

contains

  ! Lanczos_fast_diagonalize with one_step_Lanczos_fast / lanc_opt (Lanczos_plain.f90:33-134,
  ! Lanczos_plain.h:66-173) on host vectors, every H.v through Hmult__; stops on the
  ! reference's test for the lowest eigenvalue at tolerance 1e-12.
  subroutine host_lanczos(n, niter, diag, subdiag, e0, iters)
    integer, intent(in)  :: n, niter
    real(8), intent(out) :: diag(niter), subdiag(niter), e0
    integer, intent(out) :: iters
    real(8), allocatable :: vin(:), vout(:), tmp(:), w(:), z(:)
    real(8) :: normv_in, normv_out, aa, ddiag, last
    integer :: iter
    allocate(vin(n), vout(n), tmp(n), w(niter), z(niter * niter))
    call ed_check(ed_start_vector(int(n, c_int64_t), vin), 'ed_start_vector')
    tmp = 0; diag = 0; subdiag = 0
    normv_in = sqrt(sum(vin*vin)); normv_out = 0; subdiag(1) = normv_in
    last = huge(1d0); e0 = 0
    do iter = 1, niter
      if (iter == 1) then
        aa = 1d0 / normv_in
      else
        if (abs(normv_out) < 1d-25) stop '0 norm vector in lanczos, critical'
        aa = 1d0 / normv_out
      end if
      vin = vin * aa
      call Hmult__(n, vout, vin)
      if (iter > 1) vout = vout - subdiag(iter) * tmp
      ddiag = sum(vout * vin)
      diag(iter) = ddiag
      vout = vout - ddiag * vin
      normv_out = sqrt(sum(vout * vout))
      if (iter < niter) subdiag(iter + 1) = normv_out
      tmp = vin
      vin = vout
      call ed_check(ed_tridiag_eig(int(iter, c_int), diag, subdiag, w, z), 'ed_tridiag_eig')
      e0 = w(1)
      if ((abs(w(1) - last) < 1d-12 .and. iter > 2) .or. iter == niter .or. (iter == 1 .and. normv_out < 1d-24)) exit
      last = w(1)
    end do
    iters = min(iter, niter)
  end subroutine

end program ed_lanczos_demo
