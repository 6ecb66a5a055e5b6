! A main program over the drop-in modules: the initialisation and the time
! loop of src/main.f90 (:252-407 and :414-562) for the ACDI 3-D rising bubble
! (periodic x/y, no-slip walls in z, spheres from spheres.in, initflow 'zer'),
! with the reference's call sequence and argument lists. The arrays that the
! reference places on the device with `!$acc enter data` are device fields
! from fz_alloc; the full-array assignments between them (psio = psi,
! pp = p, po = pn, pn = p) are device copies.
!
! usage: fz_drive nx ny nz nstep rho1 rho2 mu1 mu2 outfile
! writes u, v, w, p, psi (halo'd, column-major, float64) and then
! time, dt, gam, divmax to outfile.
program fz_drive
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi,     only: fz_initgrid, fz_copy
  use mod_initsolver,  only: initsolver
  use mod_solver,      only: solver
  use mod_fillps,      only: fillps
  use mod_correc,      only: correc
  use mod_updatep,     only: updatep
  use mod_bound,       only: boundp, bounduvw, updt_rhs_b
  use mod_acdi,        only: acdi_set_epsilon, acdi_set_gamma, acdi_cmpt_phi
  use mod_two_fluid,   only: init2fl, cmpt_norm_curv => cmpt_norm_curv_youngs
  use mod_rk,          only: tm => rk, tm_2fl => rk_2fl
  use mod_chkdiv,      only: chkdiv
  use mod_chkdt,       only: chkdt
  implicit none
  integer :: ng(3), n(3), lo(3), hi(3), nstep, istep, icheck, k, q, iunit
  real(rp) :: l(3), dl(3), dli(3), rho12(2), mu12(2), beta12(2), sigma, gacc(3), bforce(3), rho_av(3)
  real(rp) :: cfl, dtmax, dt_f, dt, dto, dt_r, dtrk, dtrki, time, dt_cfl, gam, seps, rho0, divtot, divmax
  real(rp) :: psi_thickness_factor, acdi_gam_factor, acdi_gam_min, tm_coeff(2), normfft
  real(rp), allocatable, dimension(:) :: dzc, dzf, zc, zf, dzci, dzfi, a, b, c
  real(rp), allocatable :: lambdaxy(:,:), rhsbx(:,:,:), rhsby(:,:,:), rhsbz(:,:,:), hbuf(:,:,:)
  real(rp), pointer, contiguous, dimension(:,:,:) :: u, v, w, p, pp, pn, po, psi, psio, phi, kappa, normx, normy, &
                                                     normz, psiflx_x, psiflx_y, psiflx_z
  character(len=1), dimension(0:1,3,3) :: cbcvel, cbcnor
  character(len=1), dimension(0:1,3) :: cbcpre, cbcpsi, cbccur
  real(rp), dimension(0:1,3,3) :: bcvel, bcnor
  real(rp), dimension(0:1,3) :: bcpre, bcpsi, bccur
  integer, dimension(0:1,3) :: nb
  logical, dimension(0:1,3) :: is_bound
  type(c_ptr), dimension(2,2) :: arrplan
  logical :: is_solve_ns, is_track_interface
  character(len=256) :: arg, outfile
  integer(c_int) :: rc
  !
  do q = 1, 3
    call get_command_argument(q, arg); read(arg, *) ng(q)
  end do
  call get_command_argument(4, arg); read(arg, *) nstep
  call get_command_argument(5, arg); read(arg, *) rho12(1)
  call get_command_argument(6, arg); read(arg, *) rho12(2)
  call get_command_argument(7, arg); read(arg, *) mu12(1)
  call get_command_argument(8, arg); read(arg, *) mu12(2)
  call get_command_argument(9, outfile)
  !
  ! the case (the namelist of the rising-bubble example)
  l = [1._rp, 1._rp, 2._rp]
  cfl = 0.25_rp; dtmax = 1.e9_rp; dt_f = -1._rp; icheck = 10
  sigma = 24.5_rp; gacc = [0._rp, 0._rp, -0.98_rp]; bforce = 0._rp; beta12 = 1._rp
  psi_thickness_factor = 1._rp; acdi_gam_factor = 1._rp; acdi_gam_min = 1.e-12_rp
  is_solve_ns = .true.; is_track_interface = .true.
  do q = 1, 3
    cbcvel(:,1,q) = 'P'; cbcvel(:,2,q) = 'P'; cbcvel(:,3,q) = 'D'
  end do
  cbcpre(:,1) = 'P'; cbcpre(:,2) = 'P'; cbcpre(:,3) = 'N'
  cbcpsi = cbcpre
  cbcnor(:,:,1) = cbcpre; cbcnor(:,:,2) = cbcpre
  cbcnor(:,1,3) = 'P'; cbcnor(:,2,3) = 'P'; cbcnor(:,3,3) = 'D'
  bcvel = 0._rp; bcpre = 0._rp; bcpsi = 0._rp; bcnor = 0._rp
  cbccur = cbcpsi; bccur = bcpre                     ! src/param.f90:172-173
  rho0 = minval(rho12)
  nb = -1; is_bound = .true.
  n = ng; lo = 1; hi = ng
  !
  ! grid, src/main.f90:252-270
  allocate(dzc(0:n(3)+1), dzf(0:n(3)+1), zc(0:n(3)+1), zf(0:n(3)+1), dzci(0:n(3)+1), dzfi(0:n(3)+1))
  rc = fz_initgrid(1_c_int, int(n(3), c_int), 0._rp, l(3), dzc, dzf, zc, zf)
  if (rc /= 0) error stop 'fz_initgrid failed'
  dzci = dzc**(-1); dzfi = dzf**(-1)
  dl = l/(1.*ng); dli = dl**(-1)
  call fz_setup(ng, l, dzc, dzf, zc, zf, cbcpre)
  !
  ! fields
  do q = 1, 14
    select case(q)
    case(1); call fz_alloc(u, [0,0,0], n+1)
    case(2); call fz_alloc(v, [0,0,0], n+1)
    case(3); call fz_alloc(w, [0,0,0], n+1)
    case(4); call fz_alloc(p, [0,0,0], n+1)
    case(5); call fz_alloc(pp, [0,0,0], n+1)
    case(6); call fz_alloc(pn, [0,0,0], n+1)
    case(7); call fz_alloc(po, [0,0,0], n+1)
    case(8); call fz_alloc(psi, [0,0,0], n+1)
    case(9); call fz_alloc(psio, [0,0,0], n+1)
    case(10); call fz_alloc(phi, [0,0,0], n+1)
    case(11); call fz_alloc(kappa, [0,0,0], n+1)
    case(12); call fz_alloc(normx, [0,0,0], n+1)
    case(13); call fz_alloc(normy, [0,0,0], n+1)
    case(14); call fz_alloc(normz, [0,0,0], n+1)
    end select
  end do
  call fz_alloc(psiflx_x, [0,0,0], n+1); call fz_alloc(psiflx_y, [0,0,0], n+1); call fz_alloc(psiflx_z, [0,0,0], n+1)
  !
  ! Poisson solver, src/main.f90:281-282
  allocate(lambdaxy(n(1),n(2)), a(n(3)), b(n(3)), c(n(3)), rhsbx(n(2),n(3),0:1), rhsby(n(1),n(3),0:1), &
           rhsbz(n(1),n(2),0:1))
  call initsolver(ng, [1,1,1], [1,1,1], [1,1,1], ng, dli, dzci, dzfi, cbcpre, bcpre, lambdaxy, ['c','c','c'], &
                  a, b, c, arrplan, normfft, rhsbx, rhsby, rhsbz)
  !
  call acdi_set_epsilon(dl, dzfi, psi_thickness_factor, seps)
  time = 0._rp
  ! initflow 'zer' (src/initflow.f90:69-70), then init2fl
  allocate(hbuf(0:n(1)+1, 0:n(2)+1, 0:n(3)+1)); hbuf = 0._rp
  call fz_to_device(u, hbuf); call fz_to_device(v, hbuf); call fz_to_device(w, hbuf); call fz_to_device(p, hbuf)
  call init2fl('bub3', cbcpsi, seps, lo, hi, l, dl, zc, psi)
  call bounduvw(cbcvel, n, bcvel, nb, is_bound, .false., dl, dzc, dzf, u, v, w)
  call boundp(cbcpre, n, bcpre, nb, is_bound, dl, dzc, p)
  call devcopy(pn, p); call devcopy(po, pn)
  call boundp(cbcpsi, n, bcpsi, nb, is_bound, dl, dzc, psi)
  call acdi_cmpt_phi(n, seps, psi, phi)
  call cmpt_norm_curv(n, dli, dzci, dzfi, phi, normx, normy, normz, kappa)
  call boundp(cbccur, n, bccur, nb, is_bound, dl, dzc, kappa)
  call boundp(cbcnor(:,:,1), n, bcnor(:,:,1), nb, is_bound, dl, dzc, normx)
  call boundp(cbcnor(:,:,2), n, bcnor(:,:,2), nb, is_bound, dl, dzc, normy)
  call boundp(cbcnor(:,:,3), n, bcnor(:,:,3), nb, is_bound, dl, dzc, normz)
  call acdi_set_gamma(n, acdi_gam_factor, u, v, w, gam)
  gam = max(gam, acdi_gam_min)
  call chkdt(n, dl, dzci, dzfi, is_solve_ns, is_track_interface, mu12, rho12, sigma, gacc, u, v, w, dt_cfl, gam, seps)
  dt = min(cfl*dt_cfl, dtmax)
  dto = dt
  divmax = 0._rp
  !
  ! time loop, src/main.f90:414-562
  do istep = 1, nstep
    time = time + dt
    dt_r = dt/dto
    tm_coeff(1) = 1. + 0.5*dt_r
    tm_coeff(2) =    - 0.5*dt_r
    dtrk = dt
    dtrki = dtrk**(-1)
    call devcopy(psio, psi)
    call tm_2fl(tm_coeff, n, dli, dzci, dzfi, dt, gam, seps, 0._rp, u, v, w, normx, normy, normz, phi, psi, &
                psiflx_x, psiflx_y, psiflx_z)
    call bounduvw(cbcvel, n, bcvel, nb, is_bound, .false., dl, dzc, dzf, psiflx_x, psiflx_y, psiflx_z)
    call boundp(cbcpsi, n, bcpsi, nb, is_bound, dl, dzc, psi)
    call acdi_cmpt_phi(n, seps, psi, phi)
    call cmpt_norm_curv(n, dli, dzci, dzfi, phi, normx, normy, normz, kappa)
    call boundp(cbccur, n, bccur, nb, is_bound, dl, dzc, kappa)
    call boundp(cbcnor(:,:,1), n, bcnor(:,:,1), nb, is_bound, dl, dzc, normx)
    call boundp(cbcnor(:,:,2), n, bcnor(:,:,2), nb, is_bound, dl, dzc, normy)
    call boundp(cbcnor(:,:,3), n, bcnor(:,:,3), nb, is_bound, dl, dzc, normz)
    rho_av(:) = 0.   ! gravity acts along the walled z direction only
    call tm(tm_coeff, n, dli, dzci, dzfi, dt, dt_r, bforce, gacc, sigma, rho_av, rho12, mu12, beta12, rho0, psi, &
            kappa, p, pn, po, psio=psio, psiflx_x=psiflx_x, psiflx_y=psiflx_y, psiflx_z=psiflx_z, u=u, v=v, w=w)
    call bounduvw(cbcvel, n, bcvel, nb, is_bound, .false., dl, dzc, dzf, u, v, w)
    call devcopy(pp, p)
    call fillps(n, dli, dzfi, dtrki, rho0, u, v, w, p)
    call updt_rhs_b(['c','c','c'], cbcpre, n, is_bound, rhsbx, rhsby, rhsbz, p)
    call solver(n, ng, arrplan, normfft, lambdaxy, a, b, c, cbcpre, ['c','c','c'], p)
    call boundp(cbcpre, n, bcpre, nb, is_bound, dl, dzc, p)
    call correc(n, dli, dzci, rho0, rho12, dtrk, p, psi, u, v, w)
    call bounduvw(cbcvel, n, bcvel, nb, is_bound, .true., dl, dzc, dzf, u, v, w)
    call updatep(pp, p)
    call boundp(cbcpre, n, bcpre, nb, is_bound, dl, dzc, p)
    call devcopy(po, pn)
    call devcopy(pn, p)
    dto = dt
    call acdi_set_gamma(n, acdi_gam_factor, u, v, w, gam)
    gam = max(gam, acdi_gam_min)
    if (mod(istep, icheck) == 0) then
      call chkdt(n, dl, dzci, dzfi, is_solve_ns, is_track_interface, mu12, rho12, sigma, gacc, u, v, w, dt_cfl, gam, &
                 seps)
      dt = min(cfl*dt_cfl, dtmax)
      call chkdiv(lo, hi, dli, dzfi, u, v, w, divtot, divmax)
    end if
  end do
  call fz_sync()
  !
  open(newunit=iunit, file=trim(outfile), access='stream', form='unformatted', status='replace')
  call dump(u); call dump(v); call dump(w); call dump(p); call dump(psi)
  write(iunit) time, dt, gam, divmax
  close(iunit)
  print '(a,i0,a,es12.5,a,es10.3)', 'fz_drive: ', nstep, ' steps, time ', time, ', divmax ', divmax
contains
  subroutine devcopy(dst, src)
    real(rp), pointer, contiguous, intent(in) :: dst(:,:,:), src(:,:,:)
    call fz_check(fz_copy(fz_ctx_ptr(), c_loc(dst), c_loc(src), int(size(src), c_size_t)*8), 'devcopy')
  end subroutine devcopy
  subroutine dump(f)
    real(rp), pointer, contiguous, intent(in) :: f(:,:,:)
    call fz_to_host(hbuf, f)
    write(iunit) hbuf
  end subroutine dump
end program fz_drive
