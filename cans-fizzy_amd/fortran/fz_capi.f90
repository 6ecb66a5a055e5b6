! iso_c_binding interfaces to libfizzy (include/fizzy.h), one per C entry
! point the Fortran modules of this directory call.
module mod_fz_capi
  use, intrinsic :: iso_c_binding
  implicit none
  ! fz_params of include/fizzy.h
  type, bind(C) :: fz_params_t
    integer(c_int) :: ng(3)
    real(c_double) :: l(3)
    integer(c_int) :: gtype
    real(c_double) :: gr, cfl, dtmax, dt_f
    integer(c_int) :: is_solve_ns, is_track_interface, is_forced_hit, icheck, vof
    character(kind=c_char) :: cbcvel(18), cbcpre(6), cbcpsi(6), cbcnor(18)
    real(c_double) :: bcvel(18), bcpre(6), bcpsi(6), bcnor(18), bforce(3), gacc(3), sigma
    real(c_double) :: rho12(2), mu12(2), beta12(2), psi_thickness_factor, acdi_gam_factor, acdi_gam_min
    integer(c_int) :: nspheres, inipsi
    real(c_double) :: spheres(4*64)
  end type fz_params_t
  interface
    function fz_last_error() bind(C, name='fz_last_error') result(msg)
      import :: c_ptr
      type(c_ptr) :: msg
    end function
    integer(c_int) function fz_ctx_create(ctx, n, l, dzc, dzf, zc, zf, cbcpre, stream) bind(C, name='fz_ctx_create')
      import :: c_int, c_double, c_char, c_ptr
      type(c_ptr), intent(out) :: ctx
      integer(c_int), intent(in) :: n(3)
      real(c_double), intent(in) :: l(3), dzc(*), dzf(*), zc(*), zf(*)
      character(kind=c_char), intent(in) :: cbcpre(6)
      type(c_ptr), value :: stream
    end function
    integer(c_int) function fz_ctx_destroy(ctx) bind(C, name='fz_ctx_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function fz_ctx_sync(ctx) bind(C, name='fz_ctx_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function fz_malloc(ctx, bytes, dptr) bind(C, name='fz_malloc')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx
      integer(c_size_t), value :: bytes
      type(c_ptr), intent(out) :: dptr
    end function
    integer(c_int) function fz_free(ctx, dptr) bind(C, name='fz_free')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, dptr
    end function
    integer(c_int) function fz_upload(ctx, dst, src, bytes) bind(C, name='fz_upload')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function fz_download(ctx, dst, src, bytes) bind(C, name='fz_download')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function fz_copy(ctx, dst, src, bytes) bind(C, name='fz_copy')
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: ctx, dst, src
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function fz_initgrid(gtype, n, gr, lz, dzc, dzf, zc, zf) bind(C, name='fz_initgrid')
      import :: c_int, c_double
      integer(c_int), value :: gtype, n
      real(c_double), value :: gr, lz
      real(c_double), intent(out) :: dzc(*), dzf(*), zc(*), zf(*)
    end function
    integer(c_int) function fz_initsolver(ctx, bcpre, lambdaxy, a, b, c, normfft, rhsbx, rhsby, rhsbz) &
        bind(C, name='fz_initsolver')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: bcpre(6)
      real(c_double), intent(out) :: lambdaxy(*), a(*), b(*), c(*), normfft, rhsbx(*), rhsby(*), rhsbz(*)
    end function
    integer(c_int) function fz_solver(ctx, p) bind(C, name='fz_solver')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, p
    end function
    integer(c_int) function fz_fillps(ctx, dti, rho0, u, v, w, p) bind(C, name='fz_fillps')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, u, v, w, p
      real(c_double), value :: dti, rho0
    end function
    integer(c_int) function fz_updt_rhs_b(ctx, p) bind(C, name='fz_updt_rhs_b')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, p
    end function
    integer(c_int) function fz_correc(ctx, rho0, rho12, dt, p, psi, u, v, w) bind(C, name='fz_correc')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, p, psi, u, v, w
      real(c_double), value :: rho0, dt
      real(c_double), intent(in) :: rho12(2)
    end function
    integer(c_int) function fz_updatep(ctx, pp, p) bind(C, name='fz_updatep')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, pp, p
    end function
    integer(c_int) function fz_boundp(ctx, cbc, bc, p) bind(C, name='fz_boundp')
      import :: c_int, c_double, c_char, c_ptr
      type(c_ptr), value :: ctx, p
      character(kind=c_char), intent(in) :: cbc(6)
      real(c_double), intent(in) :: bc(6)
    end function
    integer(c_int) function fz_bounduvw(ctx, cbc, bc, is_correc, u, v, w) bind(C, name='fz_bounduvw')
      import :: c_int, c_double, c_char, c_ptr
      type(c_ptr), value :: ctx, u, v, w
      character(kind=c_char), intent(in) :: cbc(18)
      real(c_double), intent(in) :: bc(18)
      integer(c_int), value :: is_correc
    end function
    integer(c_int) function fz_acdi_transport_pf(ctx, gam, seps, u, v, w, normx, normy, normz, phi, psi, dpsidt, &
                                                 flux_x, flux_y, flux_z) bind(C, name='fz_acdi_transport_pf')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, u, v, w, normx, normy, normz, phi, psi, dpsidt, flux_x, flux_y, flux_z
      real(c_double), value :: gam, seps
    end function
    integer(c_int) function fz_acdi_cmpt_phi(ctx, seps, psi, phi) bind(C, name='fz_acdi_cmpt_phi')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, psi, phi
      real(c_double), value :: seps
    end function
    integer(c_int) function fz_acdi_set_gamma(ctx, gam_factor, u, v, w, gam) bind(C, name='fz_acdi_set_gamma')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, u, v, w
      real(c_double), value :: gam_factor
      real(c_double), intent(out) :: gam
    end function
    integer(c_int) function fz_acdi_set_epsilon(ctx, seps_factor, seps) bind(C, name='fz_acdi_set_epsilon')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: seps_factor
      real(c_double), intent(out) :: seps
    end function
    integer(c_int) function fz_vof_thinc_transport_psi(ctx, beta, u, v, w, normx, normy, normz, phi, psi, dpsidt, &
                                                       flux_x, flux_y, flux_z) bind(C, name='fz_vof_thinc_transport_psi')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, u, v, w, normx, normy, normz, phi, psi, dpsidt, flux_x, flux_y, flux_z
      real(c_double), value :: beta
    end function
    integer(c_int) function fz_vof_thinc_cmpt_phi(ctx, beta, psi, phi) bind(C, name='fz_vof_thinc_cmpt_phi')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, psi, phi
      real(c_double), value :: beta
    end function
    integer(c_int) function fz_clip_field(ctx, minmax, p) bind(C, name='fz_clip_field')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, p
      real(c_double), intent(in) :: minmax(2)
    end function
    integer(c_int) function fz_cmpt_norm_curv_youngs(ctx, psi, normx, normy, normz, kappa) &
        bind(C, name='fz_cmpt_norm_curv_youngs')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, psi, normx, normy, normz, kappa
    end function
    integer(c_int) function fz_init2fl(ctx, prm, seps, psi) bind(C, name='fz_init2fl')
      import :: c_int, c_double, c_ptr, fz_params_t
      type(c_ptr), value :: ctx, psi
      type(fz_params_t), intent(in) :: prm
      real(c_double), value :: seps
    end function
    integer(c_int) function fz_mom_xyz_ad(ctx, rho12, mu12, psiflx_x, psiflx_y, psiflx_z, u, v, w, psi, &
                                          dudt, dvdt, dwdt) bind(C, name='fz_mom_xyz_ad')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, psiflx_x, psiflx_y, psiflx_z, u, v, w, psi, dudt, dvdt, dwdt
      real(c_double), intent(in) :: rho12(2), mu12(2)
    end function
    integer(c_int) function fz_mom_xyz_oth(ctx, dt_r, rho12, beta12, bforce, gacc, sigma, rho0, rho_av, p, psi, &
                                           kappa, pn, po, dudt, dvdt, dwdt) bind(C, name='fz_mom_xyz_oth')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, p, psi, kappa, pn, po, dudt, dvdt, dwdt
      real(c_double), value :: dt_r, sigma, rho0
      real(c_double), intent(in) :: rho12(2), beta12(2), bforce(3), gacc(3), rho_av(3)
    end function
    integer(c_int) function fz_lscale_forcing(ctx, ftype, alpha, dt, u, v, w) bind(C, name='fz_lscale_forcing')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, u, v, w
      integer(c_int), value :: ftype
      real(c_double), value :: alpha, dt
    end function
    integer(c_int) function fz_bulk_mean_12_stag(ctx, idir, psi, p12, mean) bind(C, name='fz_bulk_mean_12_stag')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, psi
      integer(c_int), value :: idir
      real(c_double), intent(in) :: p12(2)
      real(c_double), intent(out) :: mean
    end function
    integer(c_int) function fz_rk_state_create(ctx, st) bind(C, name='fz_rk_state_create')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: st
    end function
    integer(c_int) function fz_rk_state_destroy(st) bind(C, name='fz_rk_state_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: st
    end function
    integer(c_int) function fz_rk(ctx, st, rkpar, dt, dt_r, bforce, gacc, sigma, rho_av, rho12, mu12, beta12, rho0, &
                                  psi, kappa, p, pn, po, psio, psiflx_x, psiflx_y, psiflx_z, u, v, w) &
        bind(C, name='fz_rk')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, st, psi, kappa, p, pn, po, psio, psiflx_x, psiflx_y, psiflx_z, u, v, w
      real(c_double), value :: dt, dt_r, sigma, rho0
      real(c_double), intent(in) :: rkpar(2), bforce(3), gacc(3), rho_av(3), rho12(2), mu12(2), beta12(2)
    end function
    integer(c_int) function fz_rk_2fl(ctx, st, rkpar, dt, gam, seps, beta, vof, u, v, w, normx, normy, normz, phi, &
                                      psi, psiflx_x, psiflx_y, psiflx_z) bind(C, name='fz_rk_2fl')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, st, u, v, w, normx, normy, normz, phi, psi, psiflx_x, psiflx_y, psiflx_z
      real(c_double), value :: dt, gam, seps, beta
      integer(c_int), value :: vof
      real(c_double), intent(in) :: rkpar(2)
    end function
    integer(c_int) function fz_chkdiv(ctx, u, v, w, divtot, divmax) bind(C, name='fz_chkdiv')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, u, v, w
      real(c_double), intent(out) :: divtot, divmax
    end function
    integer(c_int) function fz_chkdt(ctx, is_solve_ns, is_track_interface, vof, mu12, rho12, sigma, gacc, u, v, w, &
                                     gam, seps, dtmax) bind(C, name='fz_chkdt')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, u, v, w
      integer(c_int), value :: is_solve_ns, is_track_interface, vof
      real(c_double), intent(in) :: mu12(2), rho12(2), gacc(3)
      real(c_double), value :: sigma, gam, seps
      real(c_double), intent(out) :: dtmax
    end function
    integer(c_int) function fz_load_one(ctx, io, filename, p, time, istep) bind(C, name='fz_load_one')
      import :: c_int, c_double, c_char, c_ptr
      type(c_ptr), value :: ctx, p
      character(kind=c_char), value :: io
      character(kind=c_char), intent(in) :: filename(*)
      real(c_double), intent(inout) :: time
      integer(c_int), intent(inout) :: istep
    end function
    integer(c_int) function fz_out3d(ctx, filename, nskip, p) bind(C, name='fz_out3d')
      import :: c_int, c_char, c_ptr
      type(c_ptr), value :: ctx, p
      character(kind=c_char), intent(in) :: filename(*)
      integer(c_int), intent(in) :: nskip(3)
    end function
  end interface
end module mod_fz_capi
