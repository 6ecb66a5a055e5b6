! Drop-in Fortran modules for the per-timestep hot path of CaNS-Fizzy.
!
! Each module keeps the name, the public routines and the argument lists of
! the reference module it replaces (src/*.f90, cited per routine), so that a
! main program written against the reference (src/main.f90) compiles against
! these unchanged. The work is done by libfizzy on the device: array
! arguments are device-resident fields (Fortran pointers over device memory
! from fz_alloc, the stand-in for the reference's `!$acc enter data`) and are
! handed to the C ABI by address; the host never reads or writes their
! elements. Arguments that describe the grid (n, dli, dzci, ...) are part of
! the reference interface and are checked against the context set up once by
! fz_setup.
module mod_types
  use, intrinsic :: iso_c_binding, only: c_double, c_float, c_int64_t, c_int
  implicit none
  integer, parameter, public :: sp = c_float, dp = c_double, i8 = c_int64_t
  integer, parameter, public :: rp = dp
  ! the interface-capturing build switch of the reference (configs/flags.mk)
#if defined(_INTERFACE_CAPTURING_VOF)
  integer(c_int), parameter, public :: vof_build = 1
#else
  integer(c_int), parameter, public :: vof_build = 0
#endif
end module mod_types

module mod_fz
  ! the device context shared by the modules below, and device-field helpers
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz_capi
  implicit none
  private
  public fz_setup, fz_teardown, fz_check, fz_ctx_ptr, fz_alloc, fz_dealloc, fz_to_device, fz_to_host, &
         fz_sync, fz_ng
  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: ng_(3) = 0
contains
  subroutine fz_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: q
    character(len=512) :: text
    if (rc == 0) return
    call c_f_pointer(fz_last_error(), msg, [512])
    text = ''
    do q = 1, 512
      if (msg(q) == c_null_char) exit
      text(q:q) = msg(q)
    end do
    write(*, '(3a)') trim(where), ': ', trim(text)
    error stop 'libfizzy call failed'
  end subroutine fz_check

  ! one-time set-up after initgrid (src/main.f90:252-291): the grid and the
  ! pressure boundary conditions the Poisson solver is built for
  subroutine fz_setup(ng, l, dzc, dzf, zc, zf, cbcpre)
    integer , intent(in), dimension(3) :: ng
    real(rp), intent(in), dimension(3) :: l
    real(rp), intent(in), dimension(0:) :: dzc, dzf, zc, zf
    character(len=1), intent(in), dimension(0:1,3) :: cbcpre
    character(kind=c_char) :: cb(6)
    integer(c_int) :: n(3)
    integer :: q
    if (c_associated(ctx)) call fz_teardown()
    do q = 0, 5
      cb(q + 1) = cbcpre(mod(q, 2), q/2 + 1)
    end do
    n = int(ng, c_int)
    call fz_check(fz_ctx_create(ctx, n, l, dzc, dzf, zc, zf, cb, c_null_ptr), 'fz_setup')
    ng_ = ng
  end subroutine fz_setup

  subroutine fz_teardown()
    if (c_associated(ctx)) call fz_check(fz_ctx_destroy(ctx), 'fz_teardown')
    ctx = c_null_ptr
  end subroutine fz_teardown

  type(c_ptr) function fz_ctx_ptr() result(c)
    if (.not. c_associated(ctx)) error stop 'libfizzy: fz_setup has not been called'
    c = ctx
  end function fz_ctx_ptr

  function fz_ng() result(n)
    integer :: n(3)
    n = ng_
  end function fz_ng

  subroutine fz_sync()
    call fz_check(fz_ctx_sync(fz_ctx_ptr()), 'fz_sync')
  end subroutine fz_sync

  ! a device field with the given bounds (e.g. [0,0,0] to n+1 for a halo'd
  ! field, [1,1,1] to n for an AB2 right-hand side)
  subroutine fz_alloc(p, lo, hi)
    real(rp), pointer, contiguous, intent(out) :: p(:,:,:)
    integer, intent(in) :: lo(3), hi(3)
    real(rp), pointer, contiguous :: flat(:,:,:)
    type(c_ptr) :: d
    integer(c_size_t) :: bytes
    bytes = int(hi(1) - lo(1) + 1, c_size_t)*int(hi(2) - lo(2) + 1, c_size_t)*int(hi(3) - lo(3) + 1, c_size_t)*8
    call fz_check(fz_malloc(fz_ctx_ptr(), bytes, d), 'fz_alloc')
    call c_f_pointer(d, flat, hi - lo + 1)
    p(lo(1):, lo(2):, lo(3):) => flat
  end subroutine fz_alloc

  subroutine fz_dealloc(p)
    real(rp), pointer, contiguous, intent(inout) :: p(:,:,:)
    if (associated(p)) call fz_check(fz_free(fz_ctx_ptr(), c_loc(p)), 'fz_dealloc')
    nullify(p)
  end subroutine fz_dealloc

  subroutine fz_to_device(p, h)
    real(rp), pointer, contiguous, intent(in) :: p(:,:,:)
    real(rp), target, contiguous, intent(in) :: h(:,:,:)
    if (size(h) /= size(p)) error stop 'fz_to_device: size mismatch'
    call fz_check(fz_upload(fz_ctx_ptr(), c_loc(p), c_loc(h), int(size(h), c_size_t)*8), 'fz_to_device')
  end subroutine fz_to_device

  subroutine fz_to_host(h, p)
    real(rp), target, contiguous, intent(inout) :: h(:,:,:)
    real(rp), pointer, contiguous, intent(in) :: p(:,:,:)
    if (size(h) /= size(p)) error stop 'fz_to_host: size mismatch'
    call fz_check(fz_download(fz_ctx_ptr(), c_loc(h), c_loc(p), int(size(h), c_size_t)*8), 'fz_to_host')
  end subroutine fz_to_host

end module mod_fz

module mod_initsolver
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_initsolver
  implicit none
  private
  public initsolver
contains
  ! src/initsolver.f90:15-79; arrplan(1,1) holds the context that owns the
  ! device plan (the reference's plan handles)
  subroutine initsolver(ng, n_x_fft, n_y_fft, lo_z, hi_z, dli, dzci, dzfi, cbc, bc, lambdaxy, c_or_f, a, b, c, arrplan, &
                        normfft, rhsbx, rhsby, rhsbz)
    integer , intent(in), dimension(3) :: ng, n_x_fft, n_y_fft, lo_z, hi_z
    real(rp), intent(in), dimension(3) :: dli
    real(rp), intent(in), dimension(0:) :: dzci, dzfi
    character(len=1), intent(in), dimension(0:1,3) :: cbc
    real(rp), intent(in), dimension(0:1,3) :: bc
    real(rp), intent(out), dimension(lo_z(1):,lo_z(2):), target :: lambdaxy
    character(len=1), intent(in), dimension(3) :: c_or_f
    real(rp), intent(out), dimension(:) :: a, b, c
    type(c_ptr), intent(out), dimension(2,2) :: arrplan
    real(rp), intent(out) :: normfft
    real(rp), intent(out), dimension(:,:,0:) :: rhsbx, rhsby, rhsbz
    real(rp) :: bc6(6)
    if (any(c_or_f /= 'c')) error stop 'initsolver: only the cell-centred pressure solver is supported'
    if (any(ng /= fz_ng())) error stop 'initsolver: grid differs from fz_setup'
    bc6 = reshape(bc, [6])
    call fz_check(fz_initsolver(fz_ctx_ptr(), bc6, lambdaxy, a, b, c, normfft, rhsbx, rhsby, rhsbz), 'initsolver')
    arrplan(:,:) = c_null_ptr
    arrplan(1,1) = fz_ctx_ptr()
  end subroutine initsolver
end module mod_initsolver

module mod_solver
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_solver
  implicit none
  private
  public solver
contains
  ! src/solver.f90:19-78: x/y r2r transforms, z Thomas, inverse, in place on p
  subroutine solver(n, ng, arrplan, normfft, lambdaxy, a, b, c, bc, c_or_f, p)
    integer , intent(in), dimension(3) :: n, ng
    type(c_ptr), intent(in), dimension(2,2) :: arrplan
    real(rp), intent(in) :: normfft
    real(rp), intent(in), dimension(:,:) :: lambdaxy
    real(rp), intent(in), dimension(:) :: a, b, c
    character(len=1), dimension(0:1,3), intent(in) :: bc
    character(len=1), intent(in), dimension(3) :: c_or_f
    real(rp), intent(inout), dimension(0:,0:,0:), target, contiguous :: p
    call fz_check(fz_solver(arrplan(1,1), c_loc(p)), 'solver')
  end subroutine solver
end module mod_solver

module mod_fillps
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_fillps
  implicit none
  private
  public fillps
contains
  ! src/fillps.f90:13-51
  subroutine fillps(n, dli, dzfi, dti, rho0, u, v, w, p)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(3) :: dli
    real(rp), intent(in), dimension(0:) :: dzfi
    real(rp), intent(in) :: dti, rho0
    real(rp), intent(in ), dimension(0:,0:,0:), target, contiguous :: u, v, w
    real(rp), intent(out), dimension(0:,0:,0:), target, contiguous :: p
    call fz_check(fz_fillps(fz_ctx_ptr(), dti, rho0, c_loc(u), c_loc(v), c_loc(w), c_loc(p)), 'fillps')
  end subroutine fillps
end module mod_fillps

module mod_correc
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_correc
  implicit none
  private
  public correc
contains
  ! src/correc.f90:13-67
  subroutine correc(n, dli, dzci, rho0, rho12, dt, p, psi, u, v, w)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(3) :: dli
    real(rp), intent(in), dimension(0:) :: dzci
    real(rp), intent(in) :: rho0, rho12(2)
    real(rp), intent(in) :: dt
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous :: p, psi
    real(rp), intent(inout), dimension(0:,0:,0:), target, contiguous :: u, v, w
    call fz_check(fz_correc(fz_ctx_ptr(), rho0, rho12, dt, c_loc(p), c_loc(psi), c_loc(u), c_loc(v), c_loc(w)), &
                  'correc')
  end subroutine correc
end module mod_correc

module mod_updatep
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_updatep
  implicit none
  private
  public updatep
contains
  ! src/updatep.f90:13-24
  subroutine updatep(pp, p)
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous :: pp
    real(rp), intent(inout), dimension(0:,0:,0:), target, contiguous :: p
    call fz_check(fz_updatep(fz_ctx_ptr(), c_loc(pp), c_loc(p)), 'updatep')
  end subroutine updatep
end module mod_updatep

module mod_bound
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_boundp, fz_bounduvw, fz_updt_rhs_b
  implicit none
  private
  public boundp, bounduvw, updt_rhs_b
contains
  ! src/bound.f90:15-79
  subroutine bounduvw(cbc, n, bc, nb, is_bound, is_correc, dl, dzc, dzf, u, v, w)
    character(len=1), intent(in), dimension(0:1,3,3) :: cbc
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(0:1,3,3) :: bc
    integer , intent(in), dimension(0:1,3  ) :: nb
    logical , intent(in), dimension(0:1,3  ) :: is_bound
    logical , intent(in)                     :: is_correc
    real(rp), intent(in), dimension(3 ) :: dl
    real(rp), intent(in), dimension(0:) :: dzc, dzf
    real(rp), intent(inout), dimension(0:,0:,0:), target, contiguous :: u, v, w
    character(kind=c_char) :: cb(18)
    real(rp) :: b18(18)
    integer :: q
    do q = 0, 17
      cb(q + 1) = cbc(mod(q, 2), mod(q/2, 3) + 1, q/6 + 1)
    end do
    b18 = reshape(bc, [18])
    call fz_check(fz_bounduvw(fz_ctx_ptr(), cb, b18, merge(1_c_int, 0_c_int, is_correc), c_loc(u), c_loc(v), &
                              c_loc(w)), 'bounduvw')
  end subroutine bounduvw

  ! src/bound.f90:81-124
  subroutine boundp(cbc, n, bc, nb, is_bound, dl, dzc, p)
    character(len=1), intent(in), dimension(0:1,3) :: cbc
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(0:1,3) :: bc
    integer , intent(in), dimension(0:1,3) :: nb
    logical , intent(in), dimension(0:1,3) :: is_bound
    real(rp), intent(in), dimension(3 ) :: dl
    real(rp), intent(in), dimension(0:) :: dzc
    real(rp), intent(inout), dimension(0:,0:,0:), target, contiguous :: p
    character(kind=c_char) :: cb(6)
    real(rp) :: b6(6)
    integer :: q
    do q = 0, 5
      cb(q + 1) = cbc(mod(q, 2), q/2 + 1)
    end do
    b6 = reshape(bc, [6])
    call fz_check(fz_boundp(fz_ctx_ptr(), cb, b6, c_loc(p)), 'boundp')
  end subroutine boundp

  ! src/bound.f90:341-396 (the boundary terms are the ones initsolver built)
  subroutine updt_rhs_b(c_or_f, cbc, n, is_bound, rhsbx, rhsby, rhsbz, p)
    character(len=1), intent(in), dimension(3    ) :: c_or_f
    character(len=1), intent(in), dimension(0:1,3) :: cbc
    integer , intent(in), dimension(3) :: n
    logical , intent(in), dimension(0:1,3) :: is_bound
    real(rp), intent(in), dimension(:,:,0:), optional :: rhsbx, rhsby, rhsbz
    real(rp), intent(inout), dimension(0:,0:,0:), target, contiguous :: p
    if (any(c_or_f /= 'c')) error stop 'updt_rhs_b: only the cell-centred layout is supported'
    call fz_check(fz_updt_rhs_b(fz_ctx_ptr(), c_loc(p)), 'updt_rhs_b')
  end subroutine updt_rhs_b
end module mod_bound

module mod_acdi
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_acdi_set_epsilon, fz_acdi_set_gamma, fz_acdi_transport_pf, fz_acdi_cmpt_phi
  implicit none
  private
  public acdi_set_epsilon, acdi_set_gamma, acdi_transport_pf, acdi_cmpt_phi
contains
  ! src/acdi.f90:16-30 (the spacings are those of fz_setup)
  subroutine acdi_set_epsilon(dl, dzfi, seps_factor, seps)
    real(rp), intent(in), dimension(3) :: dl
    real(rp), intent(in), dimension(0:) :: dzfi
    real(rp), intent(in)  :: seps_factor
    real(rp), intent(out) :: seps
    call fz_check(fz_acdi_set_epsilon(fz_ctx_ptr(), seps_factor, seps), 'acdi_set_epsilon')
  end subroutine acdi_set_epsilon

  ! src/acdi.f90:32-62
  subroutine acdi_set_gamma(n, gam_factor, u, v, w, gam)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in) :: gam_factor
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous :: u, v, w
    real(rp), intent(out) :: gam
    call fz_check(fz_acdi_set_gamma(fz_ctx_ptr(), gam_factor, c_loc(u), c_loc(v), c_loc(w), gam), 'acdi_set_gamma')
  end subroutine acdi_set_gamma

  ! src/acdi.f90:64-225
  subroutine acdi_transport_pf(n, dli, dzci, dzfi, gam, seps, u, v, w, normx, normy, normz, phi, psi, dpsidt, &
                               flux_x, flux_y, flux_z)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(3) :: dli
    real(rp), intent(in), dimension(0:) :: dzci, dzfi
    real(rp), intent(in) :: gam, seps
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous :: u, v, w, normx, normy, normz, phi, psi
    real(rp), intent(out), dimension(:,:,:), target, contiguous :: dpsidt
    real(rp), intent(out), dimension(0:,0:,0:), target, contiguous :: flux_x, flux_y, flux_z
    call fz_check(fz_acdi_transport_pf(fz_ctx_ptr(), gam, seps, c_loc(u), c_loc(v), c_loc(w), c_loc(normx), &
                                       c_loc(normy), c_loc(normz), c_loc(phi), c_loc(psi), c_loc(dpsidt), &
                                       c_loc(flux_x), c_loc(flux_y), c_loc(flux_z)), 'acdi_transport_pf')
  end subroutine acdi_transport_pf

  ! src/acdi.f90:227-248
  subroutine acdi_cmpt_phi(n, seps, psi, phi)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in) :: seps
    real(rp), intent(in ), dimension(0:,0:,0:), target, contiguous :: psi
    real(rp), intent(out), dimension(0:,0:,0:), target, contiguous :: phi
    call fz_check(fz_acdi_cmpt_phi(fz_ctx_ptr(), seps, c_loc(psi), c_loc(phi)), 'acdi_cmpt_phi')
  end subroutine acdi_cmpt_phi
end module mod_acdi

module mod_vof_thinc_qq
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_vof_thinc_transport_psi, fz_vof_thinc_cmpt_phi
  implicit none
  private
  public vof_thinc_transport_psi, vof_thinc_cmpt_phi
contains
  ! src/vof_thinc_qq.f90:39-133
  subroutine vof_thinc_transport_psi(n, dli, dzfi, beta, u, v, w, normx, normy, normz, phi, psi, dpsidt, flux_x, &
                                     flux_y, flux_z)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(3) :: dli
    real(rp), intent(in), dimension(0:) :: dzfi
    real(rp), intent(in) :: beta
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous :: u, v, w, normx, normy, normz, phi, psi
    real(rp), intent(out), dimension(:,:,:), target, contiguous :: dpsidt
    real(rp), intent(out), dimension(0:,0:,0:), target, contiguous :: flux_x, flux_y, flux_z
    call fz_check(fz_vof_thinc_transport_psi(fz_ctx_ptr(), beta, c_loc(u), c_loc(v), c_loc(w), c_loc(normx), &
                                             c_loc(normy), c_loc(normz), c_loc(phi), c_loc(psi), c_loc(dpsidt), &
                                             c_loc(flux_x), c_loc(flux_y), c_loc(flux_z)), 'vof_thinc_transport_psi')
  end subroutine vof_thinc_transport_psi

  ! src/vof_thinc_qq.f90:135-157
  subroutine vof_thinc_cmpt_phi(n, beta, psi, phi)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in) :: beta
    real(rp), intent(in ), dimension(0:,0:,0:), target, contiguous :: psi
    real(rp), intent(out), dimension(0:,0:,0:), target, contiguous :: phi
    call fz_check(fz_vof_thinc_cmpt_phi(fz_ctx_ptr(), beta, c_loc(psi), c_loc(phi)), 'vof_thinc_cmpt_phi')
  end subroutine vof_thinc_cmpt_phi
end module mod_vof_thinc_qq

module mod_two_fluid
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_params_t, fz_init2fl, fz_clip_field, fz_cmpt_norm_curv_youngs
  implicit none
  private
  public init2fl, clip_field, cmpt_norm_curv_youngs
contains
  ! src/two_fluid.f90:247-437 with the spheres of 'spheres.in'
  ! (src/two_fluid.f90:439-471); 'cap-wav-1d' is not supported
  subroutine init2fl(inipsi, cbcpsi, seps, lo, hi, l, dl, zc_g, psi)
    character(len=*), intent(in) :: inipsi
    character(len=*), intent(in), dimension(0:1,3) :: cbcpsi
    real(rp), intent(in)                :: seps
    integer , intent(in), dimension(3)  :: lo, hi
    real(rp), intent(in), dimension(3)  :: l, dl
    real(rp), intent(in), dimension(0:) :: zc_g
    real(rp), dimension(lo(1)-1:,lo(2)-1:,lo(3)-1:), intent(out), target, contiguous :: psi
    type(fz_params_t) :: prm
    integer :: code, iunit, ierr, q
    logical :: exists
    character(len=1024) :: line
    real(rp) :: s4(4)
    select case(trim(inipsi))
    case('uni'); code = 0
    case('zer'); code = 1
    case('bub3'); code = 2
    case('bub2'); code = 3
    case('bub1'); code = 4
    case('drp3', 'dis3'); code = 5
    case('drp2', 'dis2'); code = 6
    case('drp1', 'dis1'); code = 7
    case('flm'); code = 8
    case default
      error stop 'init2fl: initial condition not supported'
    end select
    prm%ng = int(fz_ng(), c_int)
    prm%l = l
    prm%inipsi = int(code, c_int)
    prm%nspheres = 0
    prm%spheres = 0._rp
    do q = 0, 5
      prm%cbcpsi(q + 1) = cbcpsi(mod(q, 2), q/2 + 1)
    end do
    if (code >= 2 .and. code <= 7) then
      inquire(file='spheres.in', exist=exists)
      if (exists) then
        open(newunit=iunit, file='spheres.in', action='read', iostat=ierr)
        if (ierr /= 0) error stop 'Error reading input file spheres.in.'
        do
          read(iunit, '(A)', iostat=ierr) line
          if (ierr /= 0) exit
          if (len_trim(line) == 0) cycle
          read(line, *) s4
          if (prm%nspheres >= 64) error stop 'init2fl: at most 64 spheres are supported'
          prm%spheres(4*prm%nspheres + 1:4*prm%nspheres + 4) = s4
          prm%nspheres = prm%nspheres + 1
        end do
        close(iunit)
      end if
    end if
    call fz_check(fz_init2fl(fz_ctx_ptr(), prm, seps, c_loc(psi)), 'init2fl')
  end subroutine init2fl

  ! src/two_fluid.f90:35-58 (nh = 1)
  subroutine clip_field(nh, minmax, p)
    integer , intent(in   ), dimension(3) :: nh
    real(rp), intent(in   ), dimension(2) :: minmax
    real(rp), intent(inout), dimension(1-nh(1):,1-nh(2):,1-nh(3):), target, contiguous :: p
    if (any(nh /= 1)) error stop 'clip_field: only one halo layer is supported'
    call fz_check(fz_clip_field(fz_ctx_ptr(), minmax, c_loc(p)), 'clip_field')
  end subroutine clip_field

  ! src/two_fluid.f90:60-175
  subroutine cmpt_norm_curv_youngs(n, dli, dzci, dzfi, psi, normx, normy, normz, kappa)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(3) :: dli
    real(rp), intent(in), dimension(0:) :: dzci, dzfi
    real(rp), intent(in ), dimension(0:,0:,0:), target, contiguous :: psi
    real(rp), intent(out), dimension(0:,0:,0:), target, contiguous :: normx, normy, normz, kappa
    call fz_check(fz_cmpt_norm_curv_youngs(fz_ctx_ptr(), c_loc(psi), c_loc(normx), c_loc(normy), c_loc(normz), &
                                           c_loc(kappa)), 'cmpt_norm_curv_youngs')
  end subroutine cmpt_norm_curv_youngs
end module mod_two_fluid

module mod_mom
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_mom_xyz_ad, fz_mom_xyz_oth
  implicit none
  private
  public mom_xyz_ad, mom_xyz_oth
contains
  ! src/mom.f90:505-851
  subroutine mom_xyz_ad(n, dli, dzci, dzfi, rho12, mu12, psiflx_x, psiflx_y, psiflx_z, u, v, w, psi, dudt, dvdt, &
                        dwdt)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(3) :: dli
    real(rp), intent(in), dimension(0:) :: dzci, dzfi
    real(rp), intent(in), dimension(2) :: rho12, mu12
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous :: psiflx_x, psiflx_y, psiflx_z, u, v, w, psi
    real(rp), intent(out), dimension(:,:,:), target, contiguous :: dudt, dvdt, dwdt
    call fz_check(fz_mom_xyz_ad(fz_ctx_ptr(), rho12, mu12, c_loc(psiflx_x), c_loc(psiflx_y), c_loc(psiflx_z), &
                                c_loc(u), c_loc(v), c_loc(w), c_loc(psi), c_loc(dudt), c_loc(dvdt), c_loc(dwdt)), &
                  'mom_xyz_ad')
  end subroutine mom_xyz_ad

  ! src/mom.f90:853-1039 (without the scalar s)
  subroutine mom_xyz_oth(n, dli, dzci, dzfi, dt_r, rho12, beta12, bforce, gacc, sigma, rho0, rho_av, p, psi, kappa, &
                         s, pn, po, dudt, dvdt, dwdt)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(3) :: dli
    real(rp), intent(in), dimension(0:) :: dzci, dzfi
    real(rp), intent(in) :: dt_r
    real(rp), intent(in), dimension(2) :: rho12, beta12
    real(rp), intent(in), dimension(3) :: bforce, gacc
    real(rp), intent(in) :: sigma
    real(rp), intent(in) :: rho0, rho_av(3)
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous :: p, psi, kappa
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous, optional :: s, pn, po
    real(rp), intent(out), dimension(:,:,:), target, contiguous :: dudt, dvdt, dwdt
    if (present(s)) error stop 'mom_xyz_oth: the scalar equation is not supported'
    if (.not. (present(pn) .and. present(po))) error stop 'mom_xyz_oth: pn and po are required'
    call fz_check(fz_mom_xyz_oth(fz_ctx_ptr(), dt_r, rho12, beta12, bforce, gacc, sigma, rho0, rho_av, c_loc(p), &
                                 c_loc(psi), c_loc(kappa), c_loc(pn), c_loc(po), c_loc(dudt), c_loc(dvdt), &
                                 c_loc(dwdt)), 'mom_xyz_oth')
  end subroutine mom_xyz_oth
end module mod_mom

module mod_rk
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_rk_state_create, fz_rk, fz_rk_2fl
  implicit none
  private
  public rk, rk_2fl
  ! the AB2 histories the reference keeps in `save` arrays (src/rk.f90:44-70, :214-228)
  type(c_ptr), save :: st_mom = c_null_ptr, st_psi = c_null_ptr
contains
  ! src/rk.f90:16-131
  subroutine rk(rkpar, n, dli, dzci, dzfi, dt, dt_r, bforce, gacc, sigma, rho_av, rho12, mu12, beta12, rho0, psi, &
                kappa, p, pn, po, s, psio, psiflx_x, psiflx_y, psiflx_z, u, v, w)
    real(rp), intent(in), dimension(2) :: rkpar
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(3) :: dli
    real(rp), intent(in) :: dt, dt_r
    real(rp), intent(in   ), dimension(0:) :: dzci, dzfi
    real(rp), intent(in   ), dimension(3)  :: bforce, gacc
    real(rp), intent(in   )                :: sigma, rho_av(3)
    real(rp), intent(in   ), dimension(2)  :: rho12, mu12, beta12
    real(rp), intent(in   )                :: rho0
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous :: psi, kappa, p
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous, optional :: pn, po, s
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous, optional :: psio, psiflx_x, psiflx_y, psiflx_z
    real(rp), intent(inout), dimension(0:,0:,0:), target, contiguous :: u, v, w
    if (present(s)) error stop 'rk: the scalar equation is not supported'
    if (.not. (present(pn) .and. present(po) .and. present(psio) .and. present(psiflx_x) .and. present(psiflx_y) &
               .and. present(psiflx_z))) error stop 'rk: the two-fluid arguments are required'
    if (.not. c_associated(st_mom)) call fz_check(fz_rk_state_create(fz_ctx_ptr(), st_mom), 'rk')
    call fz_check(fz_rk(fz_ctx_ptr(), st_mom, rkpar, dt, dt_r, bforce, gacc, sigma, rho_av, rho12, mu12, beta12, &
                        rho0, c_loc(psi), c_loc(kappa), c_loc(p), c_loc(pn), c_loc(po), c_loc(psio), c_loc(psiflx_x), &
                        c_loc(psiflx_y), c_loc(psiflx_z), c_loc(u), c_loc(v), c_loc(w)), 'rk')
  end subroutine rk

  ! src/rk.f90:193-254 (ACDI, or THINC/QQ in a build with _INTERFACE_CAPTURING_VOF)
  subroutine rk_2fl(rkpar, n, dli, dzci, dzfi, dt, gam, seps, beta, u, v, w, normx, normy, normz, phi, psi, psiflx_x, &
                    psiflx_y, psiflx_z)
    real(rp), intent(in   ), dimension(2) :: rkpar
    integer , intent(in   ), dimension(3) :: n
    real(rp), intent(in   ), dimension(3) :: dli
    real(rp), intent(in   ), dimension(0:) :: dzci, dzfi
    real(rp), intent(in   ) :: gam, seps, beta, dt
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous :: u, v, w
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous :: normx, normy, normz
    real(rp), intent(in   ), dimension(0:,0:,0:), target, contiguous, optional :: phi
    real(rp), intent(inout), dimension(0:,0:,0:), target, contiguous :: psi
    real(rp), intent(out  ), dimension(0:,0:,0:), target, contiguous :: psiflx_x, psiflx_y, psiflx_z
    if (.not. present(phi)) error stop 'rk_2fl: phi is required'
    if (.not. c_associated(st_psi)) call fz_check(fz_rk_state_create(fz_ctx_ptr(), st_psi), 'rk_2fl')
    call fz_check(fz_rk_2fl(fz_ctx_ptr(), st_psi, rkpar, dt, gam, seps, beta, vof_build, c_loc(u), c_loc(v), &
                            c_loc(w), c_loc(normx), c_loc(normy), c_loc(normz), c_loc(phi), c_loc(psi), &
                            c_loc(psiflx_x), c_loc(psiflx_y), c_loc(psiflx_z)), 'rk_2fl')
  end subroutine rk_2fl
end module mod_rk

module mod_chkdiv
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_chkdiv
  implicit none
  private
  public chkdiv
contains
  ! src/chkdiv.f90:15-49
  subroutine chkdiv(lo, hi, dli, dzfi, u, v, w, divtot, divmax)
    integer , intent(in), dimension(3) :: lo, hi
    real(rp), intent(in), dimension(3) :: dli
    real(rp), intent(in), dimension(lo(3)-1:) :: dzfi
    real(rp), intent(in), dimension(lo(1)-1:,lo(2)-1:,lo(3)-1:), target, contiguous :: u, v, w
    real(rp), intent(out) :: divtot, divmax
    call fz_check(fz_chkdiv(fz_ctx_ptr(), c_loc(u), c_loc(v), c_loc(w), divtot, divmax), 'chkdiv')
  end subroutine chkdiv
end module mod_chkdiv

module mod_chkdt
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_chkdt
  implicit none
  private
  public chkdt
contains
  ! src/chkdt.f90:15-94
  subroutine chkdt(n, dl, dzci, dzfi, is_solve_ns, is_track_interface, mu12, rho12, sigma, gacc, u, v, w, dtmax, gam, &
                   seps, ka12, cp12)
    integer , intent(in), dimension(3) :: n
    real(rp), intent(in), dimension(3) :: dl
    real(rp), intent(in), dimension(0:) :: dzci, dzfi
    logical,  intent(in) :: is_solve_ns, is_track_interface
    real(rp), intent(in) :: mu12(2), rho12(2), sigma, gacc(3), gam, seps
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous :: u, v, w
    real(rp), intent(out) :: dtmax
    real(rp), intent(in ), optional :: ka12(2), cp12(2)
    call fz_check(fz_chkdt(fz_ctx_ptr(), merge(1_c_int, 0_c_int, is_solve_ns), &
                           merge(1_c_int, 0_c_int, is_track_interface), vof_build, mu12, rho12, sigma, gacc, &
                           c_loc(u), c_loc(v), c_loc(w), gam, seps, dtmax), 'chkdt')
  end subroutine chkdt
end module mod_chkdt

module mod_forcing
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_lscale_forcing
  implicit none
  private
  public lscale_forcing
contains
  ! src/forcing.f90:13-65
  subroutine lscale_forcing(ftype, lo, hi, alpha, dt, l, dl, zc, zf, u, v, w)
    integer , intent(in) :: ftype
    integer , intent(in), dimension(3) :: lo, hi
    real(rp), intent(in) :: alpha, dt
    real(rp), intent(in), dimension(3) :: l, dl
    real(rp), intent(in), dimension(lo(3)-1:) :: zc, zf
    real(rp), intent(inout), dimension(lo(1)-1:,lo(2)-1:,lo(3)-1:), target, contiguous :: u, v, w
    call fz_check(fz_lscale_forcing(fz_ctx_ptr(), int(ftype, c_int), alpha, dt, c_loc(u), c_loc(v), c_loc(w)), &
                  'lscale_forcing')
  end subroutine lscale_forcing
end module mod_forcing

module mod_utils
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_bulk_mean_12_stag
  implicit none
  private
  public bulk_mean_12_stag, swap
contains
  ! src/utils.f90:68-136 (grid_vol_ratio is the one fz_setup's grid gives)
  subroutine bulk_mean_12_stag(n, idir, grid_vol_ratio, psi, p12, mean)
    integer , intent(in), dimension(3) :: n
    integer , intent(in)               :: idir
    real(rp), intent(in), dimension(0:) :: grid_vol_ratio
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous :: psi
    real(rp), intent(in), dimension(2)        :: p12
    real(rp), intent(out) :: mean
    call fz_check(fz_bulk_mean_12_stag(fz_ctx_ptr(), int(idir, c_int), c_loc(psi), p12, mean), 'bulk_mean_12_stag')
  end subroutine bulk_mean_12_stag

  ! src/utils.f90:145-153
  subroutine swap(arr1, arr2)
    real(rp), intent(inout), pointer, contiguous, dimension(:,:,:) :: arr1, arr2
    real(rp), pointer, contiguous, dimension(:,:,:) :: tmp
    tmp  => arr1
    arr1 => arr2
    arr2 => tmp
  end subroutine swap
end module mod_utils

module mod_load
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_load_one
  implicit none
  private
  public load_one
contains
  ! src/load.f90:19-116: one restart file (the interior, then [time, istep]);
  ! comm is accepted for the interface and unused (one device holds the field)
  subroutine load_one(io, filename, comm, ng, nh, lo, hi, p, time, istep)
    character(len=1), intent(in) :: io
    character(len=*), intent(in) :: filename
    integer         , intent(in) :: comm
    integer , intent(in), dimension(3) :: ng, nh, lo, hi
    real(rp), intent(inout), dimension(lo(1)-nh(1):,lo(2)-nh(2):,lo(3)-nh(3):), target, contiguous :: p
    real(rp), intent(inout) :: time
    integer , intent(inout) :: istep
    integer(c_int) :: is
    if (any(ng /= fz_ng()) .or. any(nh /= 1)) error stop 'load_one: grid or halo differs from fz_setup'
    is = int(istep, c_int)
    call fz_check(fz_load_one(fz_ctx_ptr(), io, trim(filename)//c_null_char, c_loc(p), time, is), 'load_one')
    istep = int(is)
  end subroutine load_one
end module mod_load

module mod_output
  use, intrinsic :: iso_c_binding
  use mod_types
  use mod_fz
  use mod_fz_capi, only: fz_out3d
  implicit none
  private
  public out3d, write_log_output, write_visu_3d
contains
  ! src/output.f90:185-236: the interior every nskip points, no header; p is
  ! the halo'd device field
  subroutine out3d(fname, nskip, p)
    character(len=*), intent(in) :: fname
    integer , intent(in), dimension(3) :: nskip
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous :: p
    call fz_check(fz_out3d(fz_ctx_ptr(), trim(fname)//c_null_char, int(nskip, c_int), c_loc(p)), 'out3d')
  end subroutine out3d

  ! src/output.f90:238-266: one line per saved file, for the xdmf generator
  subroutine write_log_output(fname, fname_fld, varname, nmin, nmax, nskip, time, istep)
    character(len=*), intent(in) :: fname, fname_fld, varname
    integer , intent(in), dimension(3) :: nmin, nmax, nskip
    real(rp), intent(in)               :: time
    integer , intent(in)               :: istep
    integer :: iunit
    open(newunit=iunit, file=fname, position='append')
    write(iunit, '(A,A,A,9i5,E16.7e3,i8)') trim(fname_fld), ' ', trim(varname), nmin, nmax, nskip, time, istep
    close(iunit)
  end subroutine write_log_output

  ! src/output.f90:268-281; p is the whole halo'd device field (the
  ! reference passes the interior section of its host copy)
  subroutine write_visu_3d(datadir, fname_bin, fname_log, varname, nmin, nmax, nskip, time, istep, p)
    character(len=*), intent(in)       :: datadir, fname_bin, fname_log, varname
    integer , intent(in), dimension(3) :: nmin, nmax, nskip
    real(rp), intent(in)               :: time
    integer , intent(in)               :: istep
    real(rp), intent(in), dimension(0:,0:,0:), target, contiguous :: p
    call out3d(trim(datadir)//trim(fname_bin), nskip, p)
    call write_log_output(trim(datadir)//trim(fname_log), trim(fname_bin), trim(varname), nmin, nmax, nskip, time, &
                          istep)
  end subroutine write_visu_3d
end module mod_output
