! crt_hip_mod.f90 -- ISO_C_BINDING interface of libcrtomo_hip.so (include/crtomo_hip.h)
! for CRTomo's Fortran host, and the library context the shims share.
!
! One context serves every shim of a process.  The context is created on first use
! (crt_hip_ctx, under an OpenMP critical section, so the shims may be called from the
! reference's OpenMP loops); the library serialises the single-system drop-ins
! (crt_chol, crt_vre and their DC twins) on one context, so several threads may call
! them at once.  Every C entry returns the reference's errnr code (0 = ok); pointers
! that the C side accepts as NULL are passed as type(c_ptr), value.
module crt_hip_mod
  use iso_c_binding
  implicit none
  private
  public :: crt_hip_ctx
  type(c_ptr), save :: ctx_h = c_null_ptr

  interface
    integer(c_int) function crt_create(ctx, device) bind(C, name='crt_create')
      import :: c_ptr, c_int
      type(c_ptr) :: ctx
      integer(c_int), value :: device
    end function
    integer(c_int) function crt_set_grid(ctx, sanz, mb, typanz, typ, nelanz, selanz, maxn, nrel_em, x, y, &
         rnr, eanz, enr) bind(C, name='crt_set_grid')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: sanz, mb, typanz, maxn, eanz
      integer(c_int) :: typ(*), nelanz(*), selanz(*), nrel_em(*), rnr(*), enr(*)
      real(c_double) :: x(*), y(*)
    end function
    integer(c_int) function crt_set_survey(ctx, nanz, strnr, vnr) bind(C, name='crt_set_survey')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nanz
      integer(c_int) :: strnr(*), vnr(*)
    end function
    integer(c_int) function crt_set_wavenumbers(ctx, kwnanz, kwn, kwnwi) bind(C, name='crt_set_wavenumbers')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kwnanz
      real(c_double) :: kwn(*), kwnwi(*)
    end function
    integer(c_int) function crt_precal(ctx, lsytop, lbeta) bind(C, name='crt_precal')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: lsytop
      integer(c_int) :: lbeta
    end function
    integer(c_int) function crt_set_sigma(ctx, sigma) bind(C, name='crt_set_sigma')
      import :: c_ptr, c_double_complex
      type(c_ptr), value :: ctx
      complex(c_double_complex) :: sigma(*)
    end function
    integer(c_int) function crt_set_fem_switches(ctx, lsr, lana, nsink) bind(C, name='crt_set_fem_switches')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: lsr, lana, nsink
    end function
    integer(c_int) function crt_bkfak(ctx, kfak) bind(C, name='crt_bkfak')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: kfak(*)
    end function
    integer(c_int) function crt_set_inversion_kind(ctx, kind) bind(C, name='crt_set_inversion_kind')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: kind
    end function
    integer(c_int) function crt_forward(ctx) bind(C, name='crt_forward')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function crt_get_kpot(ctx, kpot) bind(C, name='crt_get_kpot')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      type(*), dimension(*) :: kpot         ! kpot (COMPLEX*16) or kpotdc (REAL*8)
    end function
    integer(c_int) function crt_voltages(ctx, mode, kfak, volt, sigmaa) bind(C, name='crt_voltages')
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      type(c_ptr), value :: kfak
      complex(c_double_complex) :: volt(*), sigmaa(*)
    end function
    integer(c_int) function crt_get_hpot(ctx, hpot) bind(C, name='crt_get_hpot')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      type(*), dimension(*) :: hpot         ! hpot or hpotdc
    end function
    integer(c_int) function crt_sens(ctx, mode, strom) bind(C, name='crt_sens')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: mode
      type(c_ptr), value :: strom
    end function
    integer(c_int) function crt_get_sens(ctx, sens) bind(C, name='crt_get_sens')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      type(*), dimension(*) :: sens         ! sens or sensdc
    end function
    integer(c_int) function crt_set_sens(ctx, nanz, manz, sens) bind(C, name='crt_set_sens')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: nanz, manz
      type(*), dimension(*) :: sens
    end function
    integer(c_int) function crt_set_regularization(ctx, ltri, nx, manz, maxn, smatm, nachbar) &
         bind(C, name='crt_set_regularization')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ltri, nx, manz, maxn
      real(c_double) :: smatm(*)
      type(c_ptr), value :: nachbar
    end function
    integer(c_int) function crt_bsmatm(ctx, ltri, nx, nz, alfx, alfz, smatm) bind(C, name='crt_bsmatm')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: ltri, nx, nz
      real(c_double), value :: alfx, alfz
      type(c_ptr), value :: smatm
    end function
    integer(c_int) function crt_bnachbar(elanz, maxn, nrel_em, nachbar) bind(C, name='crt_bnachbar')
      import :: c_int
      integer(c_int), value :: elanz, maxn
      integer(c_int) :: nrel_em(*), nachbar(*)
    end function
    integer(c_int) function crt_blam0(ctx, fpi, wmatd, wdfak, alfx, alfz, lammax) bind(C, name='crt_blam0')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: fpi
      real(c_double) :: wmatd(*)
      integer(c_int) :: wdfak(*)
      real(c_double), value :: alfx, alfz
      real(c_double) :: lammax
    end function
    integer(c_int) function crt_cjg(ctx, lam, eps, ncgmax, wmatd, wdfak, cgfac, bvec, dpar, cgres, ncg) &
         bind(C, name='crt_cjg')
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), value :: lam, eps
      integer(c_int), value :: ncgmax
      real(c_double) :: wmatd(*), cgfac(*), cgres(*)
      integer(c_int) :: wdfak(*), ncg
      complex(c_double_complex) :: bvec(*), dpar(*)
    end function
    integer(c_int) function crt_cjg_noalloc(ctx, lam, eps, ncgmax, wmatd, wdfak, cgfac, bvec, dpar, cgres, ncg) &
         bind(C, name='crt_cjg_noalloc')
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      real(c_double), value :: lam, eps
      integer(c_int), value :: ncgmax
      real(c_double) :: wmatd(*), cgfac(*), cgres(*)
      integer(c_int) :: wdfak(*), ncg
      complex(c_double_complex) :: bvec(*), dpar(*)
    end function
    integer(c_int) function crt_chol(ctx, sanz, mb, a) bind(C, name='crt_chol')
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: sanz, mb
      complex(c_double_complex) :: a(*)
    end function
    integer(c_int) function crt_vre(ctx, sanz, mb, a, b, pot) bind(C, name='crt_vre')
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: sanz, mb
      complex(c_double_complex) :: a(*), b(*), pot(*)
    end function
    integer(c_int) function crt_choldc(ctx, sanz, mb, a) bind(C, name='crt_choldc')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: sanz, mb
      real(c_double) :: a(*)
    end function
    integer(c_int) function crt_vredc(ctx, sanz, mb, a, b, pot) bind(C, name='crt_vredc')
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: sanz, mb
      real(c_double) :: a(*), b(*)
      complex(c_double_complex) :: pot(*)
    end function
    integer(c_int) function crt_cholz(ctx, n, a, p) bind(C, name='crt_cholz')
      import :: c_ptr, c_int, c_double_complex
      type(c_ptr), value :: ctx
      integer(c_int), value :: n
      complex(c_double_complex) :: a(*), p(*)
    end function
  end interface
  public :: crt_create, crt_set_grid, crt_set_survey, crt_set_wavenumbers, crt_precal, crt_set_sigma, &
       crt_set_fem_switches, crt_bkfak, &
       crt_set_inversion_kind, crt_forward, crt_get_kpot, crt_voltages, crt_get_hpot, crt_sens, crt_get_sens, &
       crt_set_sens, crt_set_regularization, crt_bsmatm, crt_bnachbar, crt_blam0, crt_cjg, crt_cjg_noalloc, &
       crt_chol, crt_vre, crt_choldc, crt_vredc, crt_cholz

contains

  ! The process's library context on device 0, created by the first caller.
  type(c_ptr) function crt_hip_ctx()
    integer(c_int) :: st
    !$omp critical (crt_hip_ctx_create)
    if (.not. c_associated(ctx_h)) then
       st = crt_create(ctx_h, 0_c_int)
       if (st /= 0) ctx_h = c_null_ptr
    end if
    !$omp end critical (crt_hip_ctx_create)
    crt_hip_ctx = ctx_h
  end function crt_hip_ctx

end module crt_hip_mod
