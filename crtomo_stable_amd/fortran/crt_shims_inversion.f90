! crt_shims_inversion.f90 -- CRTomo's no-argument routines of the voltage, Jacobian
! and normal-equation steps on the GPU through libcrtomo_hip.so, reading and writing
! the same module data as the routines they replace:
!   bvolt / bvolti  (src/bvolt.f90:7, src/bvolti.f90:8; rtrafo folded in)
!   bsens / bsensi / bsendc  (src/bsens.f90:23, src/bsensi.f90:23, src/bsendc.f90:1)
!   bnachbar  (src/bnachbar.f90:7), blam0  (src/blam0.f90:12)
! and three entries for the steps that are not separate external routines there:
!   crt_hip_setup    after relem/relectr/rdatm/rwaven: grid, survey, wavenumbers and
!                    the device precal (src/precal.f90:13)
!   crt_hip_forward  in place of the forward loop (src/fem.f90:328-397,
!                    src/inv.f90:355-484): every (electrode, wavenumber) system of the
!                    current sigma on the device; kpot stays there (and is copied into
!                    kpot / kpotdc when the host has them allocated)
!   crt_hip_cjg      the body of cg_mod's cjg (src/cg_mod.f90:84): cjggra, or cjggdc
!                    for ldc / lfpi, on the resident Jacobian
!   crt_hip_bsmatm   bsmatm (src/bsmatm_mod.f90:62) for ltri 0, 1, 2 on the device
! The library context is crt_hip_mod's.  The Jacobian stays on the device for the CG;
! the bsens* shims copy it into sens / sensdc only when the host has it allocated.
! bsens assumes one model cell per area element (mnr the identity).

subroutine crt_hip_setup(lsytop, lbeta_out)
  use iso_c_binding
  use crt_hip_mod
  use elemmod, only: sanz, mb, typanz, typ, nelanz, selanz, nrel, sx, sy, snr, rnr, max_nr_element_nodes, relanz
  use electrmod, only: eanz, enr
  use datmod, only: nanz, strnr, vnr
  use wavenmod, only: kwnanz, kwn, kwnwi
  use femmod, only: lsr
  use elemmod, only: lsink, nsink
  use errmod, only: errnr, fetxt
  implicit none
  logical, intent(in) :: lsytop
  logical, intent(out) :: lbeta_out
  integer(c_int), allocatable :: nrel_em(:), rnr_c(:)
  real(c_double), allocatable :: xn(:), yn(:)
  integer(c_int) :: lb
  integer :: i
  type(c_ptr) :: ctx
  ctx = crt_hip_ctx()
  allocate(nrel_em(size(nrel)), xn(sanz), yn(sanz), rnr_c(max(relanz, 1)))
  nrel_em = reshape(transpose(nrel), [size(nrel)])      ! element-major
  do i = 1, sanz
     xn(i) = sx(snr(i))
     yn(i) = sy(snr(i))
  end do
  rnr_c = 0
  if (relanz > 0) rnr_c(1:relanz) = rnr(1:relanz)
  errnr = crt_set_grid(ctx, sanz, mb, typanz, typ, nelanz, selanz, max_nr_element_nodes, nrel_em, xn, yn, &
       rnr_c, eanz, enr)
  if (errnr == 0) errnr = crt_set_survey(ctx, nanz, strnr, vnr)
  if (errnr == 0) errnr = crt_set_wavenumbers(ctx, kwnanz, kwn, kwnwi)
  lb = 0
  if (errnr == 0) errnr = crt_precal(ctx, merge(1_c_int, 0_c_int, lsytop), lb)
  ! the singularity removal and the fictitious sink of femmod / elemmod (lsr is off in
  ! the inversion, src/rall.f90:153; the device applies it only with mixed boundaries)
  if (errnr == 0) errnr = crt_set_fem_switches(ctx, merge(1_c_int, 0_c_int, lsr), 0_c_int, &
       merge(int(nsink, c_int), 0_c_int, lsink))
  lbeta_out = lb /= 0
  if (errnr /= 0) fetxt = 'crt_hip_setup (libcrtomo_hip)'
end subroutine crt_hip_setup

! bkfak (src/bkfak.f90:48-136): the half-space geometric factors into datmod's kfak on
! the device, and tmp.kfak as the reference writes it
subroutine bkfak()
  use iso_c_binding
  use crt_hip_mod
  use datmod, only: nanz, kfak
  use errmod, only: errnr, fetxt
  implicit none
  real(c_double), allocatable :: kf(:)
  integer :: fp, i
  allocate(kf(nanz))
  errnr = crt_bkfak(crt_hip_ctx(), kf)
  if (errnr /= 0) then
     fetxt = 'bkfak (libcrtomo_hip)'
     return
  end if
  kfak(1:nanz) = kf
  open (newunit=fp, file='tmp.kfak', status='replace')
  write (fp, '(I8,2X,G12.4)') (i, kfak(i), i = 1, nanz)
  close (fp)
end subroutine bkfak

subroutine crt_hip_forward()
  use iso_c_binding
  use crt_hip_mod
  use sigmamod, only: sigma
  use femmod, only: ldc
  use invmod, only: lfpi
  use alloci, only: kpot, kpotdc
  use errmod, only: errnr, fetxt
  implicit none
  integer(c_int) :: kind
  integer(c_int), save :: last_kind = 0
  type(c_ptr) :: ctx
  ctx = crt_hip_ctx()
  kind = 0
  if (lfpi) kind = 2
  if (ldc) kind = 1
  errnr = 0
  if (kind /= last_kind) then
     errnr = crt_set_inversion_kind(ctx, kind)
     if (errnr == 0) last_kind = kind
  end if
  if (errnr == 0) errnr = crt_set_sigma(ctx, sigma)
  if (errnr == 0) errnr = crt_forward(ctx)
  if (errnr == 0 .and. kind == 1 .and. allocated(kpotdc)) errnr = crt_get_kpot(ctx, kpotdc)
  if (errnr == 0 .and. kind /= 1 .and. allocated(kpot)) errnr = crt_get_kpot(ctx, kpot)
  if (errnr /= 0) fetxt = 'crt_hip_forward (libcrtomo_hip)'
end subroutine crt_hip_forward

subroutine bvolt()
  use iso_c_binding
  use crt_hip_mod
  use datmod, only: nanz, kfak, volt, sigmaa
  use errmod, only: errnr, fetxt
  implicit none
  real(c_double), allocatable, target :: kf(:)
  allocate(kf(nanz))
  kf = 1d0
  if (allocated(kfak)) kf = kfak(1:nanz)
  errnr = crt_voltages(crt_hip_ctx(), 0_c_int, c_loc(kf), volt, sigmaa)
  if (errnr /= 0) fetxt = 'bvolt (libcrtomo_hip)'
end subroutine bvolt

subroutine bvolti()
  use iso_c_binding
  use crt_hip_mod
  use datmod, only: volt, sigmaa
  use errmod, only: errnr, fetxt
  implicit none
  errnr = crt_voltages(crt_hip_ctx(), 1_c_int, c_null_ptr, volt, sigmaa)
  if (errnr /= 0) fetxt = 'bvolti (libcrtomo_hip)'
end subroutine bvolti

subroutine bsens()
  use iso_c_binding
  use crt_hip_mod
  use datmod, only: nanz, strom
  use alloci, only: sens
  use errmod, only: errnr, fetxt
  implicit none
  real(c_double), allocatable, target :: st(:)
  allocate(st(nanz))
  st = 1d0
  if (allocated(strom)) st = strom(1:nanz)
  errnr = crt_sens(crt_hip_ctx(), 0_c_int, c_loc(st))
  if (errnr == 0 .and. allocated(sens)) errnr = crt_get_sens(crt_hip_ctx(), sens)
  if (errnr /= 0) fetxt = 'bsens (libcrtomo_hip)'
end subroutine bsens

subroutine bsensi(tictoc)
  use iso_c_binding
  use crt_hip_mod
  use alloci, only: sens
  use errmod, only: errnr, fetxt
  implicit none
  logical :: tictoc
  errnr = crt_sens(crt_hip_ctx(), 1_c_int, c_null_ptr)
  if (errnr == 0 .and. allocated(sens)) errnr = crt_get_sens(crt_hip_ctx(), sens)
  if (errnr /= 0) fetxt = 'bsensi (libcrtomo_hip)'
end subroutine bsensi

subroutine bsendc(tictoc)
  use iso_c_binding
  use crt_hip_mod
  use alloci, only: sensdc
  use errmod, only: errnr, fetxt
  implicit none
  logical :: tictoc
  errnr = crt_sens(crt_hip_ctx(), 1_c_int, c_null_ptr)
  if (errnr == 0 .and. allocated(sensdc)) errnr = crt_get_sens(crt_hip_ctx(), sensdc)
  if (errnr /= 0) fetxt = 'bsendc (libcrtomo_hip)'
end subroutine bsendc

subroutine bnachbar()
  use iso_c_binding
  use crt_hip_mod
  use elemmod, only: nrel, elanz, max_nr_element_nodes, nachbar
  use errmod, only: errnr, fetxt
  implicit none
  integer(c_int), allocatable :: nrel_e(:), nb(:)
  integer :: m
  m = max_nr_element_nodes
  allocate(nrel_e(elanz * m), nb(elanz * (m + 1)))
  nrel_e = reshape(transpose(nrel(1:elanz, 1:m)), [elanz * m])
  errnr = crt_bnachbar(elanz, m, nrel_e, nb)
  if (errnr /= 0) then
     fetxt = 'bnachbar (libcrtomo_hip)'
     return
  end if
  if (allocated(nachbar)) deallocate(nachbar)
  allocate(nachbar(elanz, m + 1))
  nachbar = reshape(nb, [elanz, m + 1])
end subroutine bnachbar

subroutine blam0()
  use iso_c_binding
  use crt_hip_mod
  use datmod, only: nanz
  use modelmod, only: manz
  use invmod, only: lfpi, wmatd, wdfak
  use konvmod, only: lammax, llamf, lamnull_cri, lamnull_fpi, nz, alfx, alfz
  use errmod, only: errnr, fetxt
  implicit none
  ! the preset branches of src/blam0.f90:47-64 as they are
  if ((btest(llamf, 0) .or. (lamnull_cri > epsilon(lamnull_cri))) .and. .not. lfpi) then
     if (nz == -1) then
        lammax = max(real(manz), real(nanz))
     else
        lammax = dble(lamnull_cri)
     end if
     return
  else if (btest(llamf, 0) .or. (lamnull_fpi > epsilon(lamnull_fpi))) then
     lammax = dble(lamnull_fpi)
     return
  end if
  errnr = crt_blam0(crt_hip_ctx(), merge(1_c_int, 0_c_int, lfpi), wmatd, wdfak, alfx, alfz, lammax)
  if (errnr /= 0) fetxt = 'blam0 (libcrtomo_hip)'
end subroutine blam0

subroutine crt_hip_cjg()
  use iso_c_binding
  use crt_hip_mod
  use konvmod, only: lam
  use cjgmod, only: eps, ncgmax, ncg, cgres, cgfac, bvec
  use invmod, only: wmatd, wdfak, dpar
  use errmod, only: errnr, fetxt
  implicit none
  integer(c_int) :: n
  errnr = crt_cjg(crt_hip_ctx(), lam, eps, ncgmax, wmatd, wdfak, cgfac, bvec, dpar, cgres, n)
  ncg = n
  if (errnr /= 0) fetxt = 'cjg (libcrtomo_hip)'
end subroutine crt_hip_cjg

subroutine crt_hip_bsmatm()
  use iso_c_binding
  use crt_hip_mod
  use konvmod, only: ltri, nx, nz, alfx, alfz
  use alloci, only: smatm
  use errmod, only: errnr, fetxt
  implicit none
  real(c_double), allocatable, target :: sm(:)
  integer :: ncol
  if (.not. allocated(smatm)) then
     errnr = crt_bsmatm(crt_hip_ctx(), ltri, nx, nz, alfx, alfz, c_null_ptr)
  else
     ncol = size(smatm, 2)
     allocate(sm(size(smatm)))
     errnr = crt_bsmatm(crt_hip_ctx(), ltri, nx, nz, alfx, alfz, c_loc(sm))
     if (errnr == 0) smatm = reshape(sm, [size(smatm, 1), ncol])
  end if
  if (errnr /= 0) fetxt = 'bsmatm (libcrtomo_hip)'
end subroutine crt_hip_bsmatm
