! crt_shims_forward.f90 -- the reference's single-system solver routines, same names
! and arguments, running on the GPU through libcrtomo_hip.so.  They replace
! src/chol.f90, src/vre.f90, src/choldc.f90, src/vredc.f90 and src/cholz.f90 in a
! CRTomo build; their callers (the forward loops of src/fem.f90:328-397 and
! src/inv.f90:355-484) stay unchanged, including their OpenMP loop over wavenumbers:
! the calls share one library context, which runs one system at a time.
! The band layout, sizes (elemmod: sanz, mb) and error convention (errmod: errnr,
! fetxt) are the reference's.

subroutine chol(a_chol)                       ! src/chol.f90:7
  use iso_c_binding
  use crt_hip_mod
  use elemmod, only: sanz, mb
  use errmod, only: errnr, fetxt
  implicit none
  complex(kind(0d0)), dimension(*) :: a_chol
  integer(c_int) :: st
  st = crt_chol(crt_hip_ctx(), int(sanz, c_int), int(mb, c_int), a_chol)
  if (st /= 0) fetxt = 'chol (libcrtomo_hip)'
  errnr = st
end subroutine chol

subroutine vre(a_vre, b_vre, pot_vre)         ! src/vre.f90:5
  use iso_c_binding
  use crt_hip_mod
  use elemmod, only: sanz, mb
  use errmod, only: errnr, fetxt
  implicit none
  complex(kind(0d0)), dimension(*) :: a_vre, b_vre, pot_vre
  integer(c_int) :: st
  st = crt_vre(crt_hip_ctx(), int(sanz, c_int), int(mb, c_int), a_vre, b_vre, pot_vre)
  if (st /= 0) then
     fetxt = 'vre (libcrtomo_hip)'
     errnr = st
  end if
end subroutine vre

subroutine choldc(a_chol)                     ! src/choldc.f90:7
  use iso_c_binding
  use crt_hip_mod
  use elemmod, only: sanz, mb
  use errmod, only: errnr, fetxt
  implicit none
  real(kind(0d0)), dimension(*) :: a_chol
  integer(c_int) :: st
  st = crt_choldc(crt_hip_ctx(), int(sanz, c_int), int(mb, c_int), a_chol)
  if (st /= 0) fetxt = 'choldc (libcrtomo_hip)'
  errnr = st
end subroutine choldc

subroutine vredc(a_vre, b_vre, pot_vre)       ! src/vredc.f90:1
  use iso_c_binding
  use crt_hip_mod
  use elemmod, only: sanz, mb
  use errmod, only: errnr, fetxt
  implicit none
  real(kind(0d0)), dimension(*) :: a_vre, b_vre
  complex(kind(0d0)), dimension(*) :: pot_vre
  integer(c_int) :: st
  st = crt_vredc(crt_hip_ctx(), int(sanz, c_int), int(mb, c_int), a_vre, b_vre, pot_vre)
  if (st /= 0) then
     fetxt = 'vredc (libcrtomo_hip)'
     errnr = st
  end if
end subroutine vredc

subroutine cholz(a, p, n, ierr, lverb)        ! src/cholz.f90:1
  use iso_c_binding
  use crt_hip_mod
  implicit none
  integer, intent(in) :: n
  complex(kind(0d0)), dimension(n, n), intent(inout) :: a
  complex(kind(0d0)), dimension(n), intent(out) :: p
  integer, intent(out) :: ierr
  logical, intent(in) :: lverb
  ierr = crt_cholz(crt_hip_ctx(), int(n, c_int), a, p)
  if (lverb .and. ierr /= 0) print *, 'cholz (libcrtomo_hip): errnr', ierr
end subroutine cholz
