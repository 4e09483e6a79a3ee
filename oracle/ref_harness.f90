! ref_harness.f90 -- TEST INFRASTRUCTURE ONLY.
!
! A thin ISO_C_BINDING driver around the reference's own Fortran routines
! (compiled unmodified from /root/reference/src by oracle/Makefile into
! oracle/_ref/).  It fills the reference modules from C arrays and calls the
! reference subroutines, so that tests can pin the C oracle (crt_oracle.c)
! against the reference itself, and bench.py can time the reference's OpenMP
! forward loop as the CPU baseline.  Nothing in the product links this.
!
! Driver sequence follows src/fem.f90:293-397 (forward), src/fem.f90:400-428
! (rtrafo, bvolt), src/inv.f90:490-495,742 (bvolti, bsensi) and
! src/inv.f90:929 (update -> cjg).
module crt_ref_harness
  use iso_c_binding
  use alloci
  use femmod
  use elemmod
  use electrmod
  use wavenmod
  use sigmamod
  use datmod
  use errmod
  use konvmod
  use modelmod
  use invmod
  use cjgmod
  use bsmatm_mod
  use randbmod
  implicit none

contains

  integer(c_int) function ref_set_grid(sanz_, mb_, typanz_, typ_, nelanz_, selanz_, maxn_, &
       nrel_c, x_, y_, rnr_, eanz_, enr_) bind(C, name='ref_set_grid')
    integer(c_int), value :: sanz_, mb_, typanz_, maxn_, eanz_
    integer(c_int), intent(in) :: typ_(typanz_), nelanz_(typanz_), selanz_(typanz_)
    integer(c_int), intent(in) :: nrel_c(maxn_, *), rnr_(*), enr_(eanz_)
    real(c_double), intent(in) :: x_(sanz_), y_(sanz_)
    integer :: i, k, neltot, ifln, idum, j

    if (allocated(typ)) deallocate(typ, nelanz, selanz)
    if (allocated(sx)) deallocate(sx, sy, snr)
    if (allocated(nrel)) deallocate(nrel)
    if (allocated(rnr)) deallocate(rnr)
    if (allocated(enr)) deallocate(enr)
    if (allocated(espx)) deallocate(espx, espy)
    sanz = sanz_; mb = mb_; typanz = typanz_; max_nr_element_nodes = maxn_
    allocate(typ(typanz), nelanz(typanz), selanz(typanz))
    typ = typ_; nelanz = nelanz_; selanz = selanz_
    elanz = 0; relanz = 0
    do i = 1, typanz
       if (typ(i) > 10) then
          relanz = relanz + nelanz(i)
       else
          elanz = elanz + nelanz(i)
       end if
    end do
    neltot = elanz + relanz
    allocate(nrel(neltot, maxn_), rnr(max(relanz, 1)), sx(sanz), sy(sanz), snr(sanz))
    nrel = 0
    do i = 1, neltot
       do k = 1, maxn_
          nrel(i, k) = nrel_c(k, i)
       end do
    end do
    if (relanz > 0) rnr(1:relanz) = rnr_(1:relanz)
    do i = 1, sanz
       snr(i) = i
       sx(i) = x_(i)
       sy(i) = y_(i)
    end do
    eanz = eanz_
    allocate(enr(eanz))
    enr = enr_
    ! element centroids as src/relem.f90:128-138
    allocate(espx(elanz), espy(elanz))
    espx = 0d0; espy = 0d0
    ifln = 0; idum = 0
    do i = 1, typanz
       do j = 1, nelanz(i)
          if (typ(i) < 10) then
             ifln = ifln + 1
             do k = 1, selanz(i)
                espx(ifln) = espx(ifln) + sx(snr(nrel(idum + j, k)))
                espy(ifln) = espy(ifln) + sy(snr(nrel(idum + j, k)))
             end do
             espx(ifln) = espx(ifln) / selanz(i)
             espy(ifln) = espy(ifln) / selanz(i)
          end if
       end do
       idum = idum + nelanz(i)
    end do
    manz = elanz
    lsink = .false.; lrandb2 = .false.; lsr = .false.; ldc = .false.
    nr_elec_capacitances = 0; decanz = 0
    swrtr = 1
    ref_set_grid = 0
  end function ref_set_grid

  integer(c_int) function ref_set_data(nanz_, strnr_, vnr_) bind(C, name='ref_set_data')
    integer(c_int), value :: nanz_
    integer(c_int), intent(in) :: strnr_(nanz_), vnr_(nanz_)
    if (allocated(strnr)) deallocate(strnr, vnr)
    if (allocated(volt)) deallocate(volt, sigmaa, kfak, strom)
    nanz = nanz_
    allocate(strnr(nanz), vnr(nanz), volt(nanz), sigmaa(nanz), kfak(nanz), strom(nanz))
    strnr = strnr_; vnr = vnr_
    kfak = 1d0; strom = 1d0
    ref_set_data = 0
  end function ref_set_data

  ! boundary.dat values (src/rrandb.f90: rwdanz, rwdnr, rwd and the DC rwddc = DBLE(rwd));
  ! n = 0 switches the prescribed values off (lrandb2)
  integer(c_int) function ref_set_boundary(n, nodes, vals) bind(C, name='ref_set_boundary')
    integer(c_int), value :: n
    integer(c_int), intent(in) :: nodes(*)
    complex(c_double_complex), intent(in) :: vals(*)
    if (allocated(rwdnr)) deallocate(rwdnr)
    if (allocated(rwd)) deallocate(rwd)
    if (allocated(rwddc)) deallocate(rwddc)
    rwdanz = n
    lrandb2 = n > 0
    allocate(rwdnr(max(n, 1)), rwd(max(n, 1)), rwddc(max(n, 1)))
    if (n > 0) then
       rwdnr(1:n) = nodes(1:n)
       rwd(1:n) = vals(1:n)
       rwddc(1:n) = dble(vals(1:n))
    end if
    ref_set_boundary = 0
  end function ref_set_boundary

  integer(c_int) function ref_rwaven(kwnanz_, kwn_, kwnwi_, amin_, amax_) bind(C, name='ref_rwaven')
    integer(c_int), intent(out) :: kwnanz_
    real(c_double), intent(out) :: kwn_(30), kwnwi_(30), amin_, amax_
    external rwaven
    if (allocated(kwn)) deallocate(kwn)
    if (allocated(kwnwi)) deallocate(kwnwi)
    errnr = 0
    call rwaven()
    ref_rwaven = errnr
    if (errnr /= 0) return
    kwnanz_ = kwnanz
    kwn_(1:kwnanz) = kwn; kwnwi_(1:kwnanz) = kwnwi
    amin_ = amin; amax_ = amax
  end function ref_rwaven

  integer(c_int) function ref_precal(lsytop_, elbg_, relbg_, kg_, lbeta_) bind(C, name='ref_precal')
    integer(c_int), value :: lsytop_
    real(c_double), intent(out) :: elbg_(*), relbg_(*), kg_(*)
    integer(c_int), intent(out) :: lbeta_
    integer :: n
    external precal
    if (allocated(elbg)) deallocate(elbg)
    if (allocated(relbg)) deallocate(relbg)
    if (allocated(kg)) deallocate(kg)
    lsytop = (lsytop_ /= 0)
    errnr = 0
    call precal()
    ref_precal = errnr
    if (errnr /= 0) return
    n = size(elbg); elbg_(1:n) = reshape(elbg, (/n/))
    n = size(relbg); if (relanz > 0) relbg_(1:n) = reshape(relbg, (/n/))
    n = size(kg); if (relanz > 0) kg_(1:n) = reshape(kg, (/n/))
    lbeta_ = merge(1, 0, lbeta)
  end function ref_precal

  ! forward loop over (k, l) exactly as src/fem.f90:328-397 (lsr off, no Dirichlet)
  integer(c_int) function ref_forward(sigma_, nthreads, kpot_) bind(C, name='ref_forward')
    complex(c_double_complex), intent(in) :: sigma_(*)
    integer(c_int), value :: nthreads
    complex(c_double_complex), intent(out) :: kpot_(*)
    complex(kind(0d0)), allocatable :: la(:), lb(:), lpot(:)
    real(kind(0d0)), allocatable :: lfak(:)
    integer :: j, l, k, n
    external kompab, randb, randb2, scalab, chol, kompb, vre
    if (allocated(sigma)) deallocate(sigma)
    allocate(sigma(elanz))
    sigma = sigma_(1:elanz)
    if (allocated(kpot)) deallocate(kpot)
    allocate(kpot(sanz, eanz, kwnanz))
    allocate(la((mb + 1) * sanz), lb(sanz), lpot(sanz), lfak(sanz))
    call omp_set_num_threads(nthreads)
    !$OMP PARALLEL DEFAULT (none) &
    !$OMP FIRSTPRIVATE (lfak,lpot,la,lb) &
    !$OMP PRIVATE (j,l,k) &
    !$OMP SHARED (kwnanz,eanz,lbeta,lrandb,lrandb2,sanz,kpot)
    !$OMP DO
    do k = 1, kwnanz
       do l = 1, eanz
          if (lbeta .or. l == 1) then
             call kompab(l, k, la, lb)
             if (lrandb) call randb(la, lb)
             if (lrandb2) call randb2(la, lb)
             call scalab(la, lb, lfak)
             call chol(la)
          else
             call kompb(l, lb, lfak)
          end if
          call vre(la, lb, lpot)
          do j = 1, sanz
             kpot(j, l, k) = lpot(j) * dcmplx(lfak(j))
          end do
       end do
    end do
    !$OMP END PARALLEL
    n = size(kpot)
    kpot_(1:n) = reshape(kpot, (/n/))
    ref_forward = errnr
  end function ref_forward

  ! reference conductivity of sigma_ (src/refsig.f90)
  integer(c_int) function ref_refsig(sigma_, s0_) bind(C, name='ref_refsig')
    complex(c_double_complex), intent(in) :: sigma_(*)
    complex(c_double_complex), intent(out) :: s0_
    external refsig
    if (allocated(sigma)) deallocate(sigma)
    allocate(sigma(elanz))
    sigma = sigma_(1:elanz)
    call refsig()
    s0_ = sigma0
    ref_refsig = 0
  end function ref_refsig

  ! analytic half-space potential of electrode l at wavenumber k (src/potana.f90), with
  ! the sigma0 of the last ref_refsig
  integer(c_int) function ref_potana(l, k, pota_) bind(C, name='ref_potana')
    integer(c_int), value :: l, k
    complex(c_double_complex), intent(out) :: pota_(*)
    complex(kind(0d0)), allocatable :: lp(:)
    external potana
    allocate(lp(sanz))
    call potana(l, k, lp)
    pota_(1:sanz) = lp
    ref_potana = 0
  end function ref_potana

  ! the forward loop of src/fem.f90:302-397 with its switches (serial): refsig, lsr off
  ! without mixed boundaries, potana for lsr or lana, the assembly for lbeta, l = 1 or
  ! lsr (not lana), kompb with the sink otherwise, kpot = pota (lana) or pot fak (+ pota)
  integer(c_int) function ref_forward_sw(sigma_, lsr_, lana_, lsink_, nsink_, kpot_) &
       bind(C, name='ref_forward_sw')
    complex(c_double_complex), intent(in) :: sigma_(*)
    integer(c_int), value :: lsr_, lana_, lsink_, nsink_
    complex(c_double_complex), intent(out) :: kpot_(*)
    complex(kind(0d0)), allocatable :: la(:), lb(:), lpot(:)
    real(kind(0d0)), allocatable :: lfak(:)
    logical :: lana
    integer :: j, l, k, n
    external kompab, randb, randb2, scalab, chol, kompb, vre, potana, refsig
    if (allocated(sigma)) deallocate(sigma)
    allocate(sigma(elanz))
    sigma = sigma_(1:elanz)
    lsr = lsr_ /= 0
    lana = lana_ /= 0
    if (lsr) lana = .false.
    lsink = lsink_ /= 0
    nsink = nsink_
    errnr = 0
    call refsig()
    if (.not. lbeta) lsr = .false.
    if (allocated(kpot)) deallocate(kpot)
    allocate(kpot(sanz, eanz, kwnanz))
    allocate(la((mb + 1) * sanz), lb(sanz), lpot(sanz), lfak(sanz))
    ! kompab's singularity removal reads femmod's pota, which fem.f90 fills by potana
    if (allocated(pota)) deallocate(pota)
    allocate(pota(sanz))
    do k = 1, kwnanz
       do l = 1, eanz
          if (lsr .or. lana) call potana(l, k, pota)
          if ((lbeta .or. l == 1 .or. lsr) .and. .not. lana) then
             call kompab(l, k, la, lb)
             if (lrandb) call randb(la, lb)
             if (lrandb2) call randb2(la, lb)
             call scalab(la, lb, lfak)
             call chol(la)
          else if (.not. lana) then
             call kompb(l, lb, lfak)
          end if
          if (.not. lana) call vre(la, lb, lpot)
          do j = 1, sanz
             if (lana) then
                kpot(j, l, k) = pota(j)
             else
                kpot(j, l, k) = lpot(j) * dcmplx(lfak(j))
                if (lsr) kpot(j, l, k) = kpot(j, l, k) + pota(j)
             end if
          end do
       end do
    end do
    n = size(kpot)
    kpot_(1:n) = reshape(kpot, (/n/))
    ref_forward_sw = errnr
    lsr = .false.; lsink = .false.
  end function ref_forward_sw

  ! half-space geometric factors of the measurements (src/bkfak.f90; writes tmp.kfak
  ! into the working directory)
  integer(c_int) function ref_bkfak(kfak_) bind(C, name='ref_bkfak')
    real(c_double), intent(out) :: kfak_(*)
    external bkfak
    errnr = 0
    call bkfak()
    ref_bkfak = errnr
    if (errnr == 0) kfak_(1:nanz) = kfak(1:nanz)
  end function ref_bkfak

  ! Bounded sample of the same OpenMP loop (k = 1..nk, l = 1..nl) for timing;
  ! returns the wall time in seconds.
  real(c_double) function ref_forward_sample(sigma_, nthreads, nk, nl) bind(C, name='ref_forward_sample')
    complex(c_double_complex), intent(in) :: sigma_(*)
    integer(c_int), value :: nthreads, nk, nl
    complex(kind(0d0)), allocatable :: la(:), lb(:), lpot(:)
    real(kind(0d0)), allocatable :: lfak(:)
    integer :: j, l, k
    integer(kind=8) :: c0, c1, rate
    external kompab, scalab, chol, kompb, vre
    if (allocated(sigma)) deallocate(sigma)
    allocate(sigma(elanz))
    sigma = sigma_(1:elanz)
    if (allocated(kpot)) deallocate(kpot)
    allocate(kpot(sanz, eanz, kwnanz))
    allocate(la((mb + 1) * sanz), lb(sanz), lpot(sanz), lfak(sanz))
    call omp_set_num_threads(nthreads)
    call system_clock(c0, rate)
    !$OMP PARALLEL DEFAULT (none) &
    !$OMP FIRSTPRIVATE (lfak,lpot,la,lb) &
    !$OMP PRIVATE (j,l,k) &
    !$OMP SHARED (nk,nl,lbeta,sanz,kpot)
    !$OMP DO
    do k = 1, nk
       do l = 1, nl
          if (lbeta .or. l == 1) then
             call kompab(l, k, la, lb)
             call scalab(la, lb, lfak)
             call chol(la)
          else
             call kompb(l, lb, lfak)
          end if
          call vre(la, lb, lpot)
          do j = 1, sanz
             kpot(j, l, k) = lpot(j) * dcmplx(lfak(j))
          end do
       end do
    end do
    !$OMP END PARALLEL
    call system_clock(c1)
    ref_forward_sample = dble(c1 - c0) / dble(rate)
  end function ref_forward_sample

  ! bsensi over the first nmeas measurements (OpenMP over measurements as
  ! src/bsensi.f90:90-96); returns the wall time in seconds.
  real(c_double) function ref_sens_sample(nthreads, nmeas) bind(C, name='ref_sens_sample')
    integer(c_int), value :: nthreads, nmeas
    integer :: j, nsave
    integer(kind=8) :: c0, c1, rate
    external bsensi
    if (allocated(mnr)) deallocate(mnr)
    allocate(mnr(elanz))
    do j = 1, elanz
       mnr(j) = j
    end do
    nsave = nanz
    nanz = nmeas
    ! timing only: a sampled forward leaves most voltages unset
    where (abs(volt) == 0d0) volt = dcmplx(1d0, 0d0)
    if (allocated(sens)) deallocate(sens)
    allocate(sens(nanz, manz))
    call omp_set_num_threads(nthreads)
    call system_clock(c0, rate)
    call bsensi(.false.)
    call system_clock(c1)
    nanz = nsave
    ref_sens_sample = dble(c1 - c0) / dble(rate)
  end function ref_sens_sample

  integer(c_int) function ref_voltages(complex_inv, hpot_, volt_, sigmaa_) bind(C, name='ref_voltages')
    integer(c_int), value :: complex_inv
    complex(c_double_complex), intent(out) :: hpot_(*), volt_(*), sigmaa_(*)
    integer :: n
    external rtrafo, bvolt, bvolti
    if (allocated(hpot)) deallocate(hpot)
    allocate(hpot(sanz, eanz))
    call rtrafo()
    errnr = 0
    if (complex_inv /= 0) then
       call bvolti()
    else
       call bvolt()
    end if
    n = size(hpot); hpot_(1:n) = reshape(hpot, (/n/))
    volt_(1:nanz) = volt; sigmaa_(1:nanz) = sigmaa
    ref_voltages = errnr
  end function ref_voltages

  integer(c_int) function ref_sens(inversion, sens_) bind(C, name='ref_sens')
    integer(c_int), value :: inversion
    complex(c_double_complex), intent(out) :: sens_(*)
    integer :: j, n
    external bsensi, bsens
    if (allocated(mnr)) deallocate(mnr)
    allocate(mnr(elanz))
    do j = 1, elanz
       mnr(j) = j
    end do
    if (allocated(sens)) deallocate(sens)
    allocate(sens(nanz, manz))
    if (inversion /= 0) then
       call bsensi(.false.)
    else
       call bsens()
    end if
    n = size(sens); sens_(1:n) = reshape(sens, (/n/))
    ref_sens = 0
  end function ref_sens

  ! DC forward loop exactly as src/inv.f90:355-398 (lsr off, no Dirichlet), serial
  integer(c_int) function ref_forward_dc(sigma_, kpotdc_) bind(C, name='ref_forward_dc')
    complex(c_double_complex), intent(in) :: sigma_(*)
    real(c_double), intent(out) :: kpotdc_(*)
    real(kind(0d0)), allocatable :: la(:), lb(:), lfak(:)
    complex(kind(0d0)), allocatable :: lpot(:)
    integer :: j, l, k, n
    external kompadc, randdc, randbdc2, scaldc, choldc, kompbdc, vredc
    if (allocated(sigma)) deallocate(sigma)
    allocate(sigma(elanz))
    sigma = sigma_(1:elanz)
    if (allocated(kpotdc)) deallocate(kpotdc)
    allocate(kpotdc(sanz, eanz, kwnanz))
    allocate(la((mb + 1) * sanz), lb(sanz), lpot(sanz), lfak(sanz))
    errnr = 0
    do k = 1, kwnanz
       do l = 1, eanz
          if (lbeta .or. l == 1) then
             call kompadc(l, k, la, lb)
             if (lrandb) call randdc(la, lb)
             if (lrandb2) call randbdc2(la, lb)
             call scaldc(la, lb, lfak)
             call choldc(la)
          else
             call kompbdc(l, lb, lfak)
          end if
          call vredc(la, lb, lpot)
          do j = 1, sanz
             kpotdc(j, l, k) = dble(lpot(j)) * lfak(j)
          end do
       end do
    end do
    n = size(kpotdc)
    kpotdc_(1:n) = reshape(kpotdc, (/n/))
    ref_forward_dc = errnr
  end function ref_forward_dc

  ! rtrafo + bvolti with ldc set (src/rtrafo.f90:48-66, src/bvolti.f90:55-66)
  integer(c_int) function ref_voltages_dc(hpotdc_, volt_, sigmaa_) bind(C, name='ref_voltages_dc')
    real(c_double), intent(out) :: hpotdc_(*)
    complex(c_double_complex), intent(out) :: volt_(*), sigmaa_(*)
    integer :: n
    external rtrafo, bvolti
    if (allocated(hpotdc)) deallocate(hpotdc)
    allocate(hpotdc(sanz, eanz))
    ldc = .true.
    errnr = 0
    call rtrafo()
    call bvolti()
    ldc = .false.
    n = size(hpotdc); hpotdc_(1:n) = reshape(hpotdc, (/n/))
    volt_(1:nanz) = volt; sigmaa_(1:nanz) = sigmaa
    ref_voltages_dc = errnr
  end function ref_voltages_dc

  ! bsendc (src/bsendc.f90:1) on the state of ref_forward_dc / ref_voltages_dc
  integer(c_int) function ref_sens_dc(sensdc_) bind(C, name='ref_sens_dc')
    real(c_double), intent(out) :: sensdc_(*)
    integer :: j, n
    external bsendc
    if (allocated(mnr)) deallocate(mnr)
    allocate(mnr(elanz))
    do j = 1, elanz
       mnr(j) = j
    end do
    if (allocated(sensdc)) deallocate(sensdc)
    allocate(sensdc(nanz, manz))
    call bsendc(.false.)
    n = size(sensdc); sensdc_(1:n) = reshape(sensdc, (/n/))
    ref_sens_dc = 0
  end function ref_sens_dc

  integer(c_int) function ref_chol(sanz_, mb_, a_) bind(C, name='ref_chol')
    integer(c_int), value :: sanz_, mb_
    complex(c_double_complex), intent(inout) :: a_(*)
    external chol
    sanz = sanz_; mb = mb_
    errnr = 0
    call chol(a_)
    ref_chol = errnr
  end function ref_chol

  integer(c_int) function ref_vre(sanz_, mb_, a_, b_, pot_) bind(C, name='ref_vre')
    integer(c_int), value :: sanz_, mb_
    complex(c_double_complex), intent(in) :: a_(*), b_(*)
    complex(c_double_complex), intent(out) :: pot_(*)
    external vre
    sanz = sanz_; mb = mb_
    call vre(a_, b_, pot_)
    ref_vre = 0
  end function ref_vre

  ! besp_elem (src/besp_elem.f90) on the grid of ref_set_grid: centroids as
  ! src/relem.f90:122-143 forms them, bnachbar, then the statistics (inv.gstat goes to
  ! the working directory)
  integer(c_int) function ref_besp(stats_) bind(C, name='ref_besp')
    real(c_double), intent(out) :: stats_(11)
    integer :: i, j, k, ifln, idum
    external bnachbar, besp_elem
    if (allocated(espx)) deallocate(espx, espy)
    allocate(espx(elanz), espy(elanz))
    espx = 0.; espy = 0.
    idum = 0; ifln = 0
    do i = 1, typanz
       do j = 1, nelanz(i)
          if (typ(i) < 10) then
             ifln = ifln + 1
             do k = 1, selanz(i)
                espx(ifln) = espx(ifln) + sx(snr(nrel(idum + j, k)))
                espy(ifln) = espy(ifln) + sy(snr(nrel(idum + j, k)))
             end do
             espx(ifln) = espx(ifln) / selanz(i)
             espy(ifln) = espy(ifln) / selanz(i)
          end if
       end do
       idum = idum + nelanz(i)
    end do
    if (allocated(nachbar)) deallocate(nachbar)
    call bnachbar()
    errnr = 0
    call besp_elem()
    stats_ = (/esp_min, esp_max, esp_mit, esp_med, esp_std, grid_min, grid_max, grid_minx, grid_maxx, &
         grid_miny, grid_maxy/)
    ref_besp = errnr
  end function ref_besp

  ! regularisation matrix through the reference dispatcher bsmatm (src/bsmatm_mod.f90:62)
  integer(c_int) function ref_bsmatm(ltri_, nx_, nz_, alfx_, alfz_, smatm_, nachbar_) &
       bind(C, name='ref_bsmatm')
    integer(c_int), value :: ltri_, nx_, nz_
    real(c_double), value :: alfx_, alfz_
    real(c_double), intent(out) :: smatm_(*)
    integer(c_int), intent(out) :: nachbar_(*)
    logical :: lb
    integer :: n
    external bnachbar
    ltri = ltri_; nx = nx_; nz = nz_; alfx = alfx_; alfz = alfz_
    lverb = .false.
    if (allocated(smatm)) deallocate(smatm)
    if (ltri == 1 .or. ltri == 2) then
       if (allocated(nachbar)) deallocate(nachbar)
       call bnachbar()
       n = size(nachbar); nachbar_(1:n) = reshape(nachbar, (/n/))
    end if
    call bsmatm(1, lb)
    n = size(smatm); smatm_(1:n) = reshape(smatm, (/n/))
    ref_bsmatm = 0
  end function ref_bsmatm

  ! one model update through src/update.f90:27 (complex, ltri 0/1/2, no prior)
  integer(c_int) function ref_update(ncgmax_, eps_, lam_, sens_, wmatd_, wdfak_, dat_, sigmaa_, &
       par_, cgfac_, bvec_, dpar_, cgres_, ncg_, kind_, sensdc_) bind(C, name='ref_update')
    integer(c_int), value :: ncgmax_, kind_
    real(c_double), intent(in) :: sensdc_(*)
    real(c_double), value :: eps_, lam_
    complex(c_double_complex), intent(in) :: sens_(*), dat_(*), sigmaa_(*), par_(*)
    real(c_double), intent(in) :: wmatd_(*)
    integer(c_int), intent(in) :: wdfak_(*)
    real(c_double), intent(out) :: cgfac_(*), cgres_(*)
    complex(c_double_complex), intent(out) :: bvec_(*), dpar_(*)
    integer(c_int), intent(out) :: ncg_
    integer :: n
    external update
    if (allocated(sens)) deallocate(sens)
    allocate(sens(nanz, manz))
    sens = reshape(sens_(1:nanz * manz), (/nanz, manz/))
    if (allocated(sensdc)) deallocate(sensdc)
    allocate(sensdc(nanz, manz))
    if (kind_ == 1) then
       sensdc = reshape(sensdc_(1:nanz * manz), (/nanz, manz/))
    else
       sensdc = 0d0
    end if
    if (allocated(wmatd)) deallocate(wmatd, wdfak, dat, par, dpar, dpar2)
    allocate(wmatd(nanz), wdfak(nanz), dat(nanz), par(manz), dpar(manz), dpar2(manz))
    wmatd = wmatd_(1:nanz); wdfak = wdfak_(1:nanz); dat = dat_(1:nanz)
    sigmaa = sigmaa_(1:nanz); par = par_(1:manz)
    if (allocated(w_ref_re)) deallocate(w_ref_re, w_ref_im)
    n = max(manz, nanz + 1)
    allocate(w_ref_re(n), w_ref_im(n))
    w_ref_re = 0d0; w_ref_im = 0d0
    if (allocated(cgres)) deallocate(cgres, cgres2, bvec, cgfac)
    llam = .false.; lprior = .false.; lw_ref = .false.; ldiff = .false.; lfpi = .false.
    ldc = .false.; lphi0 = .false.; lverb = .false.
    ! kind 1: the DC inversion (ldc, sensdc); kind 2: final phase improvement (lfpi)
    ldc = kind_ == 1; lfpi = kind_ == 2
    lam = lam_; eps = eps_; step = 1d0
    ! pass 1: ncgmax = 0 leaves the scaled RHS in bvec
    ncgmax = 0
    call con_cjgmod(1, fetxt, errnr)
    call update()
    bvec_(1:manz) = bvec; cgfac_(1:manz) = cgfac
    deallocate(cgres, cgres2, bvec, cgfac)
    ! pass 2: the real solve
    par = par_(1:manz)
    ncgmax = ncgmax_
    call con_cjgmod(1, fetxt, errnr)
    call update()
    dpar_(1:manz) = dpar
    cgres_(1:ncgmax + 1) = cgres(1:ncgmax + 1)
    ncg_ = ncg
    ldc = .false.; lfpi = .false.
    ref_update = errnr
  end function ref_update

  ! starting regularisation parameter through src/blam0.f90:12, on the state that
  ! ref_update left (sens, wmatd, wdfak); computed, not preset (llamf = 0)
  real(c_double) function ref_blam0(lfpi_, alfx_, alfz_, ldc_) bind(C, name='ref_blam0')
    integer(c_int), value :: lfpi_, ldc_
    real(c_double), value :: alfx_, alfz_
    external blam0
    llamf = 0; lamnull_cri = 0d0; lamnull_fpi = 0d0
    lfpi = lfpi_ /= 0; ldc = ldc_ /= 0; lverb = .false.
    alfx = alfx_; alfz = alfz_
    call blam0()
    lfpi = .false.; ldc = .false.
    ref_blam0 = lammax
  end function ref_blam0

  ! linearised model uncertainty through the reference driver buncert
  ! (src/bmcm_mod.f90:49) with lamalt = lam_, on the state of the last ref_update
  ! (sens, wmatd, wdfak) and ref_bsmatm (ltri, nx, smatm, nachbar) calls; sigma_
  ! (manz) feeds bpar and the cov1_m/cov2_m columns.  The reference writes its
  ! .diag files (ata, ata_reg, cov1_m, res_m, cov2_m) into the directory dir_.
  ! Its chold stops the process on a non-positive pivot: callers pass positive
  ! definite systems only.
  integer(c_int) function ref_buncert(lam_, lres_, lcov2_, lgauss_, sigma_, dir_, ndir_) &
       bind(C, name='ref_buncert')
    use bmcm_mod, only: buncert
    use pathmod, only: ramd, lnramd, slash
    real(c_double), value :: lam_
    integer(c_int), value :: lres_, lcov2_, lgauss_, ndir_
    complex(c_double_complex), intent(in) :: sigma_(*)
    character(kind=c_char), intent(in) :: dir_(*)
    integer :: j
    ramd = ' '
    do j = 1, ndir_
       ramd(j:j) = dir_(j)
    end do
    lnramd = ndir_; slash = '/'
    lres = lres_ /= 0; lcov2 = lcov2_ /= 0; lgauss = lgauss_ /= 0
    lfpi = .false.; lelerr = .false.; ldc = .false.; lverb = .false.; lverb_dat = .false.
    mswitch = 0
    if (allocated(sigma)) deallocate(sigma)
    allocate(sigma(manz))
    sigma = sigma_(1:manz)
    if (allocated(mnr)) deallocate(mnr)
    allocate(mnr(elanz))
    do j = 1, elanz
       mnr(j) = j
    end do
    if (allocated(par)) deallocate(par)
    allocate(par(manz))
    ! bata reads wmatdp (the FPI weights) for its dig_fpi column whatever lfpi is
    if (allocated(wmatdp)) deallocate(wmatdp)
    allocate(wmatdp(nanz))
    wmatdp = wmatd
    fprun = 93
    open(fprun, file=ramd(1:lnramd)//'/run.ctr', status='replace')
    errnr = 0
    call buncert(94, lam_)
    close(fprun)
    ref_buncert = errnr
  end function ref_buncert

end module crt_ref_harness
