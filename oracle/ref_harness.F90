! ref_harness.F90 -- driver that runs the reference AFiD routines on given
! inputs and dumps their outputs as raw little-endian float64 streams.
!
! TEST INFRASTRUCTURE ONLY.  The reference routines themselves are compiled
! unchanged from the reference source tree by oracle/ref.mk; this file is
! the project's own driver around them.  Because FFTW3 is absent from this
! image, the reference SolvePressureCorrection.F90 (and InitPressureSolver.F90,
! which builds FFTW plans) cannot be built; this driver supplies a
! SolvePressureCorrection that hands dph to the C oracle's restatement, and
! takes the Poisson coefficients (ak1, ak2, amphk, acphk, apphk) from the
! oracle.  Everything else called here is reference code.
!
! Input file harness.in (list-directed):
!   nxm nym nzm nsst
!   alx3 istr3 str3
!   ylen zlen
!   ray pra
!   inslws inslwn
!   mode nsteps dt idtv dtmin dtmax limitcfl
! mode 1: dump grid.  mode 2: per-routine stages from fields_in.bin.
! mode 3: march nsteps steps from fields_in.bin; the step loop's wall time
!         goes to march_time.txt.
! mode 4: the reference interp of interp_old.bin (Fortran (nxo, nyo-1, nzo-1))
!         onto this grid; interp.in holds nxo nyo nzo istro3 stro3 intvar.
!         The old window is ReadFlowField's at one rank; the padded old array
!         starts at zero, as a fresh allocation does.

module harness_glue
  use iso_c_binding
  implicit none
  interface
    function afo_create(nxm, nym, nzm, nsst, alx3, istr3, str3, ylen, zlen, ray, pra, &
                        inslws, inslwn) bind(C, name='afo_create')
      import
      type(c_ptr) :: afo_create
      integer(c_int), value :: nxm, nym, nzm, nsst, istr3, inslws, inslwn
      real(c_double), value :: alx3, str3, ylen, zlen, ray, pra
    end function
    function afo_array(o, name, cnt, is_int) bind(C, name='afo_array')
      import
      type(c_ptr) :: afo_array
      type(c_ptr), value :: o
      character(kind=c_char), dimension(*) :: name
      integer(c_long) :: cnt
      integer(c_int) :: is_int
    end function
    function afo_solve_pressure_correction(o) bind(C, name='afo_solve_pressure_correction')
      import
      integer(c_int) :: afo_solve_pressure_correction
      type(c_ptr), value :: o
    end function
  end interface
  type(c_ptr) :: oracle = c_null_ptr
contains
  subroutine oracle_vec(name, n, out)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real(8), intent(out) :: out(n)
    real(c_double), pointer :: p(:)
    integer(c_long) :: cnt
    integer(c_int) :: isint
    type(c_ptr) :: cp
    cp = afo_array(oracle, trim(name)//c_null_char, cnt, isint)
    call c_f_pointer(cp, p, [cnt])
    out(1:n) = p(2:n+1)
  end subroutine
end module harness_glue

subroutine SolvePressureCorrection
  use param
  use local_arrays, only: dph
  use harness_glue
  implicit none
  real(c_double), pointer :: p(:)
  integer(c_long) :: cnt
  integer(c_int) :: isint, info
  type(c_ptr) :: cp
  cp = afo_array(oracle, 'dph'//c_null_char, cnt, isint)
  call c_f_pointer(cp, p, [cnt])
  p = reshape(dph, [cnt])
  info = afo_solve_pressure_correction(oracle)
  dph = reshape(p, shape(dph))
end subroutine

program ref_harness
  use param
  use local_arrays
  use stat_arrays
  use decomp_2d
  use harness_glue
  implicit none
  integer :: mode, nsteps, u
  integer(8) :: clk0, clk1, clkrate
  real(8) :: dtin, dtmin_, dtmax_, limitcfl_, instcfl, dmax
  real(8) :: nu_diag(2)

  open(newunit=u, file='harness.in', status='old')
  read(u, *) nxm, nym, nzm, nsst
  read(u, *) alx3, istr3, str3
  read(u, *) ylen, zlen
  read(u, *) ray, pra
  read(u, *) inslws, inslwn
  read(u, *) mode, nsteps, dtin, idtv, dtmin_, dtmax_, limitcfl_
  close(u)
  nx = nxm + 1; ny = nym + 1; nz = nzm + 1
  ren = dsqrt(ray/pra)
  pec = dsqrt(pra*ray)
  pi = 2.d0*dasin(1.d0)
  statcal = .true.
  disscal = .true.
  dt = dtin

  call decomp_2d_init(nxm, nym, nzm, 1, 1, (/ .false., .true., .true. /))
  call InitTimeMarchScheme
  call InitVariables
  call CreateGrid
  ismaster = .true.

  oracle = afo_create(int(nxm,c_int), int(nym,c_int), int(nzm,c_int), int(nsst,c_int), &
                      alx3, int(istr3,c_int), str3, ylen, zlen, ray, pra, &
                      int(inslws,c_int), int(inslwn,c_int))
  call oracle_vec('ak1', nzm, ak1)
  call oracle_vec('ak2', nym, ak2)
  call oracle_vec('amphk', nxm, amphk)
  call oracle_vec('acphk', nxm, acphk)
  call oracle_vec('apphk', nxm, apphk)
  call SetTempBCs

  if (mode == 1) then
    open(newunit=u, file='grid_out.bin', access='stream', form='unformatted', status='replace')
    write(u) xc(1:nx), xm(1:nx), g3rc(1:nx), g3rm(1:nx), udx3c(1:nx), udx3m(1:nx)
    write(u) ap3ck(1:nx), ac3ck(1:nx), am3ck(1:nx), ap3sk(1:nx), ac3sk(1:nx), am3sk(1:nx)
    write(u) ap3ssk(1:nx), ac3ssk(1:nx), am3ssk(1:nx)
    write(u) dble(kmv(1:nxm)), dble(kpv(1:nxm)), dble(kmc(1:nxm)), dble(kpc(1:nxm))
    write(u) dx, dy, dz, ren, pec, gam(1:3), rom(1:3), alm(1:3)
    write(u) yc(1:ny), ym(1:nym), zc(1:nz), zm(1:nzm)
    close(u)
  else if (mode == 2) then
    call read_fields()
    ! stage A: explicit terms (RK coefficients of substep nsteps)
    al = alm(nsteps); ga = gam(nsteps); ro = rom(nsteps)
    beta = dt/ren*0.5d0
    call ExplicitTermsVX
    call ExplicitTermsVY
    call ExplicitTermsVZ
    call ExplicitTermsTemp
    open(newunit=u, file='stageA.bin', access='stream', form='unformatted', status='replace')
    write(u) qcap, dph, dq, hro
    close(u)
    ! stage B: implicit updates
    call ImplicitAndUpdateVX
    call ImplicitAndUpdateVY
    call ImplicitAndUpdateVZ
    call ImplicitAndUpdateTemp
    call update_halo(vy, lvlhalo)
    call update_halo(vz, lvlhalo)
    open(newunit=u, file='stageB.bin', access='stream', form='unformatted', status='replace')
    write(u) vx, vy, vz, temp, rhs, rux, ruy, ruz, rutemp
    close(u)
    ! stage C: local divergence
    call CalcLocalDivergence
    open(newunit=u, file='stageC.bin', access='stream', form='unformatted', status='replace')
    write(u) dph
    close(u)
    ! stage D: corrections from a given dph (read from dph_in.bin)
    open(newunit=u, file='dph_in.bin', access='stream', form='unformatted', status='old')
    read(u) dph
    close(u)
    dphhalo(1:nxm, 1:nym, 1:nzm) = dph
    call update_halo(dphhalo, lvlhalo)
    call CorrectVelocity
    call CorrectPressure
    open(newunit=u, file='stageD.bin', access='stream', form='unformatted', status='replace')
    write(u) vx, vy, vz, pr
    close(u)
    ! stage E: diagnostics on the updated fields
    call update_halo(vx, lvlhalo)
    call update_halo(vy, lvlhalo)
    call update_halo(vz, lvlhalo)
    call update_halo(temp, lvlhalo)
    call CalcMaxCFL(instcfl)
    call CheckDivergence(dmax)
    disste = 0.d0
    dissth = 0.d0
    call GlobalQuantities
    call CalcPlateNu
    call CalcDissipationNu
    open(newunit=u, file='stageE.bin', access='stream', form='unformatted', status='replace')
    write(u) instcfl, dmax, vmax(1:3), tempm, tempmax, tempmin, disste(1:nxm), dissth(1:nxm)
    close(u)
  else if (mode == 3) then
    call read_fields()
    open(newunit=u, file='march_diag.bin', access='stream', form='unformatted', status='replace')
    call system_clock(clk0, clkrate)
    do ntime = 0, nsteps - 1
      ! dt selection restated from main.F90:171-189
      call CalcMaxCFL(instcfl)
      if (idtv == 1) then
        if (ntime > 1) then
          if (instcfl < 1.0d-8) then
            dt = dtmax_
          else
            dt = limitcfl_/instcfl
          end if
          if (dt > dtmax_) dt = dtmax_
        else
          dt = dtmin_
        end if
      else
        dt = dtin
      end if
      call TimeMarcher
      time = time + dt
      call CheckDivergence(dmax)
      write(u) dt, instcfl, dmax
    end do
    call system_clock(clk1)
    close(u)
    ! wall time of the step loop alone (main.F90:168,226-237 times the same span)
    open(newunit=u, file='march_time.txt', status='replace')
    write(u, '(es24.16)') dble(clk1 - clk0) / dble(clkrate)
    close(u)
    open(newunit=u, file='march_out.bin', access='stream', form='unformatted', status='replace')
    write(u) vx, vy, vz, temp, pr
    close(u)
  else if (mode == 4) then
    call run_interp()
  end if
  nu_diag = 0.d0
  call decomp_2d_finalize

contains
  subroutine run_interp()
    integer :: uu, nxo, nyo, nzo, istro3, intvar, nyom, nzom, js, je, is, ie
    real(8) :: stro3
    real(8), allocatable :: old(:,:,:), arrold(:,:,:), arrnew(:,:,:)
    open(newunit=uu, file='interp.in', status='old')
    read(uu, *) nxo, nyo, nzo, istro3, stro3, intvar
    close(uu)
    nyom = nyo - 1
    nzom = nzo - 1
    allocate(old(nxo, nyom, nzom))
    open(newunit=uu, file='interp_old.bin', access='stream', form='unformatted', status='old')
    read(uu) old
    close(uu)
    js = max(floor(real(xstart(2)*nyom/nym)), 1)
    je = min(ceiling(real(xend(2)*nyom/nym)), nyom)
    is = max(floor(real(xstart(3)*nzom/nzm)), 1)
    ie = min(ceiling(real(xend(3)*nzom/nzm)), nzom)
    allocate(arrold(0:nxo+1, js-1:je+1, is-1:ie+1))
    arrold = 0.d0
    arrold(1:nxo, js:je, is:ie) = old(:, js:je, is:ie)
    allocate(arrnew(nx, nym, nzm))
    call interp(arrold, arrnew, nzo, nyo, nxo, istro3, stro3, intvar, js, je, is, ie)
    open(newunit=uu, file='interp_out.bin', access='stream', form='unformatted', status='replace')
    write(uu) arrnew
    close(uu)
  end subroutine

  subroutine read_fields()
    integer :: uu
    open(newunit=uu, file='fields_in.bin', access='stream', form='unformatted', status='old')
    read(uu) vx, vy, vz, temp, pr, rux, ruy, ruz, rutemp, rhs
    close(uu)
    call update_halo(vx, lvlhalo)
    call update_halo(vy, lvlhalo)
    call update_halo(vz, lvlhalo)
    call update_halo(temp, lvlhalo)
    call update_halo(pr, lvlhalo)
  end subroutine
end program ref_harness
