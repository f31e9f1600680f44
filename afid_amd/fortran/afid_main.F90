!-----------------------------------------------------------------------
!  afid_main.F90 -- Fortran host of the MI355X engine: reads bou.in,
!  keeps the main time loop on the host and drives the device through
!  the C ABI (module afid_hip).
!
!  Host logic followed: ReadInputFile.F90:16-36 (bou.in layout),
!  main.F90:153-279 (Factorize check, CFL/dt selection, output cadence,
!  error codes), TimeMarcher.F90:13-78 (substep sequence), the log files
!  of CalcPlateNu.F90:49-53, GlobalQuantities.F90:81-97 and
!  CalcDissipationNu.F90:114-116; ReadFlowField.F90:29-45,160-176 (nread)
!  and QuitRoutine.F90:27-31 (statistics and continuation written at a
!  normal exit); StatRoutines.F90:12-44 and SlabDumpRoutines.F90:12-70
!  at the output times after tsta (main.F90:210-216), slabs into stst3/.
!
!  Usage: afid_run [bou.in] [--seed N] [--routines] [--mixed] [--out DIR]
!                  [--parse-only] [--abi-check]
!    --routines   TimeMarcher calls the per-routine entry points in the
!                 reference's order instead of the fused substep
!    --mixed      mixed-precision mode (FP32 histories and FFT spectrum,
!                 one refinement sweep per step); not with --routines
!    --out DIR    final fields vx, vy, vz, temp, pr are written there as
!                 raw FP64 stream files, Fortran order (nx, nym, nzm)
!    --parse-only print the parsed bou.in and stop (no device use)
!    --abi-check  print the size of the bind(C) parameter block and stop
!-----------------------------------------------------------------------
      module afid_run_state
      use iso_c_binding
      use afid_hip
      implicit none
!     bou.in (names as in param.F90)
      integer :: nxm, nym, nzm, nsst, nread, ntst, ireset, istr3
      integer :: flagstat, flagbal, starea, inslws, inslwn, idtv
      integer :: stst3flag
      logical :: dumpslabs = .false.
      integer(c_int) :: kslab(9) = 2
      real(8) :: walltimemax, tout, tmax, alx3, str3, ylen, zlen
      real(8) :: ray, pra, dt, resid, limitcfl, tsta, dtmin, dtmax
      real(8) :: limitvel
!     run state
      type(c_ptr) :: ctx = c_null_ptr
      real(8) :: time = 0.d0
      logical :: per_routine = .false.
      logical :: mixed = .false.
      integer(c_int64_t) :: seed = 1

      contains

      subroutine ReadInputFile(fname)
        character(len=*), intent(in) :: fname
        character(len=4) :: skip
        integer :: u
        open(newunit=u, file=fname, status='old', action='read')
        read(u,'(a4)') skip
        read(u,*) nxm, nym, nzm, nsst, nread
        read(u,'(a4)') skip
        read(u,*) ntst, walltimemax, tout, tmax, ireset
        read(u,'(a4)') skip
        read(u,*) alx3, istr3, str3
        read(u,'(a4)') skip
        read(u,*) ylen, zlen
        read(u,'(a4)') skip
        read(u,*) ray, pra, dt, resid, limitcfl
        read(u,'(a4)') skip
        read(u,*) flagstat, flagbal, tsta, starea
        read(u,'(a4)') skip
        read(u,*) inslws, inslwn
        read(u,'(a4)') skip
        read(u,*) idtv, dtmin, dtmax, limitvel
        read(u,'(a4)') skip
        read(u,*) stst3flag
        close(u)
        if (stst3flag /= 0) then
          inquire(file='./stst3.in', exist=dumpslabs)
          if (.not. dumpslabs)                                            &
     &      write(6,*) 'stst3.in not found, turning off slab dump'
        end if
      end subroutine ReadInputFile

!     One full time step of the reference's TimeMarcher: per-routine
!     entry points in the reference order, or one fused call.
      subroutine TimeMarcher
        integer(c_int) :: ns
        if (.not. per_routine) then
          call afid_check(afid_time_marcher(ctx, dt), 'TimeMarcher')
          return
        end if
        do ns = 1, int(merge(3, 1, nsst > 1), c_int)
          call afid_check(afid_set_substep(ctx, ns, dt), 'set_substep')
          call afid_check(afid_explicit_terms_vx(ctx), 'ExplicitTermsVX')
          call afid_check(afid_explicit_terms_vy(ctx), 'ExplicitTermsVY')
          call afid_check(afid_explicit_terms_vz(ctx), 'ExplicitTermsVZ')
          call afid_check(afid_explicit_terms_temp(ctx),                  &
     &      'ExplicitTermsTemp')
          call afid_check(afid_implicit_and_update_vx(ctx),               &
     &      'ImplicitAndUpdateVX')
          call afid_check(afid_implicit_and_update_vy(ctx),               &
     &      'ImplicitAndUpdateVY')
          call afid_check(afid_implicit_and_update_vz(ctx),               &
     &      'ImplicitAndUpdateVZ')
          call afid_check(afid_implicit_and_update_temp(ctx),             &
     &      'ImplicitAndUpdateTemp')
          call afid_check(afid_update_halo(ctx, AFID_VY), 'update_halo')
          call afid_check(afid_update_halo(ctx, AFID_VZ), 'update_halo')
          call afid_check(afid_calc_local_divergence(ctx),                &
     &      'CalcLocalDivergence')
          call afid_check(afid_solve_pressure_correction(ctx),            &
     &      'SolvePressureCorrection')
          call afid_check(afid_correct_velocity(ctx), 'CorrectVelocity')
          call afid_check(afid_correct_pressure(ctx), 'CorrectPressure')
        end do
      end subroutine TimeMarcher

!     Log files of the reference diagnostics, appended per output time.
      subroutine WriteDiagnostics(vmax, dmax)
        real(8), intent(out) :: vmax(3), dmax
        real(8) :: g(11), nub, nut, numu, nuth, cfl
        integer :: u
        call afid_check(afid_global_quantities(ctx, g),                   &
     &    'GlobalQuantities')
        vmax = g(5:7)
        open(newunit=u, file='nu_vol.out', status='unknown',              &
     &    position='append')
        write(u,*) time, g(1)
        close(u)
        open(newunit=u, file='rms_vel.out', status='unknown',             &
     &    position='append')
        write(u,*) time, g(8), g(9), g(10), g(11)
        close(u)
        call afid_check(afid_calc_max_cfl(ctx, cfl), 'CalcMaxCFL')
        call afid_check(afid_check_divergence(ctx, dmax),                 &
     &    'CheckDivergence')
        call afid_check(afid_calc_plate_nu(ctx, nub, nut), 'CalcPlateNu')
        open(newunit=u, file='nu_plate.out', status='unknown',            &
     &    position='append')
        write(u,'(4(1x,e14.6))') time, nub, nut
        close(u)
        if (time > tsta .and. flagstat /= 0)                              &
     &    call afid_check(afid_calc_stats(ctx), 'CalcStats')
        if (time > tsta .and. dumpslabs)                                  &
     &    call afid_check(afid_write_slabs(ctx, 'stst3'//c_null_char,     &
     &      kslab, time, tout), 'SlabDumper')
        if (time > tsta .and. flagstat /= 0 .and. flagbal /= 0) then
          call afid_check(afid_calc_dissipation_nu(ctx, numu, nuth,       &
     &      c_null_ptr, c_null_ptr), 'CalcDissipationNu')
          open(newunit=u, file='nu_diss.out', status='unknown',           &
     &      position='append')
          write(u,*) time, numu, nuth
          close(u)
        end if
      end subroutine WriteDiagnostics

!     InitializeSlabDump (SlabDumpRoutines.F90:40-70): levels of the
!     xslab(2:9) read from stst3.in, written to stst3locs.out.
      subroutine InitializeSlabDump
        character(len=4) :: skip
        real(8) :: xslab(8)
        real(8), allocatable :: xm(:)
        integer :: u
        open(newunit=u, file='stst3.in', status='old', action='read')
        read(u,'(a4)') skip
        read(u,*) xslab
        close(u)
        allocate(xm(nxm))
        call afid_check(afid_get_metric(ctx, 'xm'//c_null_char, xm, nxm),  &
     &    'InitializeSlabDump')
        call afid_check(afid_slab_levels(xm, nxm, xslab, kslab),          &
     &    'InitializeSlabDump')
        deallocate(xm)
        open(newunit=u, file='stst3locs.out', status='unknown')
        write(u,*) kslab
        close(u)
      end subroutine InitializeSlabDump

      subroutine WriteFields(dir)
        character(len=*), intent(in) :: dir
        character(len=6), parameter :: names(5) =                         &
     &    ['vx    ', 'vy    ', 'vz    ', 'temp  ', 'pr    ']
        integer(c_int), parameter :: ids(5) =                             &
     &    [AFID_VX, AFID_VY, AFID_VZ, AFID_TEMP, AFID_PR]
        real(8), allocatable :: a(:,:,:)
        integer :: q, u, nx
        nx = nxm + 1
        allocate(a(nx, nym, nzm))
        do q = 1, 5
          call afid_check(afid_download(ctx, ids(q), a, nx, nym, 1, 1),   &
     &      'download')
          open(newunit=u, file=trim(dir)//'/'//trim(names(q))//'.bin',    &
     &      access='stream', form='unformatted', status='replace')
          write(u) a
          close(u)
        end do
        deallocate(a)
      end subroutine WriteFields

      integer function LargestFactor(n)
!       largest prime factor of n (main.F90:159-165 rejects sizes with
!       a factor above 5 through Factorize)
        integer, intent(in) :: n
        integer :: m, d
        m = n
        LargestFactor = 1
        d = 2
        do while (m > 1)
          if (mod(m, d) == 0) then
            m = m / d
            LargestFactor = d
          else
            d = d + 1
          end if
        end do
      end function LargestFactor

      end module afid_run_state


      program afid_run
      use iso_c_binding
      use afid_hip
      use afid_run_state
      implicit none
      character(len=512) :: arg, fname, outdir
      type(afid_params_t) :: p
      integer :: ia, na, ntime, errorcode
      logical :: parse_only, abi_check
      real(8) :: cfl, dmax, vmax(3), t0, t1, instcfl
      integer(8) :: c0, c1, crate

      fname = 'bou.in'
      outdir = ''
      parse_only = .false.
      abi_check = .false.
      na = command_argument_count()
      ia = 1
      do while (ia <= na)
        call get_command_argument(ia, arg)
        select case (trim(arg))
        case ('--seed')
          ia = ia + 1
          call get_command_argument(ia, arg)
          read(arg,*) seed
        case ('--routines')
          per_routine = .true.
        case ('--mixed')
          mixed = .true.
        case ('--out')
          ia = ia + 1
          call get_command_argument(ia, outdir)
        case ('--parse-only')
          parse_only = .true.
        case ('--abi-check')
          abi_check = .true.
        case default
          fname = arg
        end select
        ia = ia + 1
      end do

      if (abi_check) then
        write(6,'(a,i0)') 'sizeof(afid_params)=', c_sizeof(p)
        stop
      end if

      call ReadInputFile(trim(fname))
      if (parse_only) then
        write(6,'(a,5(1x,i0))') 'grid', nxm, nym, nzm, nsst, nread
        write(6,'(a,i0,4(1x,es13.6))') 'ntst ', ntst, walltimemax,        &
     &    tout, tmax, dble(ireset)
        write(6,'(a,es13.6,1x,i0,1x,es13.6)') 'stretch ', alx3, istr3,    &
     &    str3
        write(6,'(a,2(1x,es13.6))') 'lengths', ylen, zlen
        write(6,'(a,5(1x,es13.6))') 'physics', ray, pra, dt, resid,       &
     &    limitcfl
        write(6,'(a,2(1x,i0))') 'walls', inslws, inslwn
        write(6,'(a,i0,3(1x,es13.6))') 'timestep ', idtv, dtmin, dtmax,   &
     &    limitvel
        stop
      end if

      errorcode = 0
      if (max(LargestFactor(nym), LargestFactor(nzm)) > 5) then
        write(6,*) 'FFT size in ny or nz is not efficient'
        error stop 444
      end if

      p%nxm = nxm
      p%nym = nym
      p%nzm = nzm
      p%nsst = nsst
      p%istr3 = istr3
      p%str3 = str3
      p%alx3 = alx3
      p%ylen = ylen
      p%zlen = zlen
      p%ray = ray
      p%pra = pra
      p%inslws = inslws
      p%inslwn = inslwn
      p%tempbp = 1.d0
      p%temptp = 0.d0
      p%precision = AFID_PREC_FP64
      if (mixed) p%precision = AFID_PREC_MIXED
      call afid_check(afid_create(ctx, p, c_null_ptr), 'afid_create')
      if (nread /= 0) then
        write(6,*) 'Reading initial condition from file'
        call afid_check(afid_read_continuation(ctx, '.'//c_null_char),   &
     &    'ReadFlowField')
        call afid_check(afid_get_time(ctx, time), 'ReadFlowField')
        if (ireset /= 0) time = 0.d0
        tmax = tmax + time
      else
        write(6,*) 'Creating initial condition'
        call afid_check(afid_init_linear_noise(ctx, seed, 1.d-3),         &
     &    'CreateInitialConditions')
        time = 0.d0
      end if
      if (dumpslabs) call InitializeSlabDump
      call afid_check(afid_check_divergence(ctx, dmax), 'CheckDivergence')
      write(6,*) ' Initial maximum divergence: ', dmax

      call system_clock(c0, crate)
      t0 = dble(c0) / dble(crate)
      do ntime = 0, ntst
        call afid_check(afid_calc_max_cfl(ctx, instcfl), 'CalcMaxCFL')
        if (idtv /= 0) then
          if (ntime > 1) then
            if (instcfl < 1.d-8) then
              dt = dtmax
            else
              dt = limitcfl / instcfl
            end if
            if (dt > dtmax) dt = dtmax
          else
            dt = dtmin
          end if
          if (dt < dtmin) errorcode = 166
        else
          cfl = instcfl * dt
          if (cfl > limitcfl) errorcode = 165
        end if

        call TimeMarcher

        if (mod(time, tout) < dt) then
          write(6,*) ' ---------------------------------------- '
          write(6,*) ' T = ', time, ' NTIME = ', ntime, ' DT = ', dt
        end if
        time = time + dt

        if (ntime == 1 .or. mod(time, tout) < dt) then
          call WriteDiagnostics(vmax, dmax)
          if (vmax(1) > limitvel .and. vmax(2) > limitvel)               &
     &      errorcode = 266
          if (abs(dmax) > resid) errorcode = 169
        end if
        if (time > tmax) errorcode = 333
        call system_clock(c1)
        t1 = dble(c1) / dble(crate)
        if (t1 - t0 > walltimemax) errorcode = 334
        if (ntime == ntst) errorcode = 555
        if (errorcode /= 0) exit
      end do

      call afid_check(afid_synchronize(ctx), 'synchronize')
      call system_clock(c1)
      write(6,'(a,f10.2,a)') 'Total Iteration Time = ',                  &
     &  dble(c1 - c0) / dble(crate), ' sec.'
      select case (errorcode)
      case (166)
        write(6,'(10x,a,e14.7)') 'dt too small, DT= ', dt
      case (165)
        write(6,'(10x,a)') 'cfl too large  '
      case (266)
        write(6,'(10x,a)') 'velocities diverged'
      case (169)
        write(6,'(10x,a)') 'too large local residue for mass conservation'
      case (333)
        write(6,*) 'time greater than tmax'
      case (334)
        write(6,*) 'walltime greater than walltimemax'
      case default
        write(6,*) 'Maximum number of timesteps reached'
      end select
      write(6,'(a,i0,a,es23.16)') 'errorcode ', errorcode, ' time ', time
      if (errorcode == 333 .or. errorcode == 334 .or. errorcode == 555)   &
     &  then
        if (flagstat /= 0) then
          call afid_check(afid_write_stats(ctx, '.'//c_null_char,         &
     &      merge(1, 0, starea /= 0), merge(1, 0, flagbal /= 0)),         &
     &      'WriteStats')
        end if
        call afid_check(afid_set_time(ctx, time), 'WriteFlowField')
        call afid_check(afid_write_continuation(ctx, '.'//c_null_char),  &
     &    'WriteFlowField')
      end if
      if (len_trim(outdir) > 0) call WriteFields(trim(outdir))
      call afid_check(afid_destroy(ctx), 'afid_destroy')
      end program afid_run
