!-----------------------------------------------------------------------
!  afid_hip.F90 -- iso_c_binding interface to libafid_hip (include/afid.h)
!
!  The reference routines take no arguments and work on module globals
!  (param.F90, local_arrays).  Here one opaque context (type(c_ptr))
!  holds the device state of a cell, and each bind(C) interface below
!  replaces one reference subroutine.  Return codes: 0 OK, <0 runtime
!  error, >0 numerical condition; afid_check prints afid_strerror and
!  stops, as the reference stops through decomp_2d_abort
!  (SolvePressureCorrection.F90:130-134).
!-----------------------------------------------------------------------
      module afid_hip
      use iso_c_binding
      implicit none

      integer(c_int), parameter :: AFID_OK = 0
      integer(c_int), parameter :: AFID_ERR_ARG = -1, AFID_ERR_HIP = -2
      integer(c_int), parameter :: AFID_ERR_NOMEM = -3
      integer(c_int), parameter :: AFID_ERR_FFTSIZE = -4
      integer(c_int), parameter :: AFID_ERR_STATE = -5
      integer(c_int), parameter :: AFID_ERR_IO = -6, AFID_ERR_GRID = -7
      integer(c_int), parameter :: AFID_NUM_PIVOT = 1
      integer(c_int), parameter :: AFID_NUM_SINGULAR = 2
      integer(c_int), parameter :: AFID_PREC_FP64 = 0, AFID_PREC_MIXED = 1

!     field identifiers (local_arrays, param.F90:79-87)
      integer(c_int), parameter :: AFID_VX = 0, AFID_VY = 1
      integer(c_int), parameter :: AFID_VZ = 2, AFID_TEMP = 3
      integer(c_int), parameter :: AFID_PR = 4, AFID_RUX = 5
      integer(c_int), parameter :: AFID_RUY = 6, AFID_RUZ = 7
      integer(c_int), parameter :: AFID_RUTEMP = 8, AFID_DPH = 9
      integer(c_int), parameter :: AFID_QCAP = 10, AFID_DQ = 11
      integer(c_int), parameter :: AFID_HRO = 12, AFID_RHS = 13

!     bou.in subset (ReadInputFile.F90:16-36), C layout of afid_params
      type, bind(C) :: afid_params_t
        integer(c_int) :: nxm, nym, nzm, nsst, istr3
        real(c_double) :: str3, alx3, ylen, zlen, ray, pra
        integer(c_int) :: inslws, inslwn
        real(c_double) :: tempbp, temptp
        integer(c_int) :: precision
      end type afid_params_t

      interface
        integer(c_int) function afid_create(ctx, p, m)                  &
     &      bind(C, name='afid_create')
          import :: c_int, c_ptr, afid_params_t
          type(c_ptr), intent(out) :: ctx
          type(afid_params_t), intent(in) :: p
          type(c_ptr), value :: m
        end function
        integer(c_int) function afid_destroy(ctx)                       &
     &      bind(C, name='afid_destroy')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        type(c_ptr) function afid_strerror(code)                        &
     &      bind(C, name='afid_strerror')
          import :: c_int, c_ptr
          integer(c_int), value :: code
        end function
        integer(c_int) function afid_get_dims(ctx, out6)                &
     &      bind(C, name='afid_get_dims')
          import :: c_int, c_ptr, c_int64_t
          type(c_ptr), value :: ctx
          integer(c_int64_t), intent(out) :: out6(6)
        end function
        integer(c_int) function afid_get_metric(ctx, name, out, n)      &
     &      bind(C, name='afid_get_metric')
          import :: c_int, c_ptr, c_char, c_double
          type(c_ptr), value :: ctx
          character(kind=c_char), intent(in) :: name(*)
          real(c_double), intent(out) :: out(*)
          integer(c_int), value :: n
        end function
        integer(c_int) function afid_upload(ctx, field, host, ld1, ld2,  &
     &      j0, i0) bind(C, name='afid_upload')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: field, ld1, ld2, j0, i0
          real(c_double), intent(in) :: host(*)
        end function
        integer(c_int) function afid_download(ctx, field, host, ld1,    &
     &      ld2, j0, i0) bind(C, name='afid_download')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: field, ld1, ld2, j0, i0
          real(c_double), intent(inout) :: host(*)
        end function
        integer(c_int) function afid_init_linear_noise(ctx, seed, amp)  &
     &      bind(C, name='afid_init_linear_noise')
          import :: c_int, c_ptr, c_int64_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int64_t), value :: seed
          real(c_double), value :: amp
        end function
        integer(c_int) function afid_set_coefficients(ctx, al, ga, ro,  &
     &      dt, beta) bind(C, name='afid_set_coefficients')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), value :: al, ga, ro, dt, beta
        end function
        integer(c_int) function afid_set_substep(ctx, ns, dt)           &
     &      bind(C, name='afid_set_substep')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: ns
          real(c_double), value :: dt
        end function
        integer(c_int) function afid_explicit_terms_vx(ctx)             &
     &      bind(C, name='afid_explicit_terms_vx')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_explicit_terms_vy(ctx)             &
     &      bind(C, name='afid_explicit_terms_vy')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_explicit_terms_vz(ctx)             &
     &      bind(C, name='afid_explicit_terms_vz')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_explicit_terms_temp(ctx)           &
     &      bind(C, name='afid_explicit_terms_temp')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_and_update_vx(ctx)        &
     &      bind(C, name='afid_implicit_and_update_vx')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_and_update_vy(ctx)        &
     &      bind(C, name='afid_implicit_and_update_vy')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_and_update_vz(ctx)        &
     &      bind(C, name='afid_implicit_and_update_vz')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_implicit_and_update_temp(ctx)      &
     &      bind(C, name='afid_implicit_and_update_temp')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_calc_local_divergence(ctx)         &
     &      bind(C, name='afid_calc_local_divergence')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_solve_pressure_correction(ctx)     &
     &      bind(C, name='afid_solve_pressure_correction')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_correct_velocity(ctx)              &
     &      bind(C, name='afid_correct_velocity')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_correct_pressure(ctx)              &
     &      bind(C, name='afid_correct_pressure')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_update_halo(ctx, field)            &
     &      bind(C, name='afid_update_halo')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: field
        end function
        integer(c_int) function afid_substep(ctx, ns, dt)               &
     &      bind(C, name='afid_substep')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: ns
          real(c_double), value :: dt
        end function
        integer(c_int) function afid_time_marcher(ctx, dt)              &
     &      bind(C, name='afid_time_marcher')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), value :: dt
        end function
        integer(c_int) function afid_step(ctx, ntime, idtv, dt, dtmin,   &
     &      dtmax, limitcfl, dt_used, errorcode)                          &
     &      bind(C, name='afid_step')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: ntime, idtv
          real(c_double), value :: dt, dtmin, dtmax, limitcfl
          real(c_double), intent(out) :: dt_used
          integer(c_int), intent(out) :: errorcode
        end function
        integer(c_int) function afid_get_time(ctx, time)                &
     &      bind(C, name='afid_get_time')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: time
        end function
        integer(c_int) function afid_set_time(ctx, time)                &
     &      bind(C, name='afid_set_time')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), value :: time
        end function
!       continuation files (include/afid_io.h, libafid_io)
        integer(c_int) function afid_write_continuation(ctx, dir)       &
     &      bind(C, name='afid_write_continuation')
          import :: c_int, c_ptr, c_char
          type(c_ptr), value :: ctx
          character(kind=c_char), intent(in) :: dir(*)
        end function
        integer(c_int) function afid_read_continuation(ctx, dir)        &
     &      bind(C, name='afid_read_continuation')
          import :: c_int, c_ptr, c_char
          type(c_ptr), value :: ctx
          character(kind=c_char), intent(in) :: dir(*)
        end function
        integer(c_int) function afid_calc_max_cfl(ctx, cflm)            &
     &      bind(C, name='afid_calc_max_cfl')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: cflm
        end function
        integer(c_int) function afid_check_divergence(ctx, qmax)        &
     &      bind(C, name='afid_check_divergence')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: qmax
        end function
        integer(c_int) function afid_calc_plate_nu(ctx, nub, nut)       &
     &      bind(C, name='afid_calc_plate_nu')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: nub, nut
        end function
        integer(c_int) function afid_global_quantities(ctx, out11)      &
     &      bind(C, name='afid_global_quantities')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: out11(11)
        end function
        integer(c_int) function afid_calc_dissipation_nu(ctx, nu_mu,     &
     &      nu_th, disste, dissth) bind(C, name='afid_calc_dissipation_nu')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: nu_mu, nu_th
          type(c_ptr), value :: disste, dissth
        end function
        integer(c_int) function afid_synchronize(ctx)                   &
     &      bind(C, name='afid_synchronize')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
!       whole main.F90 steps on the device, one captured graph per step
        integer(c_int) function afid_steps(ctx, ntime0, nsteps, idtv,   &
     &      dt, dtmin, dtmax, limitcfl, dt_last, errorcode, steps_done)  &
     &      bind(C, name='afid_steps')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: ntime0, nsteps, idtv
          real(c_double), value :: dt, dtmin, dtmax, limitcfl
          real(c_double), intent(out) :: dt_last
          integer(c_int), intent(out) :: errorcode, steps_done
        end function
        integer(c_int) function afid_last_hip_error(ctx)                &
     &      bind(C, name='afid_last_hip_error')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_get_stale(ctx, plane)              &
     &      bind(C, name='afid_get_stale')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: plane(*)
        end function
        integer(c_int) function afid_set_stale(ctx, plane)              &
     &      bind(C, name='afid_set_stale')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: plane(*)
        end function
        integer(c_int) function afid_get_stats(ctx, out, nsamples)      &
     &      bind(C, name='afid_get_stats')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: out(*)
          integer(c_int), intent(out) :: nsamples
        end function
        integer(c_int) function afid_reset_stats(ctx)                   &
     &      bind(C, name='afid_reset_stats')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function afid_get_slabs(ctx, kslab, nslab, out)  &
     &      bind(C, name='afid_get_slabs')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), intent(in) :: kslab(*)
          integer(c_int), value :: nslab
          real(c_double), intent(out) :: out(*)
        end function
!       concurrent cells: asynchronous steps and interleaved batches
        integer(c_int) function afid_steps_async(ctx, ntime0, nsteps,    &
     &      idtv, dt, dtmin, dtmax, limitcfl)                            &
     &      bind(C, name='afid_steps_async')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: ntime0, nsteps, idtv
          real(c_double), value :: dt, dtmin, dtmax, limitcfl
        end function
        integer(c_int) function afid_steps_more(ctx, nsteps)           &
     &      bind(C, name='afid_steps_more')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), value :: nsteps
        end function
        integer(c_int) function afid_steps_query(ctx, busy)            &
     &      bind(C, name='afid_steps_query')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), intent(out) :: busy
        end function
        integer(c_int) function afid_steps_wait(ctx, dt_last,           &
     &      errorcode, steps_done) bind(C, name='afid_steps_wait')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: dt_last
          integer(c_int), intent(out) :: errorcode, steps_done
        end function
        integer(c_int) function afid_steps_cfl(ctx, cflm)                &
     &      bind(C, name='afid_steps_cfl')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: cflm
        end function
        integer(c_int) function afid_batch_steps(cells, ncells, nsteps) &
     &      bind(C, name='afid_batch_steps')
          import :: c_int, c_ptr
          type(c_ptr), intent(in) :: cells(*)
          integer(c_int), value :: ncells, nsteps
        end function
        integer(c_int) function afid_tune(ctx, zm)                      &
     &      bind(C, name='afid_tune')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
          integer(c_int), intent(out) :: zm
        end function
        integer(c_int) function afid_estimate_bytes(p, nbytes)          &
     &      bind(C, name='afid_estimate_bytes')
          import :: c_int, c_int64_t, afid_params_t
          type(afid_params_t), intent(in) :: p
          integer(c_int64_t), intent(out) :: nbytes
        end function
!       restart regridding (interp.F90)
        integer(c_int) function afid_interp_field(ctx, field, old, nxo,  &
     &      nyo, nzo, istro3, stro3) bind(C, name='afid_interp_field')
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          integer(c_int), value :: field, nxo, nyo, nzo, istro3
          real(c_double), intent(in) :: old(*)
          real(c_double), value :: stro3
        end function
!       statistics and slabs (StatRoutines.F90, SlabDumpRoutines.F90)
        integer(c_int) function afid_calc_stats(ctx)                    &
     &      bind(C, name='afid_calc_stats')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
!       afid_calc_stats without allocation: acc is a device array of
!       9*nxm doubles (no counterpart in the reference)
        integer(c_int) function afid_calc_stats_noalloc(ctx, acc)       &
     &      bind(C, name='afid_calc_stats_noalloc')
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx, acc
        end function
        integer(c_int) function afid_write_stats(ctx, dir, readstats,    &
     &      disscal) bind(C, name='afid_write_stats')
          import :: c_int, c_ptr, c_char
          type(c_ptr), value :: ctx
          character(kind=c_char), intent(in) :: dir(*)
          integer(c_int), value :: readstats, disscal
        end function
        integer(c_int) function afid_slab_levels(xm, nxm, xslab8,        &
     &      kslab9) bind(C, name='afid_slab_levels')
          import :: c_int, c_double
          real(c_double), intent(in) :: xm(*), xslab8(8)
          integer(c_int), value :: nxm
          integer(c_int), intent(out) :: kslab9(9)
        end function
        integer(c_int) function afid_write_slabs(ctx, dir, kslab9, time, &
     &      tout) bind(C, name='afid_write_slabs')
          import :: c_int, c_ptr, c_char, c_double
          type(c_ptr), value :: ctx
          character(kind=c_char), intent(in) :: dir(*)
          integer(c_int), intent(in) :: kslab9(9)
          real(c_double), value :: time, tout
        end function
      end interface

      contains

!     Message of a return code (afid_strerror) as a Fortran string.
      function afid_message(code) result(msg)
        integer(c_int), intent(in) :: code
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: s(:)
        type(c_ptr) :: p
        integer :: n
        p = afid_strerror(code)
        call c_f_pointer(p, s, [256])
        n = 0
        do while (n < 256)
          if (s(n + 1) == c_null_char) exit
          n = n + 1
        end do
        allocate(character(len=n) :: msg)
        msg = transfer(s(1:n), msg)
      end function afid_message

!     Stop on a nonzero return code, naming the routine that failed.
      subroutine afid_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        if (rc /= AFID_OK) then
          write(6,'(a,a,a,i0,a,a)') 'afid: ', where, ' returned ', rc,   &
     &      ': ', afid_message(rc)
          error stop 1
        end if
      end subroutine afid_check

      end module afid_hip
