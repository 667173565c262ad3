!>
!! @brief GPU drop-in for the per-time-step procedures of module wPFAW3D.
!!
!! The procedures keep the names and argument lists of the reference's
!! module procedures (wPFAW3D.f90:561-1657) and forward to libwpf_hip.so
!! through ISO_C_BINDING.  The fields live on the GPU; the host arrays
!! pressure, phase and concent of this module are refreshed by syncHost()
!! before output routines need them.  A non-zero status from the library
!! stops the program, as the reference stops on a KSPSolve error
!! (wPFAW3D.f90:1005).
module wpfaw3d_gpu
   use, intrinsic :: iso_c_binding
   implicit none
   private

   !> wpf_params of include/wpf.h (subset of Param_type, wPFAW3D.f90:27-76)
   type, bind(C), public :: wpf_params
      integer(c_int32_t) :: nx, ny, nz, ix_ref, iy_ref, iz_ref
      real(c_double) :: dtime, tfac, epsilon, dx, lambda, porosity, rhos, rhoa, viscosity
      real(c_double) :: cwtp, diffusion, b100, ral, rat, permeab_hole, alpha, pini
   end type wpf_params

   !> wpf_well of include/wpf.h (Well_type, wPFAW3D.f90:82-90)
   type, bind(C), public :: wpf_well
      integer(c_int32_t) :: ix_g, iy_g, iz_g, pad
      real(c_double) :: pressure, injection, phase, concent
   end type wpf_well

   type, bind(C), public :: wpf_solver_opts
      integer(c_int32_t) :: solver, maxit
      real(c_double) :: rtol, atol
      integer(c_int32_t) :: mg_levels, mg_sweeps
      real(c_double) :: mg_omega, mg_coarse_scale, inner_rtol
      integer(c_int32_t) :: inner_maxit, flags
   end type wpf_solver_opts

   type, bind(C), public :: wpf_solve_stats
      integer(c_int32_t) :: iterations, outer, converged, learned
      real(c_double) :: rnorm0, rnorm, bnorm, seconds
   end type wpf_solve_stats

   type, bind(C), public :: wpf_diag
      real(c_double) :: pv, vcal, rwhmax, phase_max, phase_min, concent_max, concent_min, pinj
   end type wpf_diag

   integer(c_int32_t), parameter, public :: WPF_FIELD_PERMEAB_ROCK = 0, WPF_FIELD_PRESSURE = 1, &
      WPF_FIELD_PHASE = 2, WPF_FIELD_CONCENT = 3, WPF_FIELD_DPHASEDT = 4, WPF_FIELD_DCONCENTDT = 5, &
      WPF_FIELD_PERMEAB = 6
   integer(c_int32_t), parameter :: WPF_OK = 0, WPF_ERR_NOCONV = 3

   interface
      integer(c_int32_t) function wpf_config_load(path, p, nwell, tstart, ttotal) bind(C, name="wpf_config_load")
         import :: c_char, c_int32_t, c_double, wpf_params
         character(kind=c_char), intent(in) :: path(*)
         type(wpf_params), intent(out) :: p
         integer(c_int32_t), intent(out) :: nwell
         real(c_double), intent(out) :: tstart, ttotal
      end function
      integer(c_int32_t) function wpf_config_wells(w) bind(C, name="wpf_config_wells")
         import :: c_int32_t, wpf_well
         type(wpf_well), intent(out) :: w(*)
      end function
      integer(c_int32_t) function wpf_config_fields(lw, kr, ph, cc) bind(C, name="wpf_config_fields")
         import :: c_int32_t, c_double
         integer(c_int32_t), intent(out) :: lw(*)
         real(c_double), intent(out) :: kr(*), ph(*), cc(*)
      end function
      integer(c_int32_t) function wpf_field_write(path, t, p, kr, ph, cc) bind(C, name="wpf_field_write")
         import :: c_char, c_int32_t, c_double, wpf_params
         character(kind=c_char), intent(in) :: path(*)
         real(c_double), value :: t
         type(wpf_params), intent(in) :: p
         real(c_double), intent(in) :: kr(*), ph(*), cc(*)
      end function
      integer(c_int32_t) function wpf_config_output(nres, ndat) bind(C, name="wpf_config_output")
         import :: c_int32_t
         integer(c_int32_t), intent(out) :: nres, ndat
      end function
      integer(c_int32_t) function wpf_result_write(path, t, p, first, pr, ph, cc, k0) &
            bind(C, name="wpf_result_write")
         import :: c_char, c_int32_t, c_double, wpf_params
         character(kind=c_char), intent(in) :: path(*)
         real(c_double), value :: t
         type(wpf_params), intent(in) :: p
         integer(c_int32_t), value :: first
         real(c_double), intent(in) :: pr(*), ph(*), cc(*), k0(*)
      end function
      integer(c_int32_t) function wpf_create(p, lw, w, nwell, dev, ctx) bind(C, name="wpf_create")
         import :: c_int32_t, c_ptr, wpf_params, wpf_well
         type(wpf_params), intent(in) :: p
         integer(c_int32_t), intent(in) :: lw(*)
         type(wpf_well), intent(in) :: w(*)
         integer(c_int32_t), value :: nwell, dev
         type(c_ptr), intent(out) :: ctx
      end function
      integer(c_int32_t) function wpf_destroy(ctx) bind(C, name="wpf_destroy")
         import :: c_int32_t, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int32_t) function wpf_upload(ctx, f, a) bind(C, name="wpf_upload")
         import :: c_int32_t, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: f
         real(c_double), intent(in) :: a(*)
      end function
      integer(c_int32_t) function wpf_download(ctx, f, a) bind(C, name="wpf_download")
         import :: c_int32_t, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: f
         real(c_double), intent(out) :: a(*)
      end function
      integer(c_int32_t) function wpf_calc_permeability(ctx) bind(C, name="wpf_calc_permeability")
         import :: c_int32_t, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int32_t) function wpf_solve_pressure(ctx, o, s) bind(C, name="wpf_solve_pressure")
         import :: c_int32_t, c_ptr, wpf_solver_opts, wpf_solve_stats
         type(c_ptr), value :: ctx
         type(wpf_solver_opts), intent(in) :: o
         type(wpf_solve_stats), intent(out) :: s
      end function
      integer(c_int32_t) function wpf_calc_velocity(ctx) bind(C, name="wpf_calc_velocity")
         import :: c_int32_t, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int32_t) function wpf_calc_phase_deriv(ctx, taumin) bind(C, name="wpf_calc_phase_deriv")
         import :: c_int32_t, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: taumin
      end function
      integer(c_int32_t) function wpf_calc_concent_deriv(ctx) bind(C, name="wpf_calc_concent_deriv")
         import :: c_int32_t, c_ptr
         type(c_ptr), value :: ctx
      end function
      integer(c_int32_t) function wpf_evolve_phase(ctx, dt) bind(C, name="wpf_evolve_phase")
         import :: c_int32_t, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: dt
      end function
      integer(c_int32_t) function wpf_evolve_concent(ctx, dt) bind(C, name="wpf_evolve_concent")
         import :: c_int32_t, c_ptr, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: dt
      end function
      integer(c_int32_t) function wpf_get_pressure(ctx, i, j, k, p) bind(C, name="wpf_get_pressure")
         import :: c_int32_t, c_ptr, c_double
         type(c_ptr), value :: ctx
         integer(c_int32_t), value :: i, j, k
         real(c_double), intent(out) :: p
      end function
      integer(c_int32_t) function wpf_diagnostics(ctx, d) bind(C, name="wpf_diagnostics")
         import :: c_int32_t, c_ptr, wpf_diag
         type(c_ptr), value :: ctx
         type(wpf_diag), intent(out) :: d
      end function
      type(c_ptr) function wpf_last_error() bind(C, name="wpf_last_error")
         import :: c_ptr
      end function
   end interface

   !> module state (the parts of wPFAW3D's state the driver reads)
   type(wpf_params), public :: gparam
   type(wpf_solver_opts), public :: solver
   type(wpf_solve_stats), public :: solve_stats
   type(wpf_well), allocatable, public :: well(:)
   ! param%nwell: the number of wells loaded (well(:) holds at least one
   ! record, so loops over the wells run to nwell, not size(well))
   integer(c_int32_t), public :: nwell = 0
   integer(c_int32_t), allocatable, public :: lw(:,:,:)
   real(c_double), allocatable, public :: permeab_rock(:,:,:), pressure(:,:,:), phase(:,:,:), concent(:,:,:)
   real(c_double), public :: time_start = 0.0d0, time_total = 0.0d0
   ! gid_output and scr_output of the .cfg (0 when absent)
   integer(c_int32_t), public :: n_res_out = 0, n_dat_out = 0
   ! the .dat file of showProgress and its start clocks
   integer :: progress_unit = 0
   logical :: progress_open = .false.  ! (newunit= numbers are negative)
   integer(8) :: progress_c0 = 0, progress_rate = 1
   real(c_double) :: progress_cpu0 = 0.0d0
   type(c_ptr) :: ctx = c_null_ptr

   public :: loadConfig, syncHost, saveField, saveResult, doneResult, freeField, showDiag, getPressure
   public :: showProgress, doneProgress
   public :: calcPermeability, solvePressure, calcVelocity, calcPhaseDeriv, calcConcentDeriv
   public :: evolvePhase, evolveConcent, updateHalo

contains

   subroutine check(ierr, where)
      integer(c_int32_t), intent(in) :: ierr
      character(len=*), intent(in) :: where
      character(kind=c_char), pointer :: msg(:)
      integer :: n
      if (ierr == WPF_OK) return
      call c_f_pointer(wpf_last_error(), msg, [512])
      n = 0
      do while (n < 512)
         if (msg(n + 1) == c_null_char) exit
         n = n + 1
      end do
      write(*,'(a,a,i3,a,512a)') "ERROR ", where//": status", ierr, ": ", msg(1:n)
      error stop 1
   end subroutine check

   !> loadConfig (wPFAW3D.f90:264-505) through the native reader, then the
   !! GPU context and the upload of the loaded state.
   subroutine loadConfig(filename)
      character(len=*), intent(in) :: filename
      call check(wpf_config_load(trim(filename)//c_null_char, gparam, nwell, time_start, time_total), &
                 "loadConfig")
      call check(wpf_config_output(n_res_out, n_dat_out), "loadConfig(output)")
      allocate(well(max(1, nwell)))
      call check(wpf_config_wells(well), "loadConfig(wells)")
      allocate(lw(0:gparam%nx+1, 0:gparam%ny+1, 0:gparam%nz+1))
      allocate(permeab_rock(0:gparam%nx+1, 0:gparam%ny+1, 0:gparam%nz+1))
      allocate(pressure(0:gparam%nx+1, 0:gparam%ny+1, 0:gparam%nz+1))
      allocate(phase(0:gparam%nx+1, 0:gparam%ny+1, 0:gparam%nz+1))
      allocate(concent(0:gparam%nx+1, 0:gparam%ny+1, 0:gparam%nz+1))
      call check(wpf_config_fields(lw, permeab_rock, phase, concent), "loadConfig(fields)")
      pressure(:,:,:) = gparam%pini
      call check(wpf_create(gparam, lw, well, nwell, -1_c_int32_t, ctx), "wpf_create")
      call check(wpf_upload(ctx, WPF_FIELD_PERMEAB_ROCK, permeab_rock), "upload(permeab_rock)")
      call check(wpf_upload(ctx, WPF_FIELD_PHASE, phase), "upload(phase)")
      call check(wpf_upload(ctx, WPF_FIELD_CONCENT, concent), "upload(concent)")
      ! default solver: FP64 CG + multigrid, the reference's rtol and maxit (1:989-997)
      solver = wpf_solver_opts(1, 1000, 1.0d-10, 1.0d-50, 5, 2, 0.8d0, 2.0d0, 1.0d-4, 100, 0)
   end subroutine loadConfig

   !> refresh the host copies of pressure, phase and concent
   subroutine syncHost()
      call check(wpf_download(ctx, WPF_FIELD_PRESSURE, pressure), "download(pressure)")
      call check(wpf_download(ctx, WPF_FIELD_PHASE, phase), "download(phase)")
      call check(wpf_download(ctx, WPF_FIELD_CONCENT, concent), "download(concent)")
   end subroutine syncHost

   !> saveField (wPFAW3D.f90:2096-2141) from the current device state
   subroutine saveField(filename, time)
      character(len=*), intent(in) :: filename
      real(c_double), intent(in) :: time
      call syncHost()
      call check(wpf_field_write(trim(filename)//c_null_char, time, gparam, permeab_rock, phase, &
                                 concent), "saveField")
   end subroutine saveField

   !> saveResult (wPFAW3D.f90:1842-2091) from the current device state: the
   !> first call writes the GiD mesh, every call appends the four results
   subroutine saveResult(filename, time)
      character(len=*), intent(in) :: filename
      real(c_double), intent(in) :: time
      logical, save :: firsttime = .true.
      real(c_double), allocatable :: perm(:,:,:,:), k0(:,:,:)
      call syncHost()
      allocate(perm(0:3, 0:gparam%nx+1, 0:gparam%ny+1, 0:gparam%nz+1))
      call check(wpf_download(ctx, WPF_FIELD_PERMEAB, perm), "download(permeab)")
      k0 = perm(0,:,:,:)
      call check(wpf_result_write(trim(filename)//c_null_char, time, gparam, &
                                  merge(1_c_int32_t, 0_c_int32_t, firsttime), pressure, phase, concent, k0), &
                 "saveResult")
      firsttime = .false.
      deallocate(perm, k0)
   end subroutine saveResult

   !> doneResult: the result file is closed after every saveResult
   subroutine doneResult()
   end subroutine doneResult

   subroutine freeField()
      integer(c_int32_t) :: ierr
      if (c_associated(ctx)) ierr = wpf_destroy(ctx)
      ctx = c_null_ptr
      if (allocated(well)) deallocate(well, lw, permeab_rock, pressure, phase, concent)
   end subroutine freeField

   subroutine showDiag(d)
      type(wpf_diag), intent(out) :: d
      call check(wpf_diagnostics(ctx, d), "showProgress")
   end subroutine showDiag

   !> showProgress (wPFAW3D.f90:1665-1810): the .dat file (input summary on
   !> the first call, then one line per call) and the progress line on
   !> stdout, from the device reductions of wpf_diagnostics
   subroutine showProgress(filename, time, itime, pinj)
      character(len=*), intent(in) :: filename
      real(c_double), intent(in) :: time, pinj
      integer, intent(in) :: itime
      logical, save :: firsttime = .true.
      real(c_double), save :: pv
      real(c_double) :: xchl, qinj, vacid, vcaco3
      type(wpf_diag) :: d
      integer :: iw
      call check(wpf_diagnostics(ctx, d), "showProgress")
      xchl = gparam%cwtp*gparam%b100*gparam%rhoa/gparam%rhos
      if (firsttime) then
         open(newunit=progress_unit, file=filename, status="replace")
         progress_open = .true.
         call system_clock(progress_c0, progress_rate)
         call cpu_time(progress_cpu0)
         pv = d%pv
         call progress_header(pv, xchl)
         firsttime = .false.
      end if
      qinj = 0.0d0
      do iw = 1, nwell
         if (well(iw)%injection /= 0.0d0) qinj = qinj + well(iw)%injection
      end do
      vacid = qinj*time
      vcaco3 = qinj*time*xchl
      write(progress_unit, '(1x,i8,1x,7es20.10)') itime, time, vacid/pv, pinj, d%rwhmax, vacid, vcaco3, d%vcal
      write(*, "(i8,es13.5,es11.3,1x,4(f8.3),es13.5,es12.4,es10.2)") itime, time, vacid/pv, &
         d%phase_max, d%phase_min, d%concent_max, d%concent_min, pinj, d%rwhmax, gparam%dtime
   end subroutine showProgress

   subroutine progress_header(pv, xchl)
      real(c_double), intent(in) :: pv, xchl
      real(c_double) :: b, kc
      integer :: u
      u = progress_unit
      b = gparam%rhoa/gparam%rhos
      kc = gparam%cwtp*gparam%b100/(1.0d0 - gparam%porosity)
      write(u,*) "***** CALCULATION SUMMARY FOR PFACID PROGRAM *****"
      write(u,*) "/// Input data ///"
      write(u,*) "*** Grid data"
      write(u,"(3x,'Num of grids, nx,ny,nz =',3i6)") gparam%nx, gparam%ny, gparam%nz
      write(u,"(3x,'Grid size, dx       =',es14.6,' [m]')") gparam%dx
      write(u,"(3x,'Core length, Lx, Ly, Lz =',3f10.4,' [m]')") gparam%dx*gparam%nx, gparam%dx*gparam%ny, &
         gparam%dx*gparam%nz
      write(u,*) "*** Rock and fluid data"
      write(u,"(3x,'acid con., cwtp      =',f12.4,' [kg HCl/kg solution]')") gparam%cwtp
      write(u,"(3x,'acid density, rhoa   =',f12.1,' [kg/m3]')") gparam%rhoa
      write(u,"(3x,'rock density, rhos   =',f12.1,' [kg/m3]')") gparam%rhos
      write(u,"(3x,'viscosity, visco     =',es14.6,' [Pa*s]')") gparam%viscosity
      write(u,"(3x,'rock porosity, poro  =',f12.6,' [dec]')") gparam%porosity
      write(u,"(3x,'diffusivity coef. ,d =',es12.4,' [m2/s]')") gparam%diffusion
      write(u,"(3x,'gravimetric dissolving power, b100 =',f8.3,' [kg CaCO3/kg HCl]')") gparam%b100
      write(u,"(3x,'volumetric dissolving power, Xacid =',f8.3,' [m3 CaCO3/m3 HCl]')") xchl
      write(u,*) "*** Phase-field related parameters"
      write(u,"(3x,'lambda =',f12.6,' [-]')") gparam%lambda
      write(u,"(3x,'kc     =',f12.6,' [-]')") kc
      write(u,"(3x,'b      =',f12.6,' [-]')") b
      write(u,"(3x,'alpha  =',f12.6,' [-]')") gparam%alpha
      write(u,*) "*** Injection data"
      write(u,"(3x,'time step size, dt     =',es12.4,' [s]')") gparam%dtime
      write(u,"(3x,'initial pressure, pini =',es12.4,' [m/s]')") gparam%pini
      write(u,"(1x,'pore volume =',es20.6,' [m3]')") pv
      write(u,"(//'/// SIMULATION RESULTS ///')")
      write(u,"(13x,'ITIME',2x,'TIME',15x,'PVINJ',13x,'WELL_PRESS'," // &
         "10x,'WH_LENGTH',10x,'INJ_ACID',11x,'DISSOLVED_CaCO3',10x,'DISS.CaCO3_c')")
      write(u,"(13x,'(-)',4x,'(sec)',15x,'(v/v)',15x,'(Pa)'," // &
         "16x,'(m)',14x,'(m3/m)',14x,'(m3/m)',14x,'(m3/m)')")
      write(u,"(5x,'-----',2x,6('   --------------   '))")
   end subroutine progress_header

   !> doneProgress (wPFAW3D.f90:1812-1825): the run times, then close
   subroutine doneProgress()
      integer(8) :: c1
      real(c_double) :: cpu1
      if (.not. progress_open) return
      call system_clock(c1)
      call cpu_time(cpu1)
      write(progress_unit, "(/'CPU time =',es12.4,' [min]')") (cpu1 - progress_cpu0)/60.0d0
      write(progress_unit, "('Cal time =',es12.4,' [min]')") dble(c1 - progress_c0)/progress_rate/60.0d0
      close(progress_unit)
      progress_open = .false.
   end subroutine doneProgress

   !> getPressure (wPFAW3D.f90:2151-2166)
   function getPressure(ix_g, iy_g, iz_g) result(p)
      integer, intent(in) :: ix_g, iy_g, iz_g
      real(c_double) :: p
      call check(wpf_get_pressure(ctx, int(ix_g, c_int32_t), int(iy_g, c_int32_t), &
                                  int(iz_g, c_int32_t), p), "getPressure")
   end function getPressure

   !> calcPermeability (wPFAW3D.f90:561-599)
   subroutine calcPermeability
      call check(wpf_calc_permeability(ctx), "calcPermeability")
   end subroutine calcPermeability

   !> solvePressure (wPFAW3D.f90:603-1160); maxit without convergence is
   !! reported and the run goes on, as after KSPSolve
   subroutine solvePressure
      integer(c_int32_t) :: ierr
      ierr = wpf_solve_pressure(ctx, solver, solve_stats)
      if (ierr == WPF_ERR_NOCONV) then
         write(*,'(a,i6,a,es12.4)') " solvePressure: no convergence in", solve_stats%iterations, &
            " iterations, residual", solve_stats%rnorm / max(solve_stats%bnorm, tiny(1.0d0))
         return
      end if
      call check(ierr, "solvePressure")
   end subroutine solvePressure

   !> calcVelocity (wPFAW3D.f90:1164-1193)
   subroutine calcVelocity
      call check(wpf_calc_velocity(ctx), "calcVelocity")
   end subroutine calcVelocity

   !> calcPhaseDeriv (wPFAW3D.f90:1197-1475)
   subroutine calcPhaseDeriv(taumin)
      real(8), intent(out) :: taumin
      call check(wpf_calc_phase_deriv(ctx, taumin), "calcPhaseDeriv")
   end subroutine calcPhaseDeriv

   !> calcConcentDeriv (wPFAW3D.f90:1554-1624)
   subroutine calcConcentDeriv
      call check(wpf_calc_concent_deriv(ctx), "calcConcentDeriv")
   end subroutine calcConcentDeriv

   !> evolvePhase (wPFAW3D.f90:1628-1632), step size gparam%dtime
   subroutine evolvePhase
      call check(wpf_evolve_phase(ctx, gparam%dtime), "evolvePhase")
   end subroutine evolvePhase

   !> evolveConcent (wPFAW3D.f90:1636-1640), step size gparam%dtime
   subroutine evolveConcent
      call check(wpf_evolve_concent(ctx, gparam%dtime), "evolveConcent")
   end subroutine evolveConcent

   !> updateHalo (wPFAW3D.f90:1645-1657): device fields need nothing (the
   !! kernels clamp); a host array gets the single-rank clamp copies.
   subroutine updateHalo(field)
      real(8), intent(inout) :: field(0:, 0:, 0:)
      integer :: nx, ny, nz
      nx = size(field, 1) - 2
      ny = size(field, 2) - 2
      nz = size(field, 3) - 2
      field(0, :, :) = field(1, :, :)
      field(nx + 1, :, :) = field(nx, :, :)
      field(:, 0, :) = field(:, 1, :)
      field(:, ny + 1, :) = field(:, ny, :)
      field(:, :, 0) = field(:, :, 1)
      field(:, :, nz + 1) = field(:, :, nz)
   end subroutine updateHalo

end module wpfaw3d_gpu
