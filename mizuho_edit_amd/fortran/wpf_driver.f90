!>
!! @brief Time loop of program main (wPFAW3D.f90:2172-2325) over the GPU
!! drop-in module: same procedure calls in the same order, same dt rule.
!!
!! Usage: wpf_driver <name>.cfg [nsteps]
!! The loop ends at time_end as in the reference, or after nsteps steps when
!! given.  With gid_output = n in the .cfg, results go to <name>.res with
!! saveResult at the start and every time_total/n (1:2217-2221, 2288-2292).
!! At the end the state is written with saveField to <name>.last.fld and a
!! progress line per step goes to stdout.
program wpf_driver
   use wpfaw3d_gpu
   implicit none
   integer :: itime, nsteps, i, nargs
   real(8) :: time, time_end, taumin, time_res, dtime_res, time_dat, dtime_dat
   character(len=512) :: filename, arg
   type(wpf_diag) :: d

   nargs = command_argument_count()
   if (nargs < 1 .or. nargs > 2) then
      write(*,*) "Error: usage: wpf_driver <name>.cfg [nsteps]"
      call exit(1)
   end if
   call get_command_argument(1, filename)
   nsteps = 0
   if (nargs == 2) then
      call get_command_argument(2, arg)
      read(arg, *) nsteps
   end if
   i = index(filename, ".cfg")
   if (i <= 0) then
      write(*,*) "Error: input file is not a .cfg"
      call exit(1)
   end if

   call loadConfig(filename)

   itime = 0
   time = time_start
   time_end = time_start + time_total
   taumin = 0.0d0
   dtime_res = 0.0d0
   if (n_res_out > 0) then
      dtime_res = time_total/n_res_out
      time_res = time + dtime_res
      call saveResult(filename(1:i-1)//".res", time)
   end if
   dtime_dat = 0.0d0
   if (n_dat_out > 0) then
      dtime_dat = time_total/n_dat_out
      time_dat = time + dtime_dat
      call showProgress(filename(1:i-1)//".dat", time, itime, getPressure(1, 1, 1))
   end if
   do while (time < time_end)
      if (nsteps > 0 .and. itime >= nsteps) exit
      itime = itime + 1
      if (itime >= 3) then
         gparam%dtime = gparam%tfac*0.5d0*gparam%dx**2/gparam%epsilon**2 &
            * min(gparam%epsilon**2/gparam%diffusion, taumin)
      end if
      time = time + gparam%dtime

      call calcPermeability
      call solvePressure
      call calcVelocity
      call calcPhaseDeriv(taumin)
      call calcConcentDeriv

      call evolvePhase
      call evolveConcent

      write(*,'(a,i8,a,es22.15,a,es22.15,a,es22.15,a,i6)') " ITIME", itime, " TIME", time, &
         " DT", gparam%dtime, " TAUMIN", taumin, " ITS", solve_stats%iterations

      if (n_dat_out > 0 .and. time > time_dat) then
         call showProgress(filename(1:i-1)//".dat", time, itime, &
                           getPressure(gparam%ix_ref, gparam%iy_ref, gparam%iz_ref))
         time_dat = time_dat + dtime_dat
      end if
      if (n_res_out > 0 .and. time > time_res) then
         call saveResult(filename(1:i-1)//".res", time)
         time_res = time_res + dtime_res
      end if
   end do
   if (n_res_out > 0) call doneResult()
   if (n_dat_out > 0) call doneProgress()

   call showDiag(d)
   write(*,'(a,4es22.14)') " FINAL pv vcal rwhmax pinj", d%pv, d%vcal, d%rwhmax, d%pinj
   call saveField(filename(1:i-1)//".last.fld", time)
   call freeField()
end program wpf_driver
