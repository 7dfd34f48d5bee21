!! rakali_driver -- thin namelist driver over librakali.so.
!!
!!   rakali_driver run.nml [--chunks K] [--out h.bin]
!!
!! Follows the structured branch of the reference's driver
!! (driver_run_coastal, rki_driver.F90:284-522) for the barotropic path:
!! create from the namelist, march to t_end (the loop guard
!! t < t_end - 1e-10 and dt_max = t_end - t of rki_driver.F90:415-476,
!! here inside rki_advance), print a status line at K evenly spaced times,
!! report the throughput as cells x steps / wall time (report_throughput,
!! rki_driver.F90:264-282) and, with --out, write the interior depth as
!! raw FP64 (nx*ny values, x fastest).
program rakali_driver
   use, intrinsic :: iso_c_binding, only: c_ptr, c_int, c_double, c_char, c_null_char, &
                                          c_f_pointer, c_associated
   use, intrinsic :: iso_fortran_env, only: int64, output_unit, error_unit
   use rakali_amd_api
   implicit none

   character(len=4096) :: nml, arg, out_file
   type(c_ptr) :: sim, hptr, dev
   integer(c_int) :: ierr, nx, ny, ng, nxt, nyt
   real(c_double) :: dx, dy, t_end, t, mass
   real(c_double), pointer :: h(:, :)
   integer :: nchunks, k, i, nargs, u
   integer(int64) :: c0, c1, rate
   real(c_double) :: wall, mcells
   integer :: steps

   nargs = command_argument_count()
   if (nargs < 1) then
      write (error_unit, "(A)") "usage: rakali_driver run.nml [--chunks K] [--out h.bin]"
      stop 2
   end if
   call get_command_argument(1, nml)
   nchunks = 10
   out_file = ""
   i = 2
   do while (i <= nargs)
      call get_command_argument(i, arg)
      select case (trim(arg))
      case ("--chunks")
         call get_command_argument(i + 1, arg)
         read (arg, *) nchunks
         i = i + 2
      case ("--out")
         call get_command_argument(i + 1, out_file)
         i = i + 2
      case default
         write (error_unit, "(A)") "unknown argument: "//trim(arg)
         stop 2
      end select
   end do

   ierr = rki_create_from_namelist(trim(nml)//c_null_char, len_trim(nml, kind=c_int), sim)
   if (ierr /= 0) then
      write (error_unit, "(A,I0,A)") "rki_create_from_namelist failed (", ierr, ")"
      stop 1
   end if
   call rki_get_grid_info(sim, nx, ny, ng, dx, dy)
   t_end = namelist_t_end(trim(nml))
   dev = rki_get_device_grid(sim)
   ierr = rkg_total_mass(dev, mass)
   write (output_unit, "(A,I0,A,I0,A,ES14.6)") "grid ", nx, " x ", ny, ", initial mass ", mass
   write (output_unit, "(A9,A14,A14,A13,A13)") "step", "t (s)", "dt (s)", "mass", "wall (s)"

   call system_clock(c0, rate)
   do k = 1, nchunks
      ierr = rki_advance(sim, t_end*real(k, c_double)/real(nchunks, c_double))
      if (ierr /= 0) then
         write (error_unit, "(A,I0,A)") "rki_advance failed (", ierr, ")"
         stop 1
      end if
      ierr = rkg_total_mass(dev, mass)
      call system_clock(c1)
      t = rki_get_time(sim)
      write (output_unit, "(I9,F14.2,F14.4,ES13.3,F13.4)") rki_get_step_count(sim), t, &
         rki_get_dt(sim), mass, real(c1 - c0, c_double)/real(rate, c_double)
   end do
   call system_clock(c1)
   wall = real(c1 - c0, c_double)/real(rate, c_double)
   steps = rki_get_step_count(sim)
   mcells = 0.0_c_double
   if (wall > 0.0_c_double) mcells = real(nx, c_double)*real(ny, c_double)*real(steps, c_double)/wall/1.0e6_c_double
   write (output_unit, "(A,F12.3,A)") "throughput: ", mcells, " Mcells/s"

   if (len_trim(out_file) > 0) then
      ierr = rki_get_h_ptr(sim, hptr, nxt, nyt)
      if (ierr /= 0 .or. .not. c_associated(hptr)) then
         write (error_unit, "(A,I0,A)") "rki_get_h_ptr failed (", ierr, ")"
         stop 1
      end if
      call c_f_pointer(hptr, h, [nxt, nyt])
      open (newunit=u, file=trim(out_file), access="stream", form="unformatted", status="replace")
      write (u) h(ng + 1:ng + nx, ng + 1:ng + ny)
      close (u)
   end if
   call rki_destroy(sim)

contains

   !! t_end of the &time_nml group (the handle API does not report it);
   !! the library has already validated the file.
   function namelist_t_end(path) result(t_end)
      character(len=*), intent(in) :: path
      real(c_double) :: t_end
      real(c_double) :: cfl, dt_max, dt_fixed
      integer :: cfl_interval, ios, unit
      character(len=16) :: time_scheme, time_unit
      namelist /time_nml/ t_end, cfl, dt_max, dt_fixed, cfl_interval, time_scheme, time_unit
      t_end = 1.0_c_double
      time_unit = "s"
      open (newunit=unit, file=path, status="old", action="read", iostat=ios)
      if (ios /= 0) return
      read (unit, nml=time_nml, iostat=ios)
      close (unit)
      select case (trim(time_unit))
      case ("min", "minute", "minutes")
         t_end = t_end*60.0_c_double
      case ("hr", "hour", "hours", "h")
         t_end = t_end*3600.0_c_double
      case ("day", "days", "d")
         t_end = t_end*86400.0_c_double
      case ("year", "years", "yr")
         t_end = t_end*365.25_c_double*86400.0_c_double
      end select
   end function namelist_t_end

end program rakali_driver
