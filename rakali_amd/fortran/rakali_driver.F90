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
!!
!! Gauges (&output_nml gauge_file, gauge_output_interval, output_dir): the
!! stations of the gauge file (read_gauge_file, rki_gauges.F90:129-206)
!! are sampled on the device at the driver's cadence (rki_driver.F90:
!! 345-352,470-474) and written at the end to output_dir/gauges.txt, a
!! text table in place of the reference's gauges.nc (no NetCDF here): one
!! "# station name x y" line per station inside the domain, then one row
!! per record of time and eta, u, v of each station.
program rakali_driver
   use, intrinsic :: iso_c_binding, only: c_ptr, c_int, c_double, c_char, c_null_char, c_long_long, &
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
   character(len=4096) :: gauge_file, output_dir
   real(c_double) :: gauge_interval
   character(len=64), allocatable :: gnames(:)
   real(c_double), allocatable :: gx(:), gy(:)
   integer(c_int), allocatable :: gidx(:)
   integer(c_int) :: ng_local

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
   call namelist_output(trim(nml), gauge_file, gauge_interval, output_dir)
   ng_local = 0
   if (len_trim(gauge_file) > 0) then
      call read_gauge_file(trim(gauge_file), gnames, gx, gy)
      if (size(gx) > 0) then
         ierr = rkg_gauges_set(dev, int(size(gx), c_int), gx, gy, ng_local)
         if (ierr /= 0) then
            write (error_unit, "(A,I0,A)") "rkg_gauges_set failed (", ierr, ")"
            stop 1
         end if
         allocate (gidx(max(ng_local, 1)))
         if (ng_local > 0) then
            ierr = rkg_gauges_local(dev, gidx)
            ierr = rkg_gauges_record(dev, gauge_interval, 4096_c_int)
         end if
         write (output_unit, "(A,I0,A,I0,A)") "gauges: ", ng_local, " of ", size(gx), " inside the domain"
      end if
   end if
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
   if (ng_local > 0) call write_gauge_table()
   call rki_destroy(sim)

contains

   !! gauge_file, gauge_output_interval and output_dir of &output_nml
   !! (rki_config.F90:244-246); every key of the group is declared so that
   !! the read accepts any valid file.
   subroutine namelist_output(path, gauge_file, gauge_output_interval, output_dir)
      character(len=*), intent(in) :: path
      character(len=*), intent(out) :: gauge_file, output_dir
      real(c_double), intent(out) :: gauge_output_interval
      logical :: output_to_file, async_output, compress_output, out_h_layer, out_velocity, &
                 out_salinity, out_temperature, out_density, out_w
      character(len=4096) :: restart_file, forcing_file, bathymetry_file
      character(len=16) :: output_format
      real(c_double) :: output_interval, restart_interval, forcing_x_origin, forcing_y_origin
      integer :: async_buffer_slots, compress_level, ios, unit
      namelist /output_nml/ output_to_file, async_output, compress_output, out_h_layer, out_velocity, &
         out_salinity, out_temperature, out_density, out_w, output_dir, restart_file, forcing_file, &
         gauge_file, bathymetry_file, output_format, output_interval, restart_interval, &
         forcing_x_origin, forcing_y_origin, gauge_output_interval, async_buffer_slots, compress_level
      gauge_file = ""
      output_dir = "."
      gauge_output_interval = 0.0_c_double
      open (newunit=unit, file=path, status="old", action="read", iostat=ios)
      if (ios /= 0) return
      read (unit, nml=output_nml, iostat=ios)
      close (unit)
      if (len_trim(output_dir) == 0) output_dir = "."
   end subroutine namelist_output

   !! Stations "name x y", one per line; blank lines and lines starting
   !! with # are skipped, malformed lines are skipped with a warning, and
   !! at most 100 are read (read_gauge_file, rki_gauges.F90:129-206).
   subroutine read_gauge_file(path, names, x, y)
      character(len=*), intent(in) :: path
      character(len=64), allocatable, intent(out) :: names(:)
      real(c_double), allocatable, intent(out) :: x(:), y(:)
      integer, parameter :: max_gauges = 100
      character(len=512) :: line
      character(len=64) :: tn(max_gauges)
      real(c_double) :: tx(max_gauges), ty(max_gauges)
      integer :: unit, ios, n
      n = 0
      open (newunit=unit, file=path, status="old", action="read", iostat=ios)
      if (ios /= 0) then
         write (error_unit, "(A)") "gauges: cannot open "//path
         allocate (names(0), x(0), y(0))
         return
      end if
      do
         read (unit, "(A)", iostat=ios) line
         if (ios /= 0) exit
         line = adjustl(line)
         if (len_trim(line) == 0) cycle
         if (line(1:1) == "#") cycle
         n = n + 1
         if (n > max_gauges) then
            write (error_unit, "(A)") "gauges: maximum of 100 gauges exceeded, ignoring remainder"
            n = max_gauges
            exit
         end if
         read (line, *, iostat=ios) tn(n), tx(n), ty(n)
         if (ios /= 0) then
            write (error_unit, "(A)") "gauges: skipping malformed line: "//trim(line)
            n = n - 1
         end if
      end do
      close (unit)
      allocate (names(n), x(n), y(n))
      names = tn(1:n)
      x = tx(1:n)
      y = ty(1:n)
   end subroutine read_gauge_file

   !! output_dir/gauges.txt: the stations, then time and (eta, u, v) of
   !! each station per record.
   subroutine write_gauge_table()
      integer(c_long_long) :: nrec, got
      real(c_double), allocatable :: tt(:), eta(:, :), uu(:, :), vv(:, :)
      integer :: unit, r, k
      ierr = rkg_gauges_series(dev, 0_c_long_long, nrec, gx, gx, gx, gx)
      allocate (tt(max(nrec, 1_c_long_long)), eta(ng_local, max(nrec, 1_c_long_long)), &
                uu(ng_local, max(nrec, 1_c_long_long)), vv(ng_local, max(nrec, 1_c_long_long)))
      ierr = rkg_gauges_series(dev, nrec, got, tt, eta, uu, vv)
      open (newunit=unit, file=trim(output_dir)//"/gauges.txt", status="replace", action="write")
      do k = 1, ng_local
         write (unit, "(A,1X,A,2ES24.16)") "# station", trim(gnames(gidx(k) + 1)), gx(gidx(k) + 1), &
            gy(gidx(k) + 1)
      end do
      write (unit, "(A)") "# time, then eta u v of each station"
      do r = 1, int(got)
         write (unit, "(ES24.16)", advance="no") tt(r)
         do k = 1, ng_local
            write (unit, "(3ES24.16)", advance="no") eta(k, r), uu(k, r), vv(k, r)
         end do
         write (unit, "(A)") ""
      end do
      close (unit)
   end subroutine write_gauge_table

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
