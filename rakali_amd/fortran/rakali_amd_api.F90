!! rakali_amd_api -- iso_c_binding interfaces to librakali.so (rki_*,
!! include/rakali.h) and librakali_amd.so (rkg_*, include/rakali_amd.h).
!!
!! The interface blocks follow the pattern the reference uses to reach C
!! from Fortran (tests/test_cross_backend.F90:18-46): bind(c) procedures
!! taking the opaque handle by value and Fortran-ordered FP64 arrays by
!! reference.  Field pointers returned by the getters are mapped onto
!! (nx_total, ny_total) arrays with c_f_pointer.
module rakali_amd_api
   use, intrinsic :: iso_c_binding, only: c_int, c_double, c_ptr, c_char, c_long_long
   implicit none
   private

   public :: rki_create_from_namelist, rki_create_from_string, rki_create, rki_finalize
   public :: rki_destroy, rki_step, rki_advance
   public :: rki_get_time, rki_get_dt, rki_get_step_count, rki_get_grid_info
   public :: rki_get_h_ptr, rki_get_hu_ptr, rki_get_hv_ptr, rki_get_b_ptr
   public :: rki_set_bathymetry, rki_set_initial_h, rki_sync_host, rki_last_error
   public :: rki_get_device_grid, rkg_total_mass
   public :: rkg_gauges_set, rkg_gauges_local, rkg_gauges_record, rkg_gauges_series

   interface
      function rki_create_from_namelist(fname, flen, handle) result(ierr) &
         bind(c, name="rki_create_from_namelist")
         import :: c_char, c_int, c_ptr
         character(kind=c_char), intent(in) :: fname(*)
         integer(c_int), value, intent(in) :: flen
         type(c_ptr), intent(out) :: handle
         integer(c_int) :: ierr
      end function

      function rki_create_from_string(text, tlen, handle) result(ierr) &
         bind(c, name="rki_create_from_string")
         import :: c_char, c_int, c_ptr
         character(kind=c_char), intent(in) :: text(*)
         integer(c_int), value, intent(in) :: tlen
         type(c_ptr), intent(out) :: handle
         integer(c_int) :: ierr
      end function

      function rki_create(fname, flen, handle) result(ierr) bind(c, name="rki_create")
         import :: c_char, c_int, c_ptr
         character(kind=c_char), intent(in) :: fname(*)
         integer(c_int), value, intent(in) :: flen
         type(c_ptr), intent(out) :: handle
         integer(c_int) :: ierr
      end function

      function rki_finalize(handle) result(ierr) bind(c, name="rki_finalize")
         import :: c_ptr, c_int
         type(c_ptr), value, intent(in) :: handle
         integer(c_int) :: ierr
      end function

      subroutine rki_destroy(handle) bind(c, name="rki_destroy")
         import :: c_ptr
         type(c_ptr), intent(inout) :: handle
      end subroutine

      function rki_step(handle, n) result(ierr) bind(c, name="rki_step")
         import :: c_ptr, c_int
         type(c_ptr), value, intent(in) :: handle
         integer(c_int), value, intent(in) :: n
         integer(c_int) :: ierr
      end function

      function rki_advance(handle, target) result(ierr) bind(c, name="rki_advance")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value, intent(in) :: handle
         real(c_double), value, intent(in) :: target
         integer(c_int) :: ierr
      end function

      function rki_get_time(handle) result(t) bind(c, name="rki_get_time")
         import :: c_ptr, c_double
         type(c_ptr), value, intent(in) :: handle
         real(c_double) :: t
      end function

      function rki_get_dt(handle) result(dt) bind(c, name="rki_get_dt")
         import :: c_ptr, c_double
         type(c_ptr), value, intent(in) :: handle
         real(c_double) :: dt
      end function

      function rki_get_step_count(handle) result(n) bind(c, name="rki_get_step_count")
         import :: c_ptr, c_int
         type(c_ptr), value, intent(in) :: handle
         integer(c_int) :: n
      end function

      subroutine rki_get_grid_info(handle, nx, ny, ng, dx, dy) bind(c, name="rki_get_grid_info")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value, intent(in) :: handle
         integer(c_int), intent(out) :: nx, ny, ng
         real(c_double), intent(out) :: dx, dy
      end subroutine

      function rki_get_h_ptr(handle, ptr, nxt, nyt) result(ierr) bind(c, name="rki_get_h_ptr")
         import :: c_ptr, c_int
         type(c_ptr), value, intent(in) :: handle
         type(c_ptr), intent(out) :: ptr
         integer(c_int), intent(out) :: nxt, nyt
         integer(c_int) :: ierr
      end function

      function rki_get_hu_ptr(handle, ptr, nxt, nyt) result(ierr) bind(c, name="rki_get_hu_ptr")
         import :: c_ptr, c_int
         type(c_ptr), value, intent(in) :: handle
         type(c_ptr), intent(out) :: ptr
         integer(c_int), intent(out) :: nxt, nyt
         integer(c_int) :: ierr
      end function

      function rki_get_hv_ptr(handle, ptr, nxt, nyt) result(ierr) bind(c, name="rki_get_hv_ptr")
         import :: c_ptr, c_int
         type(c_ptr), value, intent(in) :: handle
         type(c_ptr), intent(out) :: ptr
         integer(c_int), intent(out) :: nxt, nyt
         integer(c_int) :: ierr
      end function

      function rki_get_b_ptr(handle, ptr, nxt, nyt) result(ierr) bind(c, name="rki_get_b_ptr")
         import :: c_ptr, c_int
         type(c_ptr), value, intent(in) :: handle
         type(c_ptr), intent(out) :: ptr
         integer(c_int), intent(out) :: nxt, nyt
         integer(c_int) :: ierr
      end function

      function rki_set_bathymetry(handle, b, nx, ny) result(ierr) bind(c, name="rki_set_bathymetry")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value, intent(in) :: handle
         integer(c_int), value, intent(in) :: nx, ny
         real(c_double), intent(in) :: b(nx, ny)
         integer(c_int) :: ierr
      end function

      function rki_set_initial_h(handle, h, nx, ny) result(ierr) bind(c, name="rki_set_initial_h")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value, intent(in) :: handle
         integer(c_int), value, intent(in) :: nx, ny
         real(c_double), intent(in) :: h(nx, ny)
         integer(c_int) :: ierr
      end function

      subroutine rki_sync_host(handle) bind(c, name="rki_sync_host")
         import :: c_ptr
         type(c_ptr), value, intent(in) :: handle
      end subroutine

      function rki_last_error() result(msg) bind(c, name="rki_last_error")
         import :: c_ptr
         type(c_ptr) :: msg
      end function

      function rki_get_device_grid(handle) result(grid) bind(c, name="rki_get_device_grid")
         import :: c_ptr
         type(c_ptr), value, intent(in) :: handle
         type(c_ptr) :: grid
      end function

      function rkg_total_mass(grid, mass) result(ierr) bind(c, name="rkg_total_mass")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value, intent(in) :: grid
         real(c_double), intent(out) :: mass
         integer(c_int) :: ierr
      end function

      function rkg_gauges_set(grid, n, x, y, n_local) result(ierr) bind(c, name="rkg_gauges_set")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value, intent(in) :: grid
         integer(c_int), value, intent(in) :: n
         real(c_double), intent(in) :: x(*), y(*)
         integer(c_int), intent(out) :: n_local
         integer(c_int) :: ierr
      end function

      function rkg_gauges_local(grid, local_idx) result(ierr) bind(c, name="rkg_gauges_local")
         import :: c_ptr, c_int
         type(c_ptr), value, intent(in) :: grid
         integer(c_int), intent(out) :: local_idx(*)
         integer(c_int) :: ierr
      end function

      function rkg_gauges_record(grid, interval, capacity) result(ierr) bind(c, name="rkg_gauges_record")
         import :: c_ptr, c_int, c_double
         type(c_ptr), value, intent(in) :: grid
         real(c_double), value, intent(in) :: interval
         integer(c_int), value, intent(in) :: capacity
         integer(c_int) :: ierr
      end function

      function rkg_gauges_series(grid, max_records, n_records, t, eta, u, v) result(ierr) &
         bind(c, name="rkg_gauges_series")
         import :: c_ptr, c_int, c_long_long, c_double
         type(c_ptr), value, intent(in) :: grid
         integer(c_long_long), value, intent(in) :: max_records
         integer(c_long_long), intent(out) :: n_records
         real(c_double), intent(out) :: t(*), eta(*), u(*), v(*)
         integer(c_int) :: ierr
      end function
   end interface
end module rakali_amd_api
