! Fortran binding of the C ABI in include/cuteflow_hip.h (iso_c_binding).
! A Fortran host such as the reference's driver (main.cuf /
! finite_volumes_method.cuf) uses these interfaces in place of its
! argument-less CUDA Fortran kernels.
module cuteflow_hip
  use, intrinsic :: iso_c_binding
  implicit none

  integer, parameter :: CF_MAX_IO = 20, CF_MAX_HYDRO = 64
  integer(c_int), parameter :: CF_OK = 0, CF_ERR_ARG = -1, CF_ERR_HIP = -2, CF_ERR_DT = -3, &
       CF_ERR_MESH = -4, CF_ERR_NOMEM = -5
  integer(c_int), parameter :: CF_OPT_KEEP_ORDER = 1, CF_OPT_NO_TILE_SKIP = 2

  ! struct cf_params
  type, bind(C) :: cf_params_t
     real(c_double) :: gp, tolisec, tol, cfl, ts, constant_dt, override_manning, tol_reg_perm
     integer(c_int) :: iflux, ilimiteur, friction, fricimplic, inlet, outlet
     integer(c_int) :: is_dt_constant, is_override_manning, nombre_entrees, options
     real(c_double) :: debitglob(CF_MAX_IO), h_sortie(CF_MAX_IO), long_entree(CF_MAX_IO)
     integer(c_int) :: hydro_n(CF_MAX_IO)
     real(c_double) :: hydro_t(CF_MAX_IO*CF_MAX_HYDRO), hydro_q(CF_MAX_IO*CF_MAX_HYDRO)
  end type cf_params_t

  ! struct cf_mesh
  type, bind(C) :: cf_mesh_t
     integer(c_int) :: nelt, nnt
     type(c_ptr) :: conn, coords, boundary, manning
  end type cf_mesh_t

  interface
     integer(c_int) function cf_build_boundary(nelt, nnt, conn, ndi, ndinput, num_in, ndo, ndoutput, &
          num_out, ndw, ndwall, boundary) bind(C, name="cf_build_boundary")
       import :: c_int
       integer(c_int), value :: nelt, nnt, ndi, ndo, ndw
       integer(c_int) :: conn(*), ndinput(*), num_in(*), ndoutput(*), num_out(*), ndwall(*), boundary(*)
     end function cf_build_boundary

     integer(c_int) function cf_inlet_lengths(nelt, nnt, conn, coords, boundary, nentrees, long_entree) &
          bind(C, name="cf_inlet_lengths")
       import :: c_int, c_double
       integer(c_int), value :: nelt, nnt, nentrees
       integer(c_int) :: conn(*), boundary(*)
       real(c_double) :: coords(*), long_entree(*)
     end function cf_inlet_lengths

     integer(c_int) function cf_create(mesh, params, handle) bind(C, name="cf_create")
       import :: c_int, c_ptr, cf_mesh_t, cf_params_t
       type(cf_mesh_t) :: mesh
       type(cf_params_t) :: params
       type(c_ptr) :: handle
     end function cf_create

     subroutine cf_destroy(handle) bind(C, name="cf_destroy")
       import :: c_ptr
       type(c_ptr), value :: handle
     end subroutine cf_destroy

     type(c_ptr) function cf_error_string(status) bind(C, name="cf_error_string")
       import :: c_ptr, c_int
       integer(c_int), value :: status
     end function cf_error_string

     integer(c_int) function cf_set_state(handle, u) bind(C, name="cf_set_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle
       real(c_double) :: u(*)
     end function cf_set_state

     integer(c_int) function cf_get_state(handle, u) bind(C, name="cf_get_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle
       real(c_double) :: u(*)
     end function cf_get_state

     integer(c_int) function cf_initial_dt(handle, dt) bind(C, name="cf_initial_dt")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle
       real(c_double) :: dt
     end function cf_initial_dt

     integer(c_int) function cf_get_time(handle, tc, dt) bind(C, name="cf_get_time")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle
       real(c_double) :: tc, dt
     end function cf_get_time

     integer(c_int) function cf_step(handle, nsteps, dt_hist, steps_done) bind(C, name="cf_step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle
       integer(c_int), value :: nsteps
       real(c_double) :: dt_hist(*)
       integer(c_int) :: steps_done
     end function cf_step

     integer(c_int) function cf_get_geometry(handle, surf, zm, cotemin_arr, cotemin) &
          bind(C, name="cf_get_geometry")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle
       real(c_double) :: surf(*), zm(*), cotemin_arr(*), cotemin
     end function cf_get_geometry

     integer(c_int) function cf_set_manning(handle, n) bind(C, name="cf_set_manning")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle
       real(c_double) :: n(*)
     end function cf_set_manning

     integer(c_int) function cf_set_discharge(handle, on) bind(C, name="cf_set_discharge")
       import :: c_int, c_ptr
       type(c_ptr), value :: handle
       integer(c_int), value :: on
     end function cf_set_discharge

     integer(c_int) function cf_get_discharge(handle, q_in, q_out, steady) bind(C, name="cf_get_discharge")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: handle
       real(c_double) :: q_in(*), q_out(*)
       integer(c_int) :: steady
     end function cf_get_discharge
  end interface

contains

  ! precision_kind.cuf:13-33 (cuda_glerror) behaviour: print and exit(1)
  subroutine cf_check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (status /= CF_OK) then
       call c_f_pointer(cf_error_string(status), msg, [256])
       n = 0
       do while (n < 256)
          if (msg(n + 1) == c_null_char) exit
          n = n + 1
       end do
       write(*, '(a,a,a,i0,a,256a)') 'ierr <- ', what, ' status ', status, ': ', msg(1:n)
       call exit(1)
    end if
  end subroutine cf_check

end module cuteflow_hip
