! Fortran binding of the C ABI in include/cuteflow_cgns.h (iso_c_binding):
! the CGNS mesh reader (in place of lec_mail_cgns, maillage.cuf:398-778) and
! the FlowSolution export (cgns_export_init / cgns_export,
! pre_post_traitement.cuf:223-399) for a Fortran host.
module cuteflow_cgns
  use, intrinsic :: iso_c_binding
  implicit none

  ! struct cf_cgns_info
  type, bind(C) :: cf_cgns_info_t
     integer(c_int) :: nnt, nelt, ndi, ndo, nzones, zone, has_manning, partitioned
  end type cf_cgns_info_t

  interface
     integer(c_int) function cf_cgns_mesh_info(path, info) bind(C, name="cf_cgns_mesh_info")
       import :: c_int, c_char, cf_cgns_info_t
       character(kind=c_char) :: path(*)
       type(cf_cgns_info_t) :: info
     end function cf_cgns_mesh_info

     integer(c_int) function cf_cgns_read_mesh(path, nombre_entrees, inlet_name, nombre_sorties, outlet_name, &
          name_len, is_override_manning, verbose, coords, conn, manning, ndinput, numero_ndinput, ndoutput, &
          numero_ndoutput) bind(C, name="cf_cgns_read_mesh")
       import :: c_int, c_char, c_double
       character(kind=c_char) :: path(*), inlet_name(*), outlet_name(*)
       integer(c_int), value :: nombre_entrees, nombre_sorties, name_len, is_override_manning, verbose
       real(c_double) :: coords(*), manning(*)
       integer(c_int) :: conn(*), ndinput(*), numero_ndinput(*), ndoutput(*), numero_ndoutput(*)
     end function cf_cgns_read_mesh

     integer(c_int) function cf_cgns_export_init(path, nsnap, times) bind(C, name="cf_cgns_export_init")
       import :: c_int, c_char, c_double
       character(kind=c_char) :: path(*)
       integer(c_int), value :: nsnap
       real(c_double) :: times(*)
     end function cf_cgns_export_init

     integer(c_int) function cf_cgns_export(path, k, nelt, u, zm, manning, tolisec, sol_z_offset) &
          bind(C, name="cf_cgns_export")
       import :: c_int, c_char, c_double, c_ptr
       character(kind=c_char) :: path(*)
       integer(c_int), value :: k, nelt
       real(c_double) :: u(*), zm(*)
       type(c_ptr), value :: manning
       real(c_double), value :: tolisec, sol_z_offset
     end function cf_cgns_export

     type(c_ptr) function cf_cgns_error() bind(C, name="cf_cgns_error")
       import :: c_ptr
     end function cf_cgns_error
  end interface

contains

  ! print the library's message and exit(1), like cg_error_exit_f
  subroutine cf_cgns_check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (status /= 0) then
       call c_f_pointer(cf_cgns_error(), msg, [512])
       n = 0
       do while (n < 512)
          if (msg(n + 1) == c_null_char) exit
          n = n + 1
       end do
       write(*, '(a,a,a,i0,a,512a)') 'ierr <- ', what, ' status ', status, ': ', msg(1:n)
       call exit(1)
    end if
  end subroutine cf_cgns_check

  ! byte copy of a file (the reference copies the mesh file to out_<meshfile>
  ! before writing solutions into it, maillage.cuf:427-437)
  subroutine copy_file(src, dst)
    character(len=*), intent(in) :: src, dst
    character(len=1), allocatable :: buf(:)
    integer :: sz, u1, u2
    inquire(file=src, size=sz)
    if (sz < 0) then
       write(*, *) 'cannot read ', trim(src)
       call exit(1)
    end if
    allocate(buf(max(sz, 1)))
    open(newunit=u1, file=src, access='stream', form='unformatted', status='old', action='read')
    if (sz > 0) read(u1) buf(1:sz)
    close(u1)
    open(newunit=u2, file=dst, access='stream', form='unformatted', status='replace', action='write')
    if (sz > 0) write(u2) buf(1:sz)
    close(u2)
  end subroutine copy_file

end module cuteflow_cgns
