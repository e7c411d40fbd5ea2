! Fortran driver of the St-Venant time step on one subdomain: the reference's
! program (main.cuf, solve_finite_volumes in finite_volumes_method.cuf) with
! the per-step kernels replaced by the HIP library through cuteflow_hip.
!
! Keeps the reference's inputs and outputs: the `donnees.f` namelist
! (global_data.cuf:65-94), the text mesh (maillage.cuf:196-396) or, when
! built with the CGNS library (CF_WITH_CGNS), the CGNS mesh
! (maillage.cuf:398-778) with the initial and final FlowSolutions in
! out_<meshfile> (pre_post_traitement.cuf:209-405), the
! boundary_table_<meshfile> file (maillage.cuf:55-79), the initial condition
! file (initial_condition.cuf:82-102) or the dam-line / plane initial
! condition (initial_condition.cuf:4-80), the status block every freqaffich
! steps (finite_volumes_method.cuf:345-368), the final legacy VTK, element
! text and restart files (pre_post_traitement.cuf:167-207, :407-512).
program cuteflow_main
  use, intrinsic :: iso_c_binding
  use cuteflow_hip
#ifdef CF_WITH_CGNS
  use cuteflow_cgns
#endif
  implicit none

  integer, parameter :: dp = c_double

  ! ---- namelist variables and defaults (global_data.cuf:31-62) ----
  integer :: friction = 0, is_override_manning = 0, diffusion = 0
  real(dp) :: gp = 9.81d0, override_manning = 0.d0
  character(len=60) :: meshfile = '', meshfile_path = './'
  integer :: elt_bound = 0, nombre_entrees = 1, nombre_sorties = 1, is_cgns = 0
  character(len=200) :: fich_sol_init = ''
  integer :: solinit = 0, eqlin_barrage = 0, plan = 0, sortie_finale_bluekenue = 0
  real(dp) :: h_amont = 0, u_amont = 0, v_amont = 0, h_aval = 0, u_aval = 0, v_aval = 0
  real(dp) :: xplan = 0, yplan = 0, zplan = 1, xpoint = 0, ypoint = 0, zpoint = 0
  real(dp) :: x1eqbar = 0, x2eqbar = 0, y1eqbar = 0, y2eqbar = 0
  character(len=30) :: inlet = 'inflow', outlet = 'fixedheight', loi_debitglob = ''
  integer :: iflux = 1, ilimiteur = 1, iupwind = 1, fricimplic = 1, debit_var = 0
  integer :: freqaffich = 1000, solrestart = 0
  real(dp) :: sol_z_offset = 0.
  real(dp) :: debitglob(20) = 0, h_sortie(20) = 0
  character(len=100) :: inlet_name(20) = '', outlet_name(20) = ''
  real(dp) :: ts = 1, tsstab = 0, cfl = 0.9d0, tolisec = 1d-5, tol_reg_perm = 0
  real(dp) :: constant_dt = 0.
  real(dp) :: tol = 1e-6, tolaffiche = 1e-2
  integer :: visu_snapshots = 0, solvisu = 0, is_dt_constant = 0

  namelist /donnees_namelist/ gp, is_override_manning, override_manning, is_cgns, meshfile_path, &
       meshfile, elt_bound, nombre_entrees, nombre_sorties, eqlin_barrage, x1eqbar, x2eqbar, &
       y1eqbar, y2eqbar, h_amont, u_amont, v_amont, h_aval, u_aval, v_aval, plan, xplan, yplan, &
       zplan, xpoint, ypoint, zpoint, solinit, fich_sol_init, inlet, inlet_name, debitglob, debit_var, &
       loi_debitglob, outlet, outlet_name, h_sortie, iflux, ilimiteur, iupwind, tolisec, friction, &
       fricimplic, diffusion, ts, cfl, tol_reg_perm, is_dt_constant, constant_dt, freqaffich, &
       tolaffiche, sol_z_offset, solrestart, solvisu, visu_snapshots, sortie_finale_bluekenue

  ! ---- mesh ----
  integer :: nnt, nelt, ndi, ndo, ndw, i, k, n, ios, st, done, nt, nbatch, tc_init
  real(dp), allocatable :: coords(:, :), manning_nd(:), manning(:), u(:, :), zm(:), surf(:), cot(:)
  integer, allocatable :: conn(:, :), ndinput(:), num_in(:), ndoutput(:), num_out(:), ndwall(:)
  integer(c_int), allocatable, target :: conn_c(:), bnd(:)
  real(c_double), allocatable, target :: coords_c(:), mann_c(:)
  real(c_double), allocatable :: dt_hist(:), q_in(:), q_out(:)
  integer(c_int) :: steady
  character(len=1) :: tarati
  character(len=260) :: file_bc
  type(cf_params_t) :: prm
  type(cf_mesh_t) :: msh
  type(c_ptr) :: h
  real(dp) :: cotemin, tc, dt, long_entree(20), x1, x2, x3, y1, y2, y3, v1, v2, hh, alpha, beta
  integer(8) :: c0, c1, crate
  logical :: ex
  character(len=300) :: mesh_path, out_path
  real(dp), allocatable :: tsolvisu(:), unodes(:, :), anodes(:)
  integer :: cptsolvisu
  real(dp) :: tolaff, room
  real(c_double), allocatable, target :: cg_coords(:)
  integer(c_int), allocatable :: cg_conn(:)
  integer :: nsnap
#ifdef CF_WITH_CGNS
  type(cf_cgns_info_t) :: cinfo
#endif

  open(10, file='donnees.f', form='formatted', status='old')
  read(10, nml=donnees_namelist)
  close(10)
  inlet = trim(adjustl(inlet))

  ! ---- mesh: CGNS (maillage.cuf:398-778) or text (maillage.cuf:196-396) ----
  mesh_path = trim(adjustl(meshfile_path))//trim(adjustl(meshfile))
  if (is_cgns == 1) then
#ifdef CF_WITH_CGNS
     st = cf_cgns_mesh_info(trim(mesh_path)//c_null_char, cinfo)
     call cf_cgns_check(st, 'cf_cgns_mesh_info')
     nnt = cinfo%nnt; nelt = cinfo%nelt; ndi = cinfo%ndi; ndo = cinfo%ndo; ndw = 0
     allocate(coords(nnt, 3), manning_nd(nnt), conn(nelt, 3), manning(nelt), ndwall(0))
     allocate(ndinput(max(ndi, 1)), num_in(max(ndi, 1)), ndoutput(max(ndo, 1)), num_out(max(ndo, 1)))
     allocate(cg_coords(3 * nnt), cg_conn(3 * nelt))
     st = cf_cgns_read_mesh(trim(mesh_path)//c_null_char, nombre_entrees, inlet_name, nombre_sorties, &
          outlet_name, len(inlet_name), is_override_manning, 1_c_int, cg_coords, cg_conn, manning, ndinput, &
          num_in, ndoutput, num_out)
     call cf_cgns_check(st, 'cf_cgns_read_mesh')
     do i = 1, nnt
        coords(i, :) = cg_coords(3 * i - 2:3 * i)
     end do
     do i = 1, nelt
        conn(i, :) = cg_conn(3 * i - 2:3 * i)
     end do
     manning_nd = 0
     if (solvisu == 1 .or. solvisu == 3) then
        out_path = 'out_'//trim(adjustl(meshfile))
        call copy_file(trim(mesh_path), trim(out_path))
     end if
#else
     write(*, *) 'is_cgns = 1: this driver was built without the CGNS library (no HDF5 found)'
     call exit(1)
#endif
  else
     open(11, file=trim(adjustl(meshfile_path))//trim(adjustl(meshfile)), form='formatted', status='old')
     read(11, *) tarati
     read(11, *) nnt
     allocate(coords(nnt, 3), manning_nd(nnt))
     do i = 1, nnt
        read(11, *) n, coords(i, 1), coords(i, 2), coords(i, 3), manning_nd(i)
     end do
     read(11, *) tarati
     read(11, *) nelt
     allocate(conn(nelt, 3), manning(nelt))
     do i = 1, nelt
        read(11, *) n, conn(i, 1), conn(i, 2), conn(i, 3), manning(i)
     end do
     read(11, *) tarati
     if (nombre_entrees == 1) then
        read(11, *) ndi
        allocate(ndinput(ndi), num_in(ndi))
        do i = 1, ndi
           read(11, *) ndinput(i)
           num_in(i) = 1
        end do
     else
        read(11, *) ndi, nombre_entrees
        allocate(ndinput(ndi), num_in(ndi))
        do i = 1, ndi
           read(11, *) ndinput(i), num_in(i)
        end do
     end if
     read(11, *) tarati
     if (nombre_sorties == 1) then
        read(11, *) ndo
        allocate(ndoutput(ndo), num_out(ndo))
        do i = 1, ndo
           read(11, *) ndoutput(i)
           num_out(i) = 1
        end do
     else
        read(11, *) ndo, nombre_sorties
        allocate(ndoutput(ndo), num_out(ndo))
        do i = 1, ndo
           read(11, *) ndoutput(i), num_out(i)
        end do
     end if
     read(11, *) tarati
     read(11, *) ndw
     allocate(ndwall(ndw))
     do i = 1, ndw
        read(11, *) ndwall(i)
     end do
     close(11)
  end if
  write(*, *) nnt, ' nodes, ', nelt, ' elems, ', ndi, ' inlet nodes, ', ndo, ' outlet nodes, ', &
       ndw, ' wall nodes'

  ! C layouts: connectivity and coordinates cell-/node-major
  allocate(conn_c(3 * nelt), coords_c(3 * nnt), mann_c(nelt), bnd(6 * nelt))
  do i = 1, nelt
     conn_c(3 * i - 2:3 * i) = conn(i, :)
  end do
  do i = 1, nnt
     coords_c(3 * i - 2:3 * i) = coords(i, :)
  end do
  mann_c = manning

  ! ---- boundary table: build and write, or read (maillage.cuf:42-79) ----
  file_bc = 'boundary_table_'//trim(adjustl(meshfile))
  if (elt_bound == 0) then
     st = cf_build_boundary(nelt, nnt, conn_c, ndi, ndinput, num_in, ndo, ndoutput, num_out, ndw, &
          ndwall, bnd)
     call cf_check(st, 'cf_build_boundary')
     open(12, file=trim(file_bc), status='unknown')
     do i = 1, nelt
        ! the reference compiler's list-directed layout of six integers
        write(12, '(6(1x,i12))') bnd(i), bnd(nelt + i), bnd(2 * nelt + i), bnd(3 * nelt + i), &
             bnd(4 * nelt + i), bnd(5 * nelt + i)
     end do
     close(12)
  else
     open(12, file=trim(file_bc), status='old')
     do i = 1, nelt
        read(12, *) bnd(i), bnd(nelt + i), bnd(2 * nelt + i), bnd(3 * nelt + i), bnd(4 * nelt + i), &
             bnd(5 * nelt + i)
     end do
     close(12)
  end if

  long_entree = 0
  if (nombre_entrees > 0) then
     st = cf_inlet_lengths(nelt, nnt, conn_c, coords_c, bnd, min(nombre_entrees, 20), long_entree)
     call cf_check(st, 'cf_inlet_lengths')
  end if

  ! ---- parameters (data_transfer.cuf:12-96) ----
  prm%gp = gp; prm%tolisec = tolisec; prm%tol = tol; prm%cfl = cfl; prm%ts = ts
  prm%constant_dt = constant_dt; prm%override_manning = override_manning
  prm%tol_reg_perm = tol_reg_perm
  prm%iflux = iflux; prm%ilimiteur = ilimiteur; prm%friction = friction; prm%fricimplic = fricimplic
  if (inlet == 'transm') then
     prm%inlet = 1
  else if (inlet == 'inflow') then
     prm%inlet = 2
  else if (inlet == 'phys_inflow') then
     prm%inlet = 3
  else
     print *, "Bad inlet !"
     error stop
  end if
  if (outlet == 'fixedheight') then
     prm%outlet = 1
  else if (outlet == 'transm') then
     prm%outlet = 2
  else
     print *, "Bad outlet !"
     error stop
  end if
  prm%is_dt_constant = is_dt_constant; prm%is_override_manning = is_override_manning
  prm%nombre_entrees = nombre_entrees; prm%options = 0
  prm%debitglob = debitglob; prm%h_sortie = h_sortie; prm%long_entree = long_entree
  prm%hydro_n = 0; prm%hydro_t = 0; prm%hydro_q = 0

  msh%nelt = nelt; msh%nnt = nnt
  msh%conn = c_loc(conn_c); msh%coords = c_loc(coords_c); msh%boundary = c_loc(bnd)
  msh%manning = c_loc(mann_c)
  st = cf_create(msh, prm, h)
  call cf_check(st, 'cf_create')

  allocate(zm(nelt), surf(nelt), cot(nelt), u(3, nelt))
  st = cf_get_geometry(h, surf, zm, cot, cotemin)
  call cf_check(st, 'cf_get_geometry')
  write(*, *) '======================================================================'
  write(*, *) ' Parameters'
  write(*, *) '======================================================================'
  write(*, *) 'meshfile            = ', meshfile
  write(*, *) 'cotemin             = ', cotemin
  write(*, *) 'iflux               =', iflux
  write(*, *) 'Fricimplic          =', fricimplic
  write(*, *) 'tol                 =', tol
  write(*, *) 'tolisec             =', tolisec
  write(*, *) 'nombre entree       =', nombre_entrees
  do i = 1, nombre_entrees
     write(*, *) 'debitglob       ', i, " =", debitglob(i)
     write(*, *) 'longueur entree ', i, " =", long_entree(i)
  end do

  ! ---- initial condition (initial_condition.cuf) ----
  u = 0
  if (solinit == 1) then
     open(13, file=trim(fich_sol_init), status='old')
     read(13, *) tc_init
     do i = 1, nelt
        read(13, *) u(1, i), u(2, i), u(3, i)
     end do
     close(13)
  else
     do i = 1, nelt
        x1 = coords(conn(i, 1), 1); y1 = coords(conn(i, 1), 2)
        x2 = coords(conn(i, 2), 1); y2 = coords(conn(i, 2), 2)
        x3 = coords(conn(i, 3), 1); y3 = coords(conn(i, 3), 2)
        v1 = (x1 + x2 + x3) / 3
        v2 = (y1 + y2 + y3) / 3
        if (eqlin_barrage == 1) then
           alpha = (y1eqbar - y2eqbar) / (x1eqbar - x2eqbar)
           beta = y1eqbar - alpha * x1eqbar
           if (alpha * v1 - v2 + beta >= 0.00) then
              hh = h_amont - zm(i)
              if (hh > tolisec) then
                 u(:, i) = [hh, hh * u_amont, hh * v_amont]
              else
                 u(:, i) = 0
              end if
           else
              hh = h_aval - zm(i)
              if (hh > tolisec) then
                 u(:, i) = [hh, hh * u_aval, hh * v_aval]
              else
                 u(:, i) = 0
              end if
           end if
        end if
        if (plan == 1) then
           hh = (-xplan * (xpoint - v1) - yplan * (ypoint - v2)) / zplan + zpoint - zm(i)
           if (hh > tolisec) then
              u(:, i) = [hh, hh * u_aval, hh * v_aval]
           else
              u(:, i) = 0
           end if
        end if
     end do
  end if
  st = cf_set_discharge(h, 1_c_int)   ! debit_entree / debit_sortie of the status block
  call cf_check(st, 'cf_set_discharge')
  st = cf_set_state(h, u)
  call cf_check(st, 'cf_set_state')

  ! ---- snapshot times (main.cuf:113-122) and the initial state's snapshot (main.cuf:179) ----
  nsnap = max(visu_snapshots, 1)
  allocate(tsolvisu(0:nsnap), unodes(3, nnt), anodes(nnt))
  do i = 0, nsnap
     tsolvisu(i) = i * ts / real(nsnap)
  end do
  cptsolvisu = 0
  tolaff = tolaffiche
  st = cf_get_time(h, tc, dt)
  call save_sol_visu(tc, dt)

  ! ---- time loop (finite_volumes_method.cuf:113-242): batches up to the next
  ! status print, one step at a time close to a snapshot time ----
  nbatch = max(1, freqaffich)
  allocate(dt_hist(nbatch + 1), q_in(nbatch), q_out(nbatch))
  nt = 0
  write(*, *) 'Loop over time starts'
  call system_clock(c0, crate)
  do
     n = nbatch - mod(nt, nbatch)
     if (solvisu > 0 .and. cptsolvisu <= nsnap) then
        room = tsolvisu(cptsolvisu) - 1.5d0 * dt - tc
        if (room > 0) then
           n = min(n, max(1, int(room / (4.d0 * dt))))
        else
           n = 1
        end if
     end if
     st = cf_step(h, n, dt_hist, done)
     if (st /= 0 .and. st /= CF_ERR_DT) call cf_check(st, 'cf_step')
     k = cf_get_time(h, tc, dt)
     k = cf_get_discharge(h, q_in, q_out, steady)
     call cf_check(k, 'cf_get_discharge')
     nt = nt + done
     ! global_save_sol_visu after each step (finite_volumes_method.cuf:241)
     if (done > 0) call save_sol_visu(tc, dt_hist(done))
     if (steady /= 0) then   ! check_steady_state (finite_volumes_method.cuf:380-383)
        print *, '-----------------------------------------------'
        print *, 'LE REGIME PERMANENT EST ATTEINT'
        print *, '-----------------------------------------------'
     end if
     if (mod(nt, freqaffich) == 0 .and. done > 0) then   ! print_status_in_console (:346-363)
        print *, '-----------------------------------------------'
        write(*, *) 'Iteration number ', nt, ", t = ", tc
        if (nombre_entrees == 1) print *, 'debit_entree ', 1, " = ", q_in(done)
        print *, 'debit_entree_sum =', q_in(done), 'M3/S'
        print *, 'debit_sortie     =', q_out(done), 'M3/S'
        print *, '-----------------------------------------------'
     end if
     if (st == CF_ERR_DT) then
        print *, "Erreur sur le pas de temps dt=", dt
        exit
     end if
     if (done < n) exit
  end do
  call system_clock(c1)
  write(*, *) 'Loop on time ends at iteration : ', nt

  st = cf_get_state(h, u)
  call cf_check(st, 'cf_get_state')

  ! ---- outputs ----
  if (solrestart == 1) then   ! save_sol_for_restart (pre_post_traitement.cuf:498-512)
     open(14, file='solution_elements_restart.txt', status='unknown')
     write(14, *) tc
     do i = 1, nelt
        write(14, '(4f16.6)') u(1, i), u(2, i), u(3, i)
     end do
     close(14)
  end if

  print *, '==========================================================================='
  print *, '=========================   FIN DE LA SIMULATION   ========================'
  print *, '==========================================================================='
  print *, 'DUREE DU CALCUL :', real(c1 - c0) / real(crate), ' Secondes '
  call cf_destroy(h)

contains

  ! global_save_sol_visu (pre_post_traitement.cuf:5-49): snapshot cptsolvisu
  ! is written when tcv is within +-tolaffiche of its time, tolaffiche
  ! becoming 1.5 dt (the last step's) after the first snapshot; a snapshot
  ! time stepped over is skipped with the reference's message.
  subroutine save_sol_visu(tcv, dtv)
    real(dp), intent(in) :: tcv, dtv
    integer :: stv
    if (solvisu <= 0 .or. cptsolvisu > nsnap) return
    if (cptsolvisu > 0) tolaff = 1.5d0 * dtv
    if (tcv <= tsolvisu(cptsolvisu) + tolaff .and. tcv >= tsolvisu(cptsolvisu) - tolaff) then
       stv = cf_get_state(h, u)
       call cf_check(stv, 'cf_get_state')
       call write_snapshot(cptsolvisu, tcv)
       cptsolvisu = cptsolvisu + 1
    else if (tcv > tsolvisu(cptsolvisu) + tolaff) then
       print *, "Skipped save sol ! : The timestep is too large for the number of snapshots to save!", &
            cptsolvisu
       cptsolvisu = cptsolvisu + 1
    end if
  end subroutine save_sol_visu

  ! The files of snapshot ksnap of the state u at time tcv: paraview_export
  ! (solvisu 1, 2), cgns_export (is_cgns, solvisu 1, 3), simple_export of
  ! the nodes (solvisu 1, 4: compute_sol_on_nodes of the current state) and
  ! of the elements (solvisu 1, 5) (pre_post_traitement.cuf:22-41, 167-207).
  subroutine write_snapshot(ksnap, tcv)
    integer, intent(in) :: ksnap
    real(dp), intent(in) :: tcv
    character(len=20) :: kstr
    integer :: ie, jn, nd
    write(kstr, '(I20)') ksnap
    kstr = adjustl(kstr)
    if (solvisu == 1 .or. solvisu == 2) call vtk_export('_solution-Paraview_'//trim(kstr)//'.vtk')
#ifdef CF_WITH_CGNS
    if (is_cgns == 1 .and. (solvisu == 1 .or. solvisu == 3)) then
       if (ksnap == 0) then
          st = cf_cgns_export_init(trim(out_path)//c_null_char, nsnap + 1, tsolvisu)
          call cf_cgns_check(st, 'cf_cgns_export_init')
       end if
       call cgns_snapshot(ksnap)
    end if
#endif
    if (solvisu == 1 .or. solvisu == 4) then
       ! compute_sol_on_nodes (pre_post_traitement.cuf:51-71): area-weighted
       ! average of the cells around each node, accumulated in cell order
       unodes = 0.d0
       anodes = 0.d0
       do ie = 1, nelt
          do jn = 1, 3
             nd = conn(ie, jn)
             unodes(:, nd) = unodes(:, nd) + u(:, ie) * surf(ie)
             anodes(nd) = anodes(nd) + surf(ie)
          end do
       end do
       open(15, file='solution_noeuds_simple_'//trim(kstr)//'.txt', status='unknown')
       write(15, *) tcv
       do nd = 1, nnt
          unodes(:, nd) = unodes(:, nd) / anodes(nd)
          write(15, '(4f16.6)') unodes(1, nd), unodes(1, nd) + coords(nd, 3) - sol_z_offset, unodes(2, nd), &
               unodes(3, nd)
       end do
       close(15)
    end if
    if (solvisu == 1 .or. solvisu == 5) then
       open(15, file='solution_elements_simple_'//trim(kstr)//'.txt', status='unknown')
       write(15, *) tcv
       do ie = 1, nelt
          write(15, '(4f16.6)') u(1, ie), u(1, ie) + zm(ie) - sol_z_offset, u(2, ie), u(3, ie)
       end do
       close(15)
    end if
  end subroutine write_snapshot

#ifdef CF_WITH_CGNS
  ! cgns_export (pre_post_traitement.cuf:287-399) of the current state u; the
  ! Manning field is written when the per-cell n is in use
  subroutine cgns_snapshot(k)
    integer, intent(in) :: k
    type(c_ptr) :: mp
    mp = c_null_ptr
    if (is_override_manning == 0) mp = c_loc(mann_c)
    st = cf_cgns_export(trim(out_path)//c_null_char, k, nelt, u, zm, mp, tolisec, sol_z_offset)
    call cf_cgns_check(st, 'cf_cgns_export')
  end subroutine cgns_snapshot
#endif

  ! paraview_export (pre_post_traitement.cuf:407-496)
  subroutine vtk_export(fname)
    character(len=*), intent(in) :: fname
    integer :: nd, iel
    real(dp) :: solu, solv
    open(16, file=fname, status='unknown')
    write(16, '(a)') "# vtk DataFile Version 1.0"
    write(16, '(a)') "2D Unstructured Grid"
    write(16, '(a)') "ASCII"
    write(16, '(a)') ""
    write(16, '(a)') "DATASET UNSTRUCTURED_GRID"
    write(16, '(a,I10,a)') "POINTS", nnt, " float"
    do nd = 1, nnt
       write(16, '(3f16.6)') coords(nd, 1), coords(nd, 2), 0.
    end do
    write(16, '(a)') ""
    write(16, '(a,I10,I10)') "CELLS", nelt, 4 * nelt
    do iel = 1, nelt
       write(16, '(I1,3I10)') 3, conn(iel, 1) - 1, conn(iel, 2) - 1, conn(iel, 3) - 1
    end do
    write(16, '(a)') ""
    write(16, '(a,I10)') "CELL_TYPES", nelt
    do iel = 1, nelt
       write(16, '(a)') "5"
    end do
    write(16, '(a)') ""
    write(16, '(a,I10)') "CELL_DATA", nelt
    write(16, '(a)') "SCALARS eta float"
    write(16, '(a)') "LOOKUP_TABLE default"
    do iel = 1, nelt
       write(16, '(f16.6)') u(1, iel) + zm(iel) - sol_z_offset
    end do
    write(16, '(a)') ""
    write(16, '(a)') "SCALARS h float"
    write(16, '(a)') "LOOKUP_TABLE default"
    do iel = 1, nelt
       write(16, '(f16.6)') u(1, iel)
    end do
    write(16, '(a)') ""
    write(16, '(a)') "SCALARS b float"
    write(16, '(a)') "LOOKUP_TABLE default"
    do iel = 1, nelt
       write(16, '(f16.6)') zm(iel) - sol_z_offset
    end do
    write(16, '(a)') ""
    write(16, '(a)') "VECTORS velocity float"
    do iel = 1, nelt
       if (u(1, iel) > tolisec) then
          solu = u(2, iel) / u(1, iel)
          solv = u(3, iel) / u(1, iel)
       else
          solu = 0.d0
          solv = 0.d0
       end if
       write(16, '(3f16.6)') solu, solv, 0.
    end do
    write(16, '(a)') ""
    write(16, '(a)') "VECTORS flux float"
    do iel = 1, nelt
       if (u(1, iel) > tolisec) then
          write(16, '(3f16.6)') u(2, iel), u(3, iel), 0.
       else
          write(16, '(3f16.6)') 0., 0., 0.
       end if
    end do
    close(16)
  end subroutine vtk_export

end program cuteflow_main
