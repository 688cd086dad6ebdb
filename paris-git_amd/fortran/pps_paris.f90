! pps_paris.f90 -- Fortran side of the drop-in: iso_c_binding interfaces to
! libpps (include/pps.h) and wrappers with the reference's argument lists
!   NewSolver(A,p,maxError,beta,maxit,it,ierr,norm)          newsolver.f90:34-57
!   NewSolverMG(A,p,maxError,beta,maxit,it,ierr,norm,tres2)  MGsolver.f90:346-405
!   FreeSolver(A,p,maxError,beta,maxit,it,ierr,iout,time,tres2)
!                                                 freesurface.f90:700-808
!   poi_solve(A,p,maxError,maxit,num_iterations,HYPRESolverType)
!                                                 modules.f90:2411-2593
!   init_MG / finalize_MG                                    MGsolver.f90:34-82
! The module state the reference reads implicitly (grid sizes, Ng,
! bdry_cond, MultiGrid, nrelax, sym_MG) is handed to pps_setup once.
! A and p are passed by reference exactly as the reference passes its
! explicit-shape arrays A(is:ie,js:je,ks:ke,8) and p(imin:imax,...), so the
! call sites need no copies.
module pps_paris
  use iso_c_binding
  implicit none
  private

  integer, parameter, public :: PPS_MG_REFERENCE = 0, PPS_MG_FP64 = 1, &
                                PPS_MG_MIXED = 2, PPS_SOR = 3
  integer, parameter, public :: PPS_OK = 0, PPS_ERR_NOT_CONVERGED = 6

  type, bind(C), public :: pps_grid_t
     integer(c_int) :: nx, ny, nz, ng
     integer(c_int) :: bdry_cond(6)
     real(c_double) :: xlength, ylength, zlength
  end type pps_grid_t

  type, bind(C), public :: pps_options_t
     integer(c_int) :: method, outer, nrelax, sym_mg, nu_pre, nu_post
     integer(c_int) :: check_setup, multigrid, smoother, sweep_rows, sweep_planes
     integer(c_int) :: tail_cells, transfer
  end type pps_options_t

  interface
     integer(c_int) function c_default_options(o) bind(C, name="pps_default_options")
       import :: c_int, pps_options_t
       type(pps_options_t), intent(out) :: o
     end function c_default_options

     integer(c_int) function c_init(g, o, h) bind(C, name="pps_init")
       import :: c_int, c_ptr, pps_grid_t, pps_options_t
       type(pps_grid_t), intent(in) :: g
       type(pps_options_t), intent(in) :: o
       type(c_ptr), intent(out) :: h
     end function c_init

     integer(c_int) function c_finalize(h) bind(C, name="pps_finalize")
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function c_finalize

     integer(c_int) function c_new_solver(h, A, p, maxError, beta, maxit, it, ierr, norm) &
          bind(C, name="pps_new_solver")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: A(*)
       real(c_double), intent(inout) :: p(*)
       real(c_double), value :: maxError, beta
       integer(c_int), value :: maxit, norm
       integer(c_int), intent(out) :: it, ierr
     end function c_new_solver

     integer(c_int) function c_new_solver_mg(h, A, p, maxError, beta, maxit, it, ierr, norm, &
          tres2) bind(C, name="pps_new_solver_mg")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: A(*)
       real(c_double), intent(inout) :: p(*)
       real(c_double), value :: maxError, beta
       integer(c_int), value :: maxit, norm
       integer(c_int), intent(out) :: it, ierr
       real(c_double), intent(out) :: tres2
     end function c_new_solver_mg

     integer(c_int) function c_free_solver(h, A, p, pcmask, P_gas, solver_flag, maxError, beta, &
          maxit, it, ierr, tres2) bind(C, name="pps_free_solver")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: A(*)
       real(c_double), intent(inout) :: p(*)
       integer(c_int), intent(in) :: pcmask(*)
       real(c_double), intent(in) :: P_gas(*)
       integer(c_int), value :: solver_flag, maxit
       real(c_double), value :: maxError, beta
       integer(c_int), intent(out) :: it, ierr
       real(c_double), intent(out) :: tres2
     end function c_free_solver

     integer(c_int) function c_poi_solve(h, A, p, maxError, maxit, num_iterations, solver_type) &
          bind(C, name="pps_poi_solve")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: A(*)
       real(c_double), intent(inout) :: p(*)
       real(c_double), value :: maxError
       integer(c_int), value :: maxit, solver_type
       integer(c_int), intent(out) :: num_iterations
     end function c_poi_solve

     integer(c_int) function c_set_cell_widths(h, dx, dy, dz) bind(C, name="pps_set_cell_widths")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: dx(*), dy(*), dz(*)
     end function c_set_cell_widths

     integer(c_int) function c_last_timing(h, solve_ms, setup_ms) bind(C, name="pps_last_timing")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: solve_ms, setup_ms
     end function c_last_timing
  end interface

  type(c_ptr), save :: handle = c_null_ptr

  public :: pps_setup, pps_teardown, pps_newsolver, pps_newsolvermg, pps_freesolver, &
            pps_poi_solve, pps_timing, pps_cell_widths

contains

  ! init_MG replacement.  method: PPS_MG_REFERENCE / PPS_MG_FP64 /
  ! PPS_MG_MIXED; multigrid = the reference's MultiGrid flag.
  subroutine pps_setup(nx, ny, nz, ng, bdry_cond, xlength, ylength, zlength, &
                       multigrid, nrelax, method, ierr)
    integer, intent(in) :: nx, ny, nz, ng, bdry_cond(6), nrelax, method
    real(8), intent(in) :: xlength, ylength, zlength
    logical, intent(in) :: multigrid
    integer, intent(out) :: ierr
    type(pps_grid_t) :: g
    type(pps_options_t) :: o
    ierr = c_default_options(o)
    g%nx = nx; g%ny = ny; g%nz = nz; g%ng = ng
    g%bdry_cond = bdry_cond
    g%xlength = xlength; g%ylength = ylength; g%zlength = zlength
    o%method = method
    o%nrelax = nrelax
    o%multigrid = merge(1, 0, multigrid)
    if (c_associated(handle)) ierr = c_finalize(handle)
    ierr = c_init(g, o, handle)
  end subroutine pps_setup

  ! finalize_MG replacement
  subroutine pps_teardown()
    integer :: ierr
    if (c_associated(handle)) ierr = c_finalize(handle)
    handle = c_null_ptr
  end subroutine pps_teardown

  ! NewSolver with the reference's arguments
  subroutine pps_newsolver(A, p, maxError, beta, maxit, it, ierr, norm)
    real(8), intent(in) :: A(*)
    real(8), intent(inout) :: p(*)
    real(8), intent(in) :: maxError, beta
    integer, intent(in) :: maxit, norm
    integer, intent(out) :: it, ierr
    integer :: rc
    rc = c_new_solver(handle, A, p, maxError, beta, maxit, it, ierr, norm)
    if (rc == PPS_ERR_NOT_CONVERGED) then
       write(*,*) 'Warning: LinearSolver reached maxit'
    else if (rc /= PPS_OK) then
       ierr = rc
    end if
  end subroutine pps_newsolver

  ! NewSolverMG with the reference's arguments
  subroutine pps_newsolvermg(A, p, maxError, beta, maxit, it, ierr, norm, tres2)
    real(8), intent(in) :: A(*)
    real(8), intent(inout) :: p(*)
    real(8), intent(in) :: maxError, beta
    integer, intent(in) :: maxit, norm
    integer, intent(out) :: it, ierr
    real(8), intent(out) :: tres2
    integer :: rc
    rc = c_new_solver_mg(handle, A, p, maxError, beta, maxit, it, ierr, norm, tres2)
    if (rc /= PPS_OK .and. rc /= PPS_ERR_NOT_CONVERGED) ierr = rc
  end subroutine pps_newsolvermg

  ! FreeSolver with the reference's arguments; the module_freesurface
  ! state it reads (pcmask, P_gas: arrays with p's bounds; solver_flag) is
  ! passed explicitly.  iout and time only name the reference's convergence
  ! file and are not used.
  subroutine pps_freesolver(A, p, maxError, beta, maxit, it, ierr, iout, time, tres2, &
                            pcmask, P_gas, solver_flag)
    real(8), intent(in) :: A(*)
    real(8), intent(inout) :: p(*)
    real(8), intent(in) :: maxError, beta, time
    integer, intent(in) :: maxit, iout, solver_flag
    integer, intent(out) :: it, ierr
    real(8), intent(out) :: tres2
    integer, intent(in) :: pcmask(*)
    real(8), intent(in) :: P_gas(*)
    integer :: rc
    rc = c_free_solver(handle, A, p, pcmask, P_gas, solver_flag, maxError, beta, maxit, it, &
                       ierr, tres2)
    if (rc /= PPS_OK .and. rc /= PPS_ERR_NOT_CONVERGED) ierr = rc
  end subroutine pps_freesolver

  ! poi_solve with the reference's arguments (the HYPRE entry); a failure
  ! prints its status, where the reference would stop inside HYPRE
  subroutine pps_poi_solve(A, p, maxError, maxit, num_iterations, HYPRESolverType)
    real(8), intent(in) :: A(*)
    real(8), intent(inout) :: p(*)
    real(8), intent(in) :: maxError
    integer, intent(in) :: maxit, HYPRESolverType
    integer, intent(out) :: num_iterations
    integer :: rc
    rc = c_poi_solve(handle, A, p, maxError, maxit, num_iterations, HYPRESolverType)
    if (rc /= PPS_OK .and. rc /= PPS_ERR_NOT_CONVERGED) write(*,*) 'poi_solve: status ', rc
  end subroutine pps_poi_solve

  ! Cell widths of a stretched grid, after pps_setup: the caller passes
  ! module_grid's dx(is:ie), dy(js:je), dz(ks:ke) (paris.f90:3056-3102).
  ! The fast methods then weight each row by its cell volume.
  subroutine pps_cell_widths(dx, dy, dz, ierr)
    real(8), intent(in) :: dx(*), dy(*), dz(*)
    integer, intent(out) :: ierr
    ierr = c_set_cell_widths(handle, dx, dy, dz)
  end subroutine pps_cell_widths

  ! milliseconds of the last solve (whole, and the setup part)
  subroutine pps_timing(solve_ms, setup_ms)
    real(8), intent(out) :: solve_ms, setup_ms
    integer :: rc
    rc = c_last_timing(handle, solve_ms, setup_ms)
  end subroutine pps_timing

end module pps_paris
