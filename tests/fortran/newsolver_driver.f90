! Fortran caller of the drop-in, written the way PARIS calls NewSolver
! (paris.f90:499): explicit-shape A(is:ie,js:je,ks:ke,8) and
! p(imin:imax,jmin:jmax,kmin:kmax) with Ng=2 ghost layers, passed straight
! through.  The system is a 32^3 walled box with one dense sphere (density
! ratio 100) assembled with SetupPoisson's rule (modules.f90:2737-2746).
! Usage: newsolver_driver <method 0..3> ; prints PASS/FAIL.  Method 9 calls
! FreeSolver (freesurface.f90:700-808) instead: the x-boundary planes are
! gas cells (pcmask 1) holding P_gas = 0 and 1, the source is zero, and the
! max-norm residual of the liquid cells is checked.  Method 8 calls
! poi_solve (modules.f90:2411), the HYPRE entry, with a PFMG solver type.
! Method 7 assembles the same field on a stretched grid (xform 0.6, the
! spacings of paris.f90:3056-3102) and solves it with the mixed method
! after passing the cell widths (pps_cell_widths).
program newsolver_driver
  use pps_paris
  implicit none
  integer, parameter :: N = 32, Ng = 2
  integer :: is, ie, imin, imax
  real(8), allocatable :: A(:,:,:,:), p(:,:,:), rho(:,:,:), u(:,:,:), v(:,:,:), w(:,:,:)
  integer :: i, j, k, it, ierr, method, bc(6)
  integer, allocatable :: pcmask(:,:,:)
  real(8), allocatable :: pgas(:,:,:)
  real(8) :: rmax
  real(8) :: h, x, y, z, d, c, maxError, res, rms_b, mean8, tres2, seed
  character(len=16) :: arg

  method = PPS_MG_FP64
  if (command_argument_count() >= 1) then
     call get_command_argument(1, arg)
     read(arg, *) method
  end if
  is = Ng + 1; ie = Ng + N; imin = 1; imax = N + 2 * Ng
  allocate(A(is:ie, is:ie, is:ie, 8), p(imin:imax, imin:imax, imin:imax))
  allocate(rho(imin:imax, imin:imax, imin:imax), u(imin:imax, imin:imax, imin:imax))
  allocate(v(imin:imax, imin:imax, imin:imax), w(imin:imax, imin:imax, imin:imax))
  h = 1d0 / N
  seed = 0.5d0
  do k = imin, imax; do j = imin, imax; do i = imin, imax
     x = (i - Ng - 0.5d0) * h; y = (j - Ng - 0.5d0) * h; z = (k - Ng - 0.5d0) * h
     d = sqrt((x - 0.5d0)**2 + (y - 0.5d0)**2 + (z - 0.5d0)**2) - 0.25d0
     c = 0.5d0 * (1d0 - tanh(d / (1.5d0 * h)))
     rho(i, j, k) = c * (100d0 - 1d0) + 1d0
     u(i, j, k) = frand(); v(i, j, k) = frand(); w(i, j, k) = frand()
  end do; end do; end do
  ! walls: no flux through the boundary faces
  u(is - 1, :, :) = 0d0; u(ie, :, :) = 0d0
  v(:, is - 1, :) = 0d0; v(:, ie, :) = 0d0
  w(:, :, is - 1) = 0d0; w(:, :, ie) = 0d0
  do k = is, ie; do j = is, ie; do i = is, ie
     A(i,j,k,1) = merge(0d0, 2d0 / (h * h * (rho(i-1,j,k) + rho(i,j,k))), i == is)
     A(i,j,k,2) = merge(0d0, 2d0 / (h * h * (rho(i+1,j,k) + rho(i,j,k))), i == ie)
     A(i,j,k,3) = merge(0d0, 2d0 / (h * h * (rho(i,j-1,k) + rho(i,j,k))), j == is)
     A(i,j,k,4) = merge(0d0, 2d0 / (h * h * (rho(i,j+1,k) + rho(i,j,k))), j == ie)
     A(i,j,k,5) = merge(0d0, 2d0 / (h * h * (rho(i,j,k-1) + rho(i,j,k))), k == is)
     A(i,j,k,6) = merge(0d0, 2d0 / (h * h * (rho(i,j,k+1) + rho(i,j,k))), k == ie)
     A(i,j,k,7) = sum(A(i,j,k,1:6))
     A(i,j,k,8) = -((u(i,j,k) - u(i-1,j,k)) + (v(i,j,k) - v(i,j-1,k)) + (w(i,j,k) - w(i,j,k-1))) / h
  end do; end do; end do
  mean8 = sum(A(:,:,:,8)) / dble(N)**3
  A(:,:,:,8) = A(:,:,:,8) - mean8
  rms_b = sqrt(sum(A(:,:,:,8)**2) / dble(N)**3)
  maxError = 1d-10 * rms_b
  p = 0d0

  bc = 0
  if (method == 9) then
     call free_case()
     stop
  end if
  if (method == 7) then
     call stretched_case()
     stop
  end if
  call pps_setup(N, N, N, Ng, bc, 1d0, 1d0, 1d0, .true., 3, merge(PPS_MG_FP64, method, method == 8), &
                 ierr)
  if (ierr /= 0) then
     print *, 'FAIL setup', ierr
     stop 1
  end if
  if (method == 8) then
     ! relative tolerance: ||r|| / ||A8|| < 1e-10 is the check below
     call pps_poi_solve(A, p, 1d-10, 100, it, 2)
     ierr = 0
  else
     call pps_newsolver(A, p, maxError, 1.2d0, 4000, it, ierr, 2)
  end if
  ! the reference's own residual (MGsolver.f90:118-124)
  res = 0d0
  do k = is, ie; do j = is, ie; do i = is, ie
     res = res + (A(i,j,k,8) - A(i,j,k,7) * p(i,j,k) + A(i,j,k,1) * p(i-1,j,k) + &
          A(i,j,k,2) * p(i+1,j,k) + A(i,j,k,3) * p(i,j-1,k) + A(i,j,k,4) * p(i,j+1,k) + &
          A(i,j,k,5) * p(i,j,k-1) + A(i,j,k,6) * p(i,j,k+1))**2
  end do; end do; end do
  tres2 = sqrt(res / dble(N)**3)
  call pps_teardown()
  if (ierr == 0 .and. it <= 4000 .and. tres2 < maxError) then
     print '(a,i2,a,i5,a,es10.3)', 'PASS method=', method, ' it=', it, ' rel=', tres2 / rms_b
  else
     print '(a,i2,a,i5,a,i3,a,es10.3)', 'FAIL method=', method, ' it=', it, ' ierr=', ierr, &
          ' rel=', tres2 / rms_b
     stop 1
  end if

contains

  subroutine free_case()
    allocate(pcmask(imin:imax, imin:imax, imin:imax), pgas(imin:imax, imin:imax, imin:imax))
    pcmask = 0; pcmask(is, :, :) = 1; pcmask(ie, :, :) = 1
    pgas = 0d0; pgas(ie, :, :) = 1d0
    A(:,:,:,8) = 0d0
    maxError = 1d-9 * maxval(A(:,:,:,7))
    call pps_setup(N, N, N, Ng, bc, 1d0, 1d0, 1d0, .false., 3, PPS_SOR, ierr)
    call pps_freesolver(A, p, maxError, 1.7d0, 50000, it, ierr, 0, 0d0, tres2, pcmask, pgas, 1)
    call pps_teardown()
    rmax = 0d0
    do k = is, ie; do j = is, ie; do i = is + 1, ie - 1
       rmax = max(rmax, abs(-p(i,j,k) * A(i,j,k,7) + A(i,j,k,1) * p(i-1,j,k) + &
            A(i,j,k,2) * p(i+1,j,k) + A(i,j,k,3) * p(i,j-1,k) + A(i,j,k,4) * p(i,j+1,k) + &
            A(i,j,k,5) * p(i,j,k-1) + A(i,j,k,6) * p(i,j,k+1) + A(i,j,k,8)))
    end do; end do; end do
    if (ierr == 0 .and. it <= 50000 .and. rmax < maxError .and. all(p(is,is:ie,is:ie) == 0d0) &
        .and. all(p(ie,is:ie,is:ie) == 1d0)) then
       print '(a,i6,a,es10.3)', 'PASS FreeSolver it=', it, ' max|r|/maxError=', rmax / maxError
    else
       print '(a,i6,a,i3,a,es10.3)', 'FAIL FreeSolver it=', it, ' ierr=', ierr, ' max|r|=', rmax
       stop 1
    end if
  end subroutine free_case

  subroutine stretched_case()
    real(8) :: xh(imin:imax), xc(imin:imax), dx(imin:imax), dxh(imin:imax), s, vol, sv
    integer :: q
    do q = imin, imax
       s = dble(q - Ng) / dble(N)
       xh(q) = 0.6d0 * s * (0.5d0 - s) * (1d0 - s) + s
    end do
    do q = imin + 1, imax; xc(q) = 0.5d0 * (xh(q) + xh(q - 1)); end do
    xc(imin) = 2d0 * xh(imin) - xc(imin + 1)
    do q = imin, imax - 1; dxh(q) = xc(q + 1) - xc(q); end do
    dxh(imax) = dxh(imax - 1)
    do q = imin + 1, imax; dx(q) = xh(q) - xh(q - 1); end do
    dx(imin) = dx(imin + 1)
    ! SetupPoisson (modules.f90:2737-2746) with the same spacings on all axes
    do k = is, ie; do j = is, ie; do i = is, ie
       A(i,j,k,1) = merge(0d0, 2d0 / (dx(i) * dxh(i-1) * (rho(i-1,j,k) + rho(i,j,k))), i == is)
       A(i,j,k,2) = merge(0d0, 2d0 / (dx(i) * dxh(i) * (rho(i+1,j,k) + rho(i,j,k))), i == ie)
       A(i,j,k,3) = merge(0d0, 2d0 / (dx(j) * dxh(j-1) * (rho(i,j-1,k) + rho(i,j,k))), j == is)
       A(i,j,k,4) = merge(0d0, 2d0 / (dx(j) * dxh(j) * (rho(i,j+1,k) + rho(i,j,k))), j == ie)
       A(i,j,k,5) = merge(0d0, 2d0 / (dx(k) * dxh(k-1) * (rho(i,j,k-1) + rho(i,j,k))), k == is)
       A(i,j,k,6) = merge(0d0, 2d0 / (dx(k) * dxh(k) * (rho(i,j,k+1) + rho(i,j,k))), k == ie)
       A(i,j,k,7) = sum(A(i,j,k,1:6))
       A(i,j,k,8) = -((u(i,j,k) - u(i-1,j,k)) / dx(i) + (v(i,j,k) - v(i,j-1,k)) / dx(j) + &
            (w(i,j,k) - w(i,j,k-1)) / dx(k))
    end do; end do; end do
    ! compatibility on the stretched grid: the volume-weighted mean of A8
    sv = 0d0; vol = 0d0
    do k = is, ie; do j = is, ie; do i = is, ie
       sv = sv + dx(i) * dx(j) * dx(k) * A(i,j,k,8)
       vol = vol + dx(i) * dx(j) * dx(k)
    end do; end do; end do
    A(:,:,:,8) = A(:,:,:,8) - sv / vol
    rms_b = sqrt(sum(A(:,:,:,8)**2) / dble(N)**3)
    maxError = 1d-10 * rms_b
    p = 0d0
    call pps_setup(N, N, N, Ng, bc, 1d0, 1d0, 1d0, .true., 3, PPS_MG_MIXED, ierr)
    call pps_cell_widths(dx(is:ie), dx(is:ie), dx(is:ie), ierr)
    if (ierr /= 0) then
       print *, 'FAIL cell widths', ierr
       stop 1
    end if
    call pps_newsolver(A, p, maxError, 1.2d0, 100, it, ierr, 2)
    call pps_teardown()
    res = 0d0
    do k = is, ie; do j = is, ie; do i = is, ie
       res = res + (A(i,j,k,8) - A(i,j,k,7) * p(i,j,k) + A(i,j,k,1) * p(i-1,j,k) + &
            A(i,j,k,2) * p(i+1,j,k) + A(i,j,k,3) * p(i,j-1,k) + A(i,j,k,4) * p(i,j+1,k) + &
            A(i,j,k,5) * p(i,j,k-1) + A(i,j,k,6) * p(i,j,k+1))**2
    end do; end do; end do
    tres2 = sqrt(res / dble(N)**3)
    if (ierr == 0 .and. it <= 100 .and. tres2 < maxError) then
       print '(a,i5,a,es10.3)', 'PASS stretched it=', it, ' rel=', tres2 / rms_b
    else
       print '(a,i5,a,i3,a,es10.3)', 'FAIL stretched it=', it, ' ierr=', ierr, ' rel=', tres2 / rms_b
       stop 1
    end if
  end subroutine stretched_case

  real(8) function frand()
    ! deterministic sequence in (-1, 1)
    seed = mod(seed * 997d0 + 0.123456789d0, 1d0)
    frand = 2d0 * seed - 1d0
  end function frand

end program newsolver_driver
