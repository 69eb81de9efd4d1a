! driver.f90 -- a Fortran driver in the shape of the reference's frequency loop
! (MoVFEM_3DMT.f90:62-143): per frequency it updates omega/sigma, calls
! global_vfem, removes exact zeros and solves, with every step running on the
! GPU through movfem_gpu (ISO_C_BINDING).  Odd frequencies call global_vfem
! and the zero removal separately, even ones the fused mvf_global_vfem_nz.
!
! Model: 8-node elements, 20 x 20 x (16 earth + 4 air) cells of 500 m, a
! 0.01 S/m half-space holding a 1 S/m 2 x 2 x 1 km block, Dirichlet
! boundaries; frequencies from the command line (default 1 Hz).
! Prints one line per frequency: nne, nnz after zero removal, iterations and
! the FP64 true residuals of both polarizations.
program movfem_driver
  use iso_c_binding
  use movfem_gpu
  implicit none
  integer, parameter :: nx = 20, ny = 20, nze = 16, nza = 4, nz = nze + nza
  real(c_double), parameter :: h = 500.d0, pi = 3.1415926535897932384626433d0, eps = 8.854187817d-12
  real(c_double), target, allocatable :: xp(:), yp(:), zp(:), mu(:, :)
  complex(c_double_complex), target, allocatable :: sigma(:, :), a(:), rhs(:)
  integer(c_int32_t), allocatable :: ia(:), ja(:)
  type(mvf_grid) :: grid
  type(mvf_model) :: model
  type(mvf_bc) :: bc
  type(mvf_options) :: opts
  type(mvf_solver_opts) :: so
  type(mvf_solve_info) :: info
  type(c_ptr) :: hnd
  integer(c_int32_t) :: nne
  integer(c_int64_t) :: nnze, ntriu, nnz
  integer :: i, j, k, id, nf, ifr, ierr, nnx, nny, nnz_
  real(c_double), allocatable :: freqs(:)
  real(c_double) :: ztop, f1, sg
  character(len=64) :: arg
  logical :: inblk

  nf = command_argument_count()
  if (nf == 0) then
    nf = 1
    allocate(freqs(1)); freqs(1) = 1.d0
  else
    allocate(freqs(nf))
    do i = 1, nf
      call get_command_argument(i, arg)
      read(arg, *) freqs(i)
    end do
  end if
  nnx = nx + 1; nny = ny + 1; nnz_ = nz + 1
  allocate(xp(nnx), yp(nny), zp(nnx*nny*nnz_), sigma(6, nnx*nny*nnz_), mu(6, nnx*nny*nnz_))
  do i = 1, nnx
    xp(i) = (i - 1)*h
  end do
  do j = 1, nny
    yp(j) = (j - 1)*h
  end do
  ztop = nze*h
  f1 = freqs(1)
  do i = 1, nnx
    do j = 1, nny
      do k = 1, nnz_
        id = (i - 1)*nny*nnz_ + (j - 1)*nnz_ + k
        zp(id) = (k - 1)*h
        sigma(:, id) = (0.d0, 0.d0)
        mu(:, id) = 0.d0
        mu(1, id) = 4.d0*pi*1.d-7; mu(4, id) = mu(1, id); mu(6, id) = mu(1, id)
        inblk = (i - 1 >= 8 .and. i - 1 <= 12 .and. j - 1 >= 8 .and. j - 1 <= 12 .and. k - 1 >= 13 .and. k - 1 <= 15)
        sg = 0.01d0
        if (inblk) sg = 1.d0
        if (zp(id) > ztop + 1.d-9) then
          sigma(1, id) = cmplx(0.d0, eps*2.d0*pi*f1)
        else
          sigma(1, id) = sg + cmplx(0.d0, eps*2.d0*pi*f1)
        end if
        sigma(4, id) = sigma(1, id); sigma(6, id) = sigma(1, id)
      end do
    end do
  end do
  grid = mvf_grid(nx, ny, nz, 8, c_loc(xp), c_loc(yp), c_loc(zp))
  model = mvf_model(c_loc(sigma), c_loc(mu))
  bc = mvf_bc(1, 1, 0, 2, 2, 0, 1.d0, 1.d0, 2.d0, ztop, 0.01d0, 1, 0, c_null_ptr, c_null_ptr)
  opts = mvf_options(MVF_EMU_ALL, 0, 0, 0)
  ierr = mvf_init(grid, model, bc, opts, hnd)
  if (ierr /= 0) then
    print *, 'mvf_init failed ', ierr, trim(mvf_error_string(hnd))
    stop 1
  end if
  ierr = mvf_sizes(hnd, nne, nnze, ntriu)
  allocate(ia(nnze), ja(nnze), a(nnze), rhs(2*nne))
  call mvf_default_solver_opts(so)
  do ifr = 1, nf
    ierr = mvf_set_frequency(hnd, freqs(ifr))
    if (mod(ifr, 2) == 1) then
      ! global_vfem, then find_zeros/rem_zeros on the host arrays
      if (ierr == 0) ierr = mvf_global_vfem(hnd, ia, ja, a, rhs, nnze, nne, 2)
      if (ierr == 0) ierr = mvf_remove_zeros(nnze, ia, ja, a, nnz)
    else
      ! the same in one call, the zero entries dropped on the device
      if (ierr == 0) ierr = mvf_global_vfem_nz(hnd, ia, ja, a, rhs, ntriu, nne, 2, nnz)
    end if
    if (ierr == 0) ierr = mvf_solve(hnd, rhs, so, info)
    if (ierr /= 0) then
      print *, 'frequency ', ifr, ' failed ', ierr, trim(mvf_error_string(hnd))
      stop 2
    end if
    write(*, '(a,i3,a,es12.4,a,i8,a,i10,a,i7,a,2es11.3)') 'freq ', ifr, ' f=', freqs(ifr), ' nne=', nne, &
         ' nnz=', nnz, ' it=', info%iterations, ' resid=', info%true_resid(1), info%true_resid(2)
  end do
  call mvf_free(hnd)
end program movfem_driver
