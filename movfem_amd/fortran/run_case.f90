! run_case.f90 -- the reference's main program (MoVFEM_3DMT.f90:44-143,
! 279-388) with every step on the library: the input files are read
! (input() and PARAM.INP), the lattice is built (grid_3d), and per frequency
! the system is assembled and solved on the GPU, node_solution / rx_data run
! on the device and PROBLEM.OUT and SOLUTION.OUT (inner region), RECEIVER.OUT
! and SURFACE.OUT are written as the reference writes them.
!
!   run_case [directory]     (default: the current directory)
!
! reads 3DEM_FEM.INP and PARAM.INP from the directory and writes the output
! files into the current directory.  The solve is FP64 GMRES preconditioned
! by the device factorization to a true relative residual of 1e-8 on the
! reference's complex64-rounded matrix (the reference: zmumps job 6), or on
! the FP64 operator where GMRES stalls on the rounded one.  Prints one line per frequency with the true
! residuals of both polarizations.
program run_case
  use iso_c_binding
  use movfem_gpu
  implicit none
  type(c_ptr) :: hin, hmesh, hgpu
  type(mvf_mesh_input) :: mi
  type(mvf_params) :: pr
  type(mvf_grid) :: grid
  type(mvf_model) :: model
  type(mvf_bc) :: bc
  type(mvf_options) :: opts
  type(mvf_solver_opts) :: so
  type(mvf_solve_info) :: info
  integer(c_int64_t) :: dims(12), nnze, ntriu, nnz, npt, nsel
  real(c_double) :: vals(6)
  real(c_double), pointer :: freqs(:), lsig(:), ldz(:)
  real(c_double), target, allocatable :: xp(:), yp(:), zp(:), mu(:), xsr(:), ysr(:), zsr(:), elev(:)
  real(c_double), target, allocatable :: rho(:), phi(:), nrho(:), nphi(:), xs(:), ys(:), zs(:)
  real(c_double), target :: lsig_t(64), ldz_t(64)
  complex(c_double_complex), target, allocatable :: sigma(:), a(:), rhs(:), esol(:), hsol(:), rx_e(:), rx_h(:)
  complex(c_double_complex), target, allocatable :: rx_z(:), node_z(:), pep(:), php(:), psrc(:), rhs0(:)
  integer(c_int32_t), target, allocatable :: nzl(:), ia(:), ja(:), sel(:)
  integer(c_int64_t), target, allocatable :: sids(:)
  integer(c_int32_t) :: nne, nx, ny, nz, mn, nextd, nsf, nsr, ord, ke_top, nxb, nyb, nf
  integer :: ierr, ii, i, j, k, i1, i2, j1, j2, k1, k2, ks, t
  integer(c_int64_t) :: nnx, nny, nnzl
  character(len=512) :: dir

  dir = '.'
  if (command_argument_count() > 0) call get_command_argument(1, dir)
  ierr = mvf_read_files(trim(dir)//'/3DEM_FEM.INP', trim(dir)//'/PARAM.INP', hin)
  if (ierr /= 0) call fail('reading the input files', c_null_ptr)
  ierr = mvf_inputs_view(hin, mi, pr)
  ierr = mvf_grid3d(mi, hmesh)
  if (ierr /= 0) call fail('grid_3d', c_null_ptr)
  ierr = mvf_mesh_sizes(hmesh, dims, vals)
  nx = int(dims(1)); ny = int(dims(2)); nz = int(dims(3)); mn = int(dims(4))
  nnx = dims(5); nny = dims(6); nnzl = dims(7); npt = dims(8)
  nextd = int(dims(9)); nsf = int(dims(10)); nsr = int(dims(11)); ord = int(dims(12)) - 1
  allocate(xp(nnx), yp(nny), zp(npt), sigma(6*npt), mu(6*npt), nzl(nsf), xsr(nsr), ysr(nsr), zsr(nsr))
  ierr = mvf_mesh_arrays(hmesh, c_loc(xp), c_loc(yp), c_loc(zp), c_loc(sigma), c_loc(mu), c_loc(nzl), &
       c_loc(xsr), c_loc(ysr), c_loc(zsr))
  call mvf_mesh_free(hmesh)
  write(*, '(a,3i6,a,i3,a,i3,a,i10)') 'elements ', nx, ny, nz, '  nodes/element ', mn, '  nextd ', nextd, &
       '  lattice nodes ', npt

  ! initialize_vfem's module state (MoVFEM_3DMT.f90:354-388)
  nf = pr%nf
  call c_f_pointer(pr%freqs, freqs, [nf])
  call c_f_pointer(pr%l_sigma, lsig, [pr%nl])
  lsig_t = 0.d0; ldz_t = 0.d0
  lsig_t(1:min(pr%nl, 64)) = lsig(1:min(pr%nl, 64))
  if (pr%nl > 1) then
    call c_f_pointer(pr%l_dz, ldz, [pr%nl - 1])
    ldz_t(1:min(pr%nl - 1, 64)) = ldz(1:min(pr%nl - 1, 64))
  end if
  grid = mvf_grid(nx, ny, nz, mn, c_loc(xp), c_loc(yp), c_loc(zp))
  model = mvf_model(c_loc(sigma), c_loc(mu))
  bc = mvf_bc(merge(1, 0, pr%boundary == 1), pr%inimod, pr%gpml_sch, nextd, nzl(nsf), 0, pr%a0, pr%b0, pr%nn, &
       vals(1), pr%h_sigma, pr%nl, 0, c_loc(lsig_t), c_loc(ldz_t))
  opts = mvf_options(MVF_EMU_ALL, 0, 0, 0)
  ierr = mvf_init(grid, model, bc, opts, hgpu)
  if (ierr /= 0) call fail('mvf_init', hgpu)
  ierr = mvf_sizes(hgpu, nne, nnze, ntriu)
  allocate(ia(nnze), ja(nnze), a(nnze), rhs(2*nne), rhs0(2*nne))
  allocate(esol(2*npt*3), hsol(2*npt*3), node_z(4*npt), nrho(4*npt), nphi(4*npt))
  allocate(rx_e(6*max(nsr, 1)), rx_h(6*max(nsr, 1)), rx_z(4*max(nsr, 1)), rho(4*max(nsr, 1)), phi(4*max(nsr, 1)))
  allocate(elev(max(nsr, 1)), pep(2*npt*3), php(2*npt*3), psrc(2*npt*3))
  call mvf_default_solver_opts(so)
  so%method = MVF_KRYLOV_GMRES
  so%precond = MVF_PC_DIRECT
  so%tol = 1.d-8

  ! inner region of SOLUTION.OUT (get_innergrid + inner_grid, geometry.f90:124-136,774-800; air = 0)
  i1 = nextd*ord + 1; i2 = int(nnx) - nextd*ord
  j1 = nextd*ord + 1; j2 = int(nny) - nextd*ord
  k1 = nextd*ord + 1; k2 = int(nnzl) - nzl(nsf)*ord - nzl(nsf - 1)*ord
  nsel = int(i2 - i1 + 1, c_int64_t)*(j2 - j1 + 1)*(k2 - k1 + 1)
  allocate(sel(nsel))
  t = 0
  do i = i1, i2
    do j = j1, j2
      do k = k1, k2
        t = t + 1
        sel(t) = int((i - 1)*nny*nnzl + (j - 1)*nnzl + k, c_int32_t)
      end do
    end do
  end do
  ! surface nodes of SURFACE.OUT (sur_data, receiver_data.f90:145-175): y outer, x inner
  nxb = int(nnx) - 2*nextd*ord; nyb = int(nny) - 2*nextd*ord
  ks = int(nnzl) - (nzl(nsf) + nzl(nsf - 1))*ord           ! 1-based lattice layer
  allocate(xs(nxb), ys(nyb), zs(nxb*nyb), sids(nxb*nyb))
  xs = xp(nextd*ord + 1:nextd*ord + nxb)
  ys = yp(nextd*ord + 1:nextd*ord + nyb)
  t = 0
  do j = 1, nyb
    do i = 1, nxb
      t = t + 1
      sids(t) = (int(nextd*ord + i - 1, c_int64_t)*nny + (nextd*ord + j - 1))*nnzl + ks - 1
      zs(t) = zp(sids(t) + 1)
    end do
  end do
  ke_top = nz - nzl(nsf) - nzl(nsf - 1)

  do ii = 1, nf
    ierr = mvf_set_frequency(hgpu, freqs(ii))
    ! write_problem_ugrid(ii, inner) (MoVFEM_3DMT.f90:70)
    if (ierr == 0) ierr = mvf_problem_fields(hgpu, c_loc(pep), c_loc(php), c_loc(psrc))
    if (ierr == 0) ierr = mvf_write_problem('PROBLEM.OUT'//c_null_char, ii, merge(0, 1, ii == 1), npt, c_loc(pep), &
         c_loc(php), c_loc(psrc), nsel, c_loc(sel))
    if (ierr == 0) ierr = mvf_global_vfem(hgpu, ia, ja, a, rhs, nnze, nne, 2)
    if (ierr == 0) ierr = mvf_remove_zeros(nnze, ia, ja, a, nnz)
    if (ierr == 0) then
      ! mvf_solve overwrites the RHS with the solution, as zmumps(job=6) does
      rhs0 = rhs
      so%matrix = MVF_MATRIX_C64REF
      so%maxit = 300
      ierr = mvf_solve(hgpu, rhs, so, info)
      if (ierr == MVF_E_NOCONV) then
        ! the complex64 rounding leaves the air's gradient null space
        ! near-singular instead of exactly singular; where GMRES stalls on it,
        ! solve the FP64 operator (same pattern and values before rounding)
        write(*, '(a,i4,a,2es11.3,a)') 'frequency ', ii, ': GMRES stalls at ', info%true_resid(1), &
             info%true_resid(2), ' on the complex64 matrix; solving the FP64 operator'
        rhs = rhs0
        so%matrix = MVF_MATRIX_FP64
        so%maxit = 100000
        ierr = mvf_solve(hgpu, rhs, so, info)
      end if
      if (ierr /= 0) write(*, '(a,i4,a,i7,a,2es11.3)') 'frequency ', ii, ': iterations ', info%iterations, &
           ' true residuals ', info%true_resid(1), info%true_resid(2)
    end if
    if (ierr /= 0) call fail('frequency loop', hgpu)
    ierr = mvf_readout(hgpu, c_loc(rhs), 0, ke_top, nsr, c_loc(xsr), c_loc(ysr), c_loc(esol), c_loc(hsol), &
         c_loc(rx_e), c_loc(rx_h), c_loc(rx_z), c_loc(rho), c_loc(phi), c_loc(elev), c_loc(node_z), &
         c_loc(nrho), c_loc(nphi))
    if (ierr /= 0) call fail('mvf_readout', hgpu)
    ierr = mvf_write_solution('SOLUTION.OUT'//c_null_char, ii, merge(0, 1, ii == 1), npt, c_loc(esol), &
         c_loc(hsol), c_loc(node_z), c_loc(nrho), c_loc(nphi), nsel, c_loc(sel))
    if (ierr == 0) ierr = mvf_write_receivers('RECEIVER.OUT'//c_null_char, ii, merge(0, 1, ii == 1), nf, nsr, &
         c_loc(xsr), c_loc(ysr), c_loc(elev), c_loc(rx_e), c_loc(rx_h), c_loc(rx_z), c_loc(rho), c_loc(phi))
    if (ierr == 0) ierr = mvf_write_surface('SURFACE.OUT'//c_null_char, ii, merge(0, 1, ii == 1), nf, nxb, nyb, &
         c_loc(xs), c_loc(ys), c_loc(zs), c_loc(sids), npt, c_loc(esol), c_loc(hsol), c_loc(node_z), &
         c_loc(nrho), c_loc(nphi))
    if (ierr /= 0) call fail('writing the output files', c_null_ptr)
    write(*, '(a,i3,a,es12.4,a,i9,a,i11,a,i4,a,2es11.3)') 'freq ', ii, ' f=', freqs(ii), ' nne=', nne, &
         ' nnz=', nnz, ' it=', info%iterations, ' resid=', info%true_resid(1), info%true_resid(2)
  end do
  call mvf_free(hgpu)
  call mvf_inputs_free(hin)

contains

  subroutine fail(what, h)
    character(len=*), intent(in) :: what
    type(c_ptr), intent(in) :: h
    print *, what, ' failed: ', trim(mvf_error_string(h))
    stop 1
  end subroutine fail
end program run_case
