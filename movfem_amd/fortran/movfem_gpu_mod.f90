! movfem_gpu_mod.f90 -- ISO_C_BINDING interface to libmovfem_gpu.so
! (include/movfem_gpu.h).  A Fortran caller keeps the reference's call shape:
!   call global_vfem(irn, jcn, a, rhs)   (MoVFEM_3DMT.f90:82)
! becomes
!   ierr = mvf_global_vfem(h, irn, jcn, a, rhs, nnze, nne, 2)
! and zmumps(job=6) (MoVFEM_3DMT.f90:113-115) becomes mvf_solve(h, rhs, opts, info).
module movfem_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: MVF_EMU_ALL = 3
  integer(c_int), parameter :: MVF_KRYLOV_COCG = 0, MVF_KRYLOV_BICGSTAB = 1, MVF_KRYLOV_GMRES = 2
  integer(c_int), parameter :: MVF_PREC_FP64 = 0, MVF_PREC_MIXED = 1
  integer(c_int), parameter :: MVF_MATRIX_C64REF = 0, MVF_MATRIX_FP64 = 1
  integer(c_int), parameter :: MVF_PC_NONE = 0, MVF_PC_JACOBI = 1, MVF_PC_DIRECT = 2
  integer(c_int), parameter :: MVF_E_ARG = -1, MVF_E_DEVICE = -2, MVF_E_NOCONV = -6

  type, bind(C) :: mvf_grid
    integer(c_int32_t) :: nx, ny, nz, nodes_per_elem
    type(c_ptr) :: xp, yp, zp
  end type mvf_grid

  type, bind(C) :: mvf_model
    type(c_ptr) :: sigma, mu
  end type mvf_model

  type, bind(C) :: mvf_bc
    integer(c_int32_t) :: dirichlet, inimod, gpml_sch, nextd, nzl_top, pad_
    real(c_double) :: a0, b0, nn
    real(c_double) :: ztop, h_sigma
    integer(c_int32_t) :: nl, pad2_
    type(c_ptr) :: l_sigma, l_dz
  end type mvf_bc

  type, bind(C) :: mvf_options
    integer(c_int32_t) :: emulate, device, verbose, pad_
  end type mvf_options

  type, bind(C) :: mvf_solver_opts
    integer(c_int32_t) :: method, precision, matrix, precond
    real(c_double) :: tol
    integer(c_int32_t) :: maxit, check_every
    real(c_double) :: inner_tol
    integer(c_int32_t) :: use_graph, gmres_m, freqs_per_factor, direct_tile
  end type mvf_solver_opts

  type, bind(C) :: mvf_solve_info
    integer(c_int32_t) :: iterations, refinements, converged, fallbacks
    real(c_double) :: true_resid(64)
    real(c_double) :: seconds
    integer(c_int64_t) :: spmv_launches(2)
    real(c_double) :: spmv_ms(2), spmv_bytes(2)
    integer(c_int32_t) :: spmv_ncol, spmv_lpr
    real(c_double) :: factor_seconds, factor_flops, factor_bytes
    integer(c_int32_t) :: factorizations, static_pivots, precond_applies, operator_fp32
    real(c_double) :: gemm_ms, gemm_flops
    integer(c_int64_t) :: gemm_launches, graph_launches
  end type mvf_solve_info

  ! grid_3d's arguments in the reference's layouts (geometry.f90:48-62)
  type, bind(C) :: mvf_mesh_input
    real(c_double) :: dx, dy, dz, high, depth, sigma_bg
    integer(c_int32_t) :: mn, nsr, nsf, mx, my, mz, isigma, imu, nf, pad_
    type(c_ptr) :: xsr, ysr, zsr, nsp, xto, yto, zto, xm, ym, zm, ijsigma, sigma, ijmu, mu, freqs
  end type mvf_mesh_input

  ! PARAM.INP (MoVFEM_3DMT.f90:302-334)
  type, bind(C) :: mvf_params
    integer(c_int32_t) :: nf, boundary, inimod, nl, gpml_sch, pad_
    real(c_double) :: h_sigma, a0, b0, nn
    type(c_ptr) :: freqs, l_sigma, l_dz
  end type mvf_params

  interface
    ! input() + the PARAM.INP reads; paths are null-terminated (c_null_ptr: skip)
    integer(c_int) function mvf_read_inputs(inp_path, param_path, h) bind(C, name='mvf_read_inputs')
      import :: c_int, c_ptr
      type(c_ptr), value :: inp_path, param_path
      type(c_ptr), intent(out) :: h
    end function mvf_read_inputs

    integer(c_int) function mvf_inputs_view(h, mesh, params) bind(C, name='mvf_inputs_view')
      import :: c_int, c_ptr, mvf_mesh_input, mvf_params
      type(c_ptr), value :: h
      type(mvf_mesh_input), intent(out) :: mesh
      type(mvf_params), intent(out) :: params
    end function mvf_inputs_view

    subroutine mvf_inputs_free(h) bind(C, name='mvf_inputs_free')
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine mvf_inputs_free

    ! grid_3d
    integer(c_int) function mvf_grid3d(mesh_in, m) bind(C, name='mvf_grid3d')
      import :: c_int, c_ptr, mvf_mesh_input
      type(mvf_mesh_input), intent(in) :: mesh_in
      type(c_ptr), intent(out) :: m
    end function mvf_grid3d

    integer(c_int) function mvf_mesh_sizes(m, dims, vals) bind(C, name='mvf_mesh_sizes')
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: m
      integer(c_int64_t), intent(out) :: dims(12)
      real(c_double), intent(out) :: vals(6)
    end function mvf_mesh_sizes

    integer(c_int) function mvf_mesh_arrays(m, xp, yp, zp, sigma, mu, nzl, xsr, ysr, zsr) &
         bind(C, name='mvf_mesh_arrays')
      import :: c_int, c_ptr
      type(c_ptr), value :: m, xp, yp, zp, sigma, mu, nzl, xsr, ysr, zsr
    end function mvf_mesh_arrays

    subroutine mvf_mesh_free(m) bind(C, name='mvf_mesh_free')
      import :: c_ptr
      type(c_ptr), value :: m
    end subroutine mvf_mesh_free

    integer(c_int) function mvf_init(grid, model, bc, opts, h) bind(C, name='mvf_init')
      import :: c_int, c_ptr, mvf_grid, mvf_model, mvf_bc, mvf_options
      type(mvf_grid), intent(in) :: grid
      type(mvf_model), intent(in) :: model
      type(mvf_bc), intent(in) :: bc
      type(mvf_options), intent(in) :: opts
      type(c_ptr), intent(out) :: h
    end function mvf_init

    subroutine mvf_free(h) bind(C, name='mvf_free')
      import :: c_ptr
      type(c_ptr), value :: h
    end subroutine mvf_free

    type(c_ptr) function mvf_last_error(h) bind(C, name='mvf_last_error')
      import :: c_ptr
      type(c_ptr), value :: h
    end function mvf_last_error

    integer(c_int) function mvf_sizes(h, nne, nnze, nnz_triu) bind(C, name='mvf_sizes')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: nne
      integer(c_int64_t), intent(out) :: nnze, nnz_triu
    end function mvf_sizes

    integer(c_int) function mvf_set_frequency(h, freq) bind(C, name='mvf_set_frequency')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), value :: freq
    end function mvf_set_frequency

    integer(c_int) function mvf_global_vfem(h, ia, ja, a, b, nnze, nne, ndir) bind(C, name='mvf_global_vfem')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double_complex
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: ia(*), ja(*)
      complex(c_double_complex), intent(out) :: a(*), b(*)
      integer(c_int64_t), value :: nnze
      integer(c_int32_t), value :: nne, ndir
    end function mvf_global_vfem

    integer(c_int) function mvf_global_vfem_nz(h, ia, ja, a, b, cap, nne, ndir, nnz_out) &
        bind(C, name='mvf_global_vfem_nz')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double_complex
      type(c_ptr), value :: h
      integer(c_int32_t), intent(out) :: ia(*), ja(*)
      complex(c_double_complex), intent(out) :: a(*), b(*)
      integer(c_int64_t), value :: cap
      integer(c_int32_t), value :: nne, ndir
      integer(c_int64_t), intent(out) :: nnz_out
    end function mvf_global_vfem_nz

    integer(c_int) function mvf_remove_zeros(n, ia, ja, a, n_out) bind(C, name='mvf_remove_zeros')
      import :: c_int, c_int32_t, c_int64_t, c_double_complex
      integer(c_int64_t), value :: n
      integer(c_int32_t), intent(inout) :: ia(*), ja(*)
      complex(c_double_complex), intent(inout) :: a(*)
      integer(c_int64_t), intent(out) :: n_out
    end function mvf_remove_zeros

    subroutine mvf_default_solver_opts(o) bind(C, name='mvf_default_solver_opts')
      import :: mvf_solver_opts
      type(mvf_solver_opts), intent(out) :: o
    end subroutine mvf_default_solver_opts

    integer(c_int) function mvf_solve(h, rhs, o, info) bind(C, name='mvf_solve')
      import :: c_int, c_ptr, c_double_complex, mvf_solver_opts, mvf_solve_info
      type(c_ptr), value :: h
      complex(c_double_complex), intent(inout) :: rhs(*)
      type(mvf_solver_opts), intent(in) :: o
      type(mvf_solve_info), intent(out) :: info
    end function mvf_solve

    ! mvf_solve without device allocation (COCG/BiCGStab, FP64): for callers that
    ! cannot allocate; buffers in a 64 MiB static device array per GPU
    integer(c_int) function mvf_solve_noalloc(h, rhs, o, info) bind(C, name='mvf_solve_noalloc')
      import :: c_int, c_ptr, c_double_complex, mvf_solver_opts, mvf_solve_info
      type(c_ptr), value :: h
      complex(c_double_complex), intent(inout) :: rhs(*)
      type(mvf_solver_opts), intent(in) :: o
      type(mvf_solve_info), intent(out) :: info
    end function mvf_solve_noalloc

    ! mvf_solve started from x0 (the least-squares-scaled initial guess);
    ! alpha(4): the scaling per polarization (re, im)
    integer(c_int) function mvf_solve_guess(h, x0, rhs, o, info, alpha) bind(C, name='mvf_solve_guess')
      import :: c_int, c_ptr, c_double, c_double_complex, mvf_solver_opts, mvf_solve_info
      type(c_ptr), value :: h
      complex(c_double_complex), intent(in) :: x0(*)
      complex(c_double_complex), intent(inout) :: rhs(*)
      type(mvf_solver_opts), intent(in) :: o
      type(mvf_solve_info), intent(out) :: info
      real(c_double), intent(out) :: alpha(4)
    end function mvf_solve_guess

    ! several models x frequencies: hs(nh) handles, freqs(nfreq, nh) in Hz (one
    ! column per model), infos(nh), status(nh); nthreads host threads
    integer(c_int) function mvf_solve_batch(hs, nh, freqs, nfreq, o, infos, status, nthreads) &
         bind(C, name='mvf_solve_batch')
      import :: c_int, c_ptr, c_int32_t, c_double, mvf_solver_opts, mvf_solve_info
      type(c_ptr), intent(in) :: hs(*)
      integer(c_int32_t), value :: nh, nfreq, nthreads
      real(c_double), intent(in) :: freqs(*)
      type(mvf_solver_opts), intent(in) :: o
      type(mvf_solve_info), intent(out) :: infos(*)
      integer(c_int32_t), intent(out) :: status(*)
    end function mvf_solve_batch

    ! node_solution + rx_data + z_rho_phi on the device; pass c_null_ptr for
    ! outputs not wanted and for x to read frequency f of the last batch
    integer(c_int) function mvf_readout(h, x, f, ke_top, nrx, rx, ry, esol, hsol, rx_e, rx_h, rx_z, rho, phi, &
         rx_elev, node_z, node_rho, node_phi) bind(C, name='mvf_readout')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: h, x, rx, ry, esol, hsol, rx_e, rx_h, rx_z, rho, phi, rx_elev, node_z, node_rho, node_phi
      integer(c_int32_t), value :: f, ke_top, nrx
    end function mvf_readout

    ! the reference's output files (write_solution, write_receiver_data,
    ! write_surface_data); path is a null-terminated string
    integer(c_int) function mvf_write_solution(path, nf, append, npt, esol, hsol, node_z, node_rho, node_phi, &
         nsel, sel) bind(C, name='mvf_write_solution')
      import :: c_int, c_ptr, c_char, c_int32_t, c_int64_t
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: nf, append
      integer(c_int64_t), value :: npt, nsel
      type(c_ptr), value :: esol, hsol, node_z, node_rho, node_phi, sel
    end function mvf_write_solution

    integer(c_int) function mvf_write_receivers(path, nf, append, nfreq, nrx, rx, ry, elev, rx_e, rx_h, rx_z, &
         rho, phi) bind(C, name='mvf_write_receivers')
      import :: c_int, c_ptr, c_char, c_int32_t
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: nf, append, nfreq, nrx
      type(c_ptr), value :: rx, ry, elev, rx_e, rx_h, rx_z, rho, phi
    end function mvf_write_receivers

    ! PROBLEM.OUT's fields at the current frequency (complex (2, npt, 3) each)
    integer(c_int) function mvf_problem_fields(h, ep, hp, src) bind(C, name='mvf_problem_fields')
      import :: c_int, c_ptr
      type(c_ptr), value :: h, ep, hp, src
    end function mvf_problem_fields

    ! write_problem_ugrid; sel: nsel 1-based node ids (g_nb) or c_null_ptr
    integer(c_int) function mvf_write_problem(path, nf, append, npt, ep, hp, src, nsel, sel) &
         bind(C, name='mvf_write_problem')
      import :: c_int, c_ptr, c_char, c_int32_t, c_int64_t
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: nf, append
      integer(c_int64_t), value :: npt, nsel
      type(c_ptr), value :: ep, hp, src, sel
    end function mvf_write_problem

    integer(c_int) function mvf_write_surface(path, nf, append, nfreq, nxb, nyb, xs, ys, zs, ids, npt, esol, &
         hsol, node_z, node_rho, node_phi) bind(C, name='mvf_write_surface')
      import :: c_int, c_ptr, c_char, c_int32_t, c_int64_t
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int32_t), value :: nf, append, nfreq, nxb, nyb
      integer(c_int64_t), value :: npt
      type(c_ptr), value :: xs, ys, zs, ids, esol, hsol, node_z, node_rho, node_phi
    end function mvf_write_surface
  end interface

contains

  ! mvf_read_inputs with Fortran file names; an empty name skips that file
  integer function mvf_read_files(inp, param, h) result(ierr)
    character(len=*), intent(in) :: inp, param
    type(c_ptr), intent(out) :: h
    character(kind=c_char), target :: ci(len_trim(inp) + 1), cp(len_trim(param) + 1)
    type(c_ptr) :: pi, pp
    integer :: i
    do i = 1, len_trim(inp)
      ci(i) = inp(i:i)
    end do
    ci(len_trim(inp) + 1) = c_null_char
    do i = 1, len_trim(param)
      cp(i) = param(i:i)
    end do
    cp(len_trim(param) + 1) = c_null_char
    pi = c_null_ptr
    pp = c_null_ptr
    if (len_trim(inp) > 0) pi = c_loc(ci)
    if (len_trim(param) > 0) pp = c_loc(cp)
    ierr = mvf_read_inputs(pi, pp, h)
  end function mvf_read_files

  ! the message of the last error on a handle, as a Fortran string
  function mvf_error_string(h) result(s)
    type(c_ptr), intent(in) :: h
    character(len=512) :: s
    character(kind=c_char), pointer :: p(:)
    integer :: i
    s = ' '
    call c_f_pointer(mvf_last_error(h), p, [512])
    do i = 1, 512
      if (p(i) == c_null_char) exit
      s(i:i) = p(i)
    end do
  end function mvf_error_string

end module movfem_gpu
