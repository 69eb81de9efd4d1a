! ref_harness.f90 -- drives MoVFEM_3DMT's own modules (compiled from the
! reference sources by oracle/Makefile into oracle/_ref/) through the call
! sequence of the reference's global_vfem / local_vfem (MoVFEM_3DMT.f90:167-263).
! Those two routines are contained in the main program, which cannot be compiled
! here because it includes mpif.h and zmumps_struc.h; this file re-expresses
! only their call sequence.  Every numerical routine called below is the
! reference's own.
!
! TEST INFRASTRUCTURE ONLY: the checker for the oracle restatement.  One setup
! per process (the reference's init routines allocate module arrays without
! checking), so callers run each case in a fresh process.
module mvf_ref_harness
  use iso_c_binding
  use kind_param
  use geometry, only: g_nx, g_ny, g_nz, g_nnx, g_nny, g_nnz, g_nyz, g_npt, g_nordx, g_nordy, &
       g_nordz, g_xp, g_yp, g_zp, g_sigma, g_mu, g_nsf, g_nzl, g_nf, g_freq, g_ztop, nextd, &
       omega, pi, update_omega, update_sigma, asx, asy, asz
  use n_fem, only: init_n_fem, nf_get_r
  use v_fem, only: init_v_fem
  use problem, only: init_problem, p_elem_fields, pset_pmodel
  use boundary_conds, only: gpml_sch, a0, b0, nn, bd_setmodel, bd_updatemodel, get_pml, &
       dirichlet, f_boundary
  use integration, only: init_integration, int_elem_params, alocal, blocal
  use global_assembly, only: ga_init, ga_assemble_nze, ga_sort_sparse, assign_aij, assign_bi, &
       gne, nzindx, nne, nnze, sym
  use solution, only: node_solution, esol, hsol, write_solution
  use receiver_data, only: write_receiver_data, write_surface_data
  implicit none
  integer, save :: h_me = 0
contains

  ! h_sigma, nl, l_sigma(nl), l_dz(nl-1): the boundary 1-D model handed to
  ! bd_setmodel (MoVFEM_3DMT.f90:388)
  subroutine ref_setup(nx, ny, nz, mn, xp, yp, zp, sig, mu, dirich, inimod, sch, a0_, b0_, nn_, &
       nextd_, nzl_top, ztop, nf, freqs, h_sigma, nl, l_sigma, l_dz) bind(C, name='ref_setup')
    integer(c_int), value :: nx, ny, nz, mn, dirich, inimod, sch, nextd_, nzl_top, nf, nl
    real(c_double), value :: a0_, b0_, nn_, ztop, h_sigma
    real(c_double), intent(in) :: xp(*), yp(*), zp(*), mu(*), freqs(*), l_sigma(*), l_dz(*)
    complex(c_double_complex), intent(in) :: sig(*)
    integer :: o, i, k
    real(kind=double), allocatable :: lsig(:), ldz(:)

    g_nx = nx + 1; g_ny = ny + 1; g_nz = nz + 1
    if (mn == 8) then
      g_nordx = 2; g_nordy = 2; g_nordz = 2
    else
      g_nordx = 3; g_nordy = 3; g_nordz = 3
    end if
    o = g_nordx - 1
    g_nnx = nx*o + 1; g_nny = ny*o + 1; g_nnz = nz*o + 1
    g_nyz = g_nny*g_nnz; g_npt = g_nnx*g_nyz
    allocate(g_xp(g_nnx), g_yp(g_nny), g_zp(g_npt), g_sigma(6, g_npt), g_mu(6, g_npt))
    g_xp = xp(1:g_nnx); g_yp = yp(1:g_nny); g_zp = zp(1:g_npt)
    do i = 1, g_npt
      do k = 1, 6
        g_sigma(k, i) = sig(6*(i - 1) + k)
        g_mu(k, i) = mu(6*(i - 1) + k)
      end do
    end do
    nextd = nextd_
    g_nsf = 2
    allocate(g_nzl(g_nsf)); g_nzl = nzl_top
    g_nf = nf
    allocate(g_freq(nf)); g_freq = freqs(1:nf)
    g_ztop = ztop
    ! reference coordinates of the element's lattice nodes (what the
    ! reference's private gqg_nodes sets, geometry.f90:715-740)
    allocate(asx(g_nordx), asy(g_nordy), asz(g_nordz))
    do i = 1, g_nordx
      asx(i) = -1.d0 + dble(i - 1)*2.d0/dble(g_nordx - 1)
    end do
    asy = asx; asz = asx
    omega = 2.d0*pi*g_freq(1)
    gpml_sch = sch; a0 = a0_; b0 = b0_; nn = nn_
    ! initialize_vfem order (MoVFEM_3DMT.f90:354-388)
    call init_n_fem(mn)
    select case (mn)
    case (8); call init_v_fem(12); h_me = 12
    case (20); call init_v_fem(36); h_me = 36
    case (27); call init_v_fem(54); h_me = 54
    end select
    call init_problem(1, 2, 1)
    call init_integration()
    call ga_init(.true., dirich /= 0, inimod)
    allocate(lsig(nl), ldz(max(nl - 1, 1)))
    lsig = l_sigma(1:nl); ldz = 0.d0
    if (nl > 1) ldz(1:nl - 1) = l_dz(1:nl - 1)
    call bd_setmodel(h_sigma, nl, lsig, ldz)
  end subroutine ref_setup

  ! frequency-loop state update (MoVFEM_3DMT.f90:66-69); write_problem_ugrid is
  ! output only and skipped
  subroutine ref_set_frequency(ifreq) bind(C, name='ref_set_frequency')
    integer(c_int), value :: ifreq
    call update_omega(ifreq)
    call update_sigma()
    call pset_pmodel()
    call bd_updatemodel()
  end subroutine ref_set_frequency

  subroutine ref_sizes(o_nne, o_nnze) bind(C, name='ref_sizes')
    integer(c_int), intent(out) :: o_nne
    integer(c_int64_t), intent(out) :: o_nnze
    o_nne = nne; o_nnze = nnze
  end subroutine ref_sizes

  subroutine ref_gne(out) bind(C, name='ref_gne')
    integer(c_int), intent(out) :: out(h_me, *)
    integer :: e, im
    do e = 1, size(gne, 1)
      do im = 1, h_me
        out(im, e) = gne(e, im)
      end do
    end do
  end subroutine ref_gne

  subroutine ref_nzindx(out) bind(C, name='ref_nzindx')
    integer(c_int), intent(out) :: out(*)
    out(1:size(nzindx)) = nzindx
  end subroutine ref_nzindx

  subroutine ref_get_sigma(out) bind(C, name='ref_get_sigma')
    complex(c_double_complex), intent(out) :: out(6, *)
    out(:, 1:g_npt) = g_sigma
  end subroutine ref_get_sigma

  ! one element in loop position (ie,je,ke): the per-element calls of
  ! global_vfem, then every alocal(im,jm) and blocal(im)
  subroutine ref_element(ie, je, ke, aloc, bloc) bind(C, name='ref_element')
    integer(c_int), value :: ie, je, ke
    complex(c_double_complex), intent(out) :: aloc(h_me, h_me), bloc(2, h_me)
    integer :: eno, im, jm
    complex(kind=double) :: b2(2)
    eno = (ie - 1)*g_nyz*(g_nordx - 1) + (je - 1)*g_nnz*(g_nordy - 1) + (ke - 1)*(g_nordz - 1) + 1
    call nf_get_r(ie, je, ke, eno)
    call p_elem_fields()
    call int_elem_params()
    if (.not. dirichlet) call get_pml(ie, je, ke)
    do im = 1, h_me
      do jm = 1, h_me
        aloc(jm, im) = alocal(im, jm)
      end do
      b2 = blocal(im)
      bloc(:, im) = b2
    end do
  end subroutine ref_element

  ! global_vfem (MoVFEM_3DMT.f90:167-216)
  subroutine ref_global_vfem(ia, ja, a, b) bind(C, name='ref_global_vfem')
    integer(c_int), intent(inout) :: ia(*), ja(*)
    complex(c_double_complex), intent(inout) :: a(*), b(*)
    integer :: ie, je, ke, ide, eno
    call ga_assemble_nze(ia, ja)
    a(1:nnze) = cmplx(0.d0, 0.d0, kind=double)
    b(1:2*nne) = cmplx(0.d0, 0.d0, kind=double)
    do ie = 1, g_nx - 1
      do je = 1, g_ny - 1
        do ke = 1, g_nz - 1
          eno = (ie - 1)*g_nyz*(g_nordx - 1) + (je - 1)*g_nnz*(g_nordy - 1) + (ke - 1)*(g_nordz - 1) + 1
          ide = (ie - 1)*(g_ny - 1)*(g_nz - 1) + (je - 1)*(g_nz - 1) + ke
          call nf_get_r(ie, je, ke, eno)
          call p_elem_fields()
          call int_elem_params()
          if (.not. dirichlet) call get_pml(ie, je, ke)
          call local_vfem_h(ide, ia, ja, a, b)
        end do
      end do
    end do
    deallocate(nzindx)
    call ga_sort_sparse(ia, ja, a)
  end subroutine ref_global_vfem

  ! local_vfem (MoVFEM_3DMT.f90:221-263)
  subroutine local_vfem_h(ide, ia, ja, a, b)
    integer, intent(in) :: ide
    integer, dimension(*), intent(inout) :: ia, ja
    complex(kind=double), dimension(*), intent(inout) :: a, b
    integer :: im, jm
    complex(kind=double), dimension(2) :: bda
    do im = 1, h_me
      if (gne(ide, im) < 0) cycle
      do jm = 1, h_me
        if (gne(ide, jm) < 0) cycle
        if (sym .and. gne(ide, im) < gne(ide, jm)) cycle
        call assign_aij(ide, im, jm, alocal(im, jm), a, ia, ja)
      end do
    end do
    do im = 1, h_me
      if (gne(ide, im) < 0) cycle
      bda = cmplx(0.d0, 0.d0)
      do jm = 1, h_me
        if (gne(ide, jm) >= 0) cycle
        bda = bda + f_boundary(-gne(ide, jm), jm)*alocal(im, jm)
      end do
      call assign_bi(ide, im, (blocal(im) - bda), b)
    end do
  end subroutine local_vfem_h

  ! node_solution (solution.f90:18-49) on the solution x (2*nne, polarization
  ! major): total E and H at every lattice node, (2*npt, 3) each
  subroutine ref_node_solution(x, eo, ho) bind(C, name='ref_node_solution')
    complex(c_double_complex), intent(in) :: x(*)
    complex(c_double_complex), intent(out) :: eo(2*g_npt, 3), ho(2*g_npt, 3)
    call node_solution(x(1:2*nne))
    eo = esol
    ho = hsol
  end subroutine ref_node_solution

  ! write_receiver_data(1) (receiver_data.f90:25-67) for receivers (xs, ys);
  ! it writes RECEIVER.OUT in the working directory.  nzl_a/nzl_b are the
  ! layer counts of the two top zones that place the receiver element
  ! (get_rx_element, :187-209).  Must follow ref_node_solution.
  subroutine ref_receivers(nsr, xs, ys, nzl_a, nzl_b) bind(C, name='ref_receivers')
    use geometry, only: g_nsr, g_xsr, g_ysr, g_zsr, x0, y0, z0
    integer(c_int), value :: nsr, nzl_a, nzl_b
    real(c_double), intent(in) :: xs(*), ys(*)
    g_nsr = nsr
    if (allocated(g_xsr)) deallocate(g_xsr, g_ysr, g_zsr)
    allocate(g_xsr(nsr), g_ysr(nsr), g_zsr(nsr))
    g_xsr = xs(1:nsr); g_ysr = ys(1:nsr); g_zsr = 0.d0
    x0 = 0.d0; y0 = 0.d0; z0 = 0.d0
    g_nzl(g_nsf) = nzl_a
    g_nzl(g_nsf - 1) = nzl_b
    g_nf = 1
    call write_receiver_data(1)
  end subroutine ref_receivers

  ! write_solution(1, inner) (solution.f90:73-205): SOLUTION.OUT in the
  ! working directory.  Must follow ref_node_solution; frees z, rho, phi.
  subroutine ref_write_solution(inner) bind(C, name='ref_write_solution')
    integer(c_int), value :: inner
    g_nf = 1
    call write_solution(1, inner)
  end subroutine ref_write_solution

  ! write_surface_data(1) (receiver_data.f90:69-109): SURFACE.OUT for the
  ! earth-surface lattice nodes inside the extension zones; nzl_a/nzl_b as
  ! for ref_receivers.  Must come last: it frees esol and hsol.
  subroutine ref_write_surface(nzl_a, nzl_b) bind(C, name='ref_write_surface')
    use geometry, only: x0, y0, z0
    integer(c_int), value :: nzl_a, nzl_b
    x0 = 0.d0; y0 = 0.d0; z0 = 0.d0
    g_nzl(g_nsf) = nzl_a
    g_nzl(g_nsf - 1) = nzl_b
    g_nf = 1
    call write_surface_data(1)
  end subroutine ref_write_surface
  ! write_problem_ugrid(1, 0) (problem.f90:188-240): PROBLEM.OUT in the
  ! working directory, every lattice node, both polarizations, at the current
  ! frequency.  g_nf = 1 so that the reference closes the file.
  subroutine ref_write_problem() bind(C, name='ref_write_problem')
    use problem, only: write_problem_ugrid
    g_nf = 1
    call write_problem_ugrid(1, 0)
  end subroutine ref_write_problem

  ! ---- input() and grid_3d (read_input.f90:37-43, geometry.f90:48-122) ----
  ! ref_read_input reads 3DEM_FEM.INP from the working directory with the
  ! reference's reader; ref_input_sizes / ref_input_arrays hand its arrays
  ! over before ref_grid3d (whose grid_3d modifies them) runs on them.
  subroutine ref_read_input() bind(C, name='ref_read_input')
    use read_input, only: input
    call input()
  end subroutine ref_read_input

  ! isz = nsr, nsf, size(in_xto,2), imu, isigma, mx, my, mz, mn;
  ! rsz = dx, dy, dz, high, depth, bcksigma
  subroutine ref_input_sizes(isz, rsz) bind(C, name='ref_input_sizes')
    use read_input
    integer(c_int), intent(out) :: isz(9)
    real(c_double), intent(out) :: rsz(6)
    isz = (/ in_nsr, in_nsf, size(in_xto, 2), in_imu, in_isigma, in_mx, in_my, in_mz, in_mn /)
    rsz = (/ in_dx, in_dy, in_dz, in_high, in_depth, in_bcksigma /)
  end subroutine ref_input_sizes

  subroutine ref_input_arrays(xsr, ysr, zsr, nsp, xto, yto, zto, xm, ym, zm, sig, mu, ijs, ijm) &
       bind(C, name='ref_input_arrays')
    use read_input
    real(c_double), intent(out) :: xsr(*), ysr(*), zsr(*), xto(*), yto(*), zto(*), xm(*), ym(*), zm(*)
    real(c_double), intent(out) :: sig(*), mu(*)
    integer(c_int), intent(out) :: nsp(*), ijs(*), ijm(*)
    integer :: n
    xsr(1:in_nsr) = in_xsr; ysr(1:in_nsr) = in_ysr; zsr(1:in_nsr) = in_zsr
    nsp(1:in_nsf) = in_nsp
    n = size(in_xto)
    xto(1:n) = reshape(in_xto, (/ n /)); yto(1:n) = reshape(in_yto, (/ n /))
    zto(1:n) = reshape(in_zto, (/ n /))
    xm(1:in_mx) = in_xm; ym(1:in_my) = in_ym
    n = in_mx * in_my * in_mz
    zm(1:n) = in_zm
    sig(1:in_isigma * n) = reshape(in_sigma, (/ in_isigma * n /))
    mu(1:size(in_mu)) = reshape(in_mu, (/ size(in_mu) /))
    ijs(1:2 * in_isigma) = reshape(in_ijsigma, (/ 2 * in_isigma /))
    if (in_imu > 0) ijm(1:2 * in_imu) = reshape(in_ijmu, (/ 2 * in_imu /))
  end subroutine ref_input_arrays

  ! grid_3d on the arrays input() read, as initialize_vfem calls it
  ! (MoVFEM_3DMT.f90:337-341) but with iout = 0 (no grid files)
  subroutine ref_grid3d(nf, freqs) bind(C, name='ref_grid3d')
    use read_input
    use geometry, only: grid_3d
    integer(c_int), value :: nf
    real(c_double), intent(in) :: freqs(nf)
    real(kind=double), allocatable :: f(:)
    allocate(f(nf))
    f = freqs(1:nf)
    call grid_3d(in_dx, in_dy, in_dz, in_high, in_depth, in_mn, in_bcksigma, in_nsf, &
         in_nsp, in_xto, in_yto, in_zto, in_nsr, in_xsr, in_ysr, in_zsr, in_mx, in_my, &
         in_mz, in_xm, in_ym, in_zm, in_isigma, in_ijsigma, in_sigma, in_imu, in_ijmu, in_mu, &
         nf, f, 0, 0)
  end subroutine ref_grid3d

  ! isz = g_nx, g_ny, g_nz, g_nnx, g_nny, g_nnz, g_npt, nextd, g_nsf, g_nordx, g_nsr;
  ! rsz = g_ztop, x0, y0, z0, omega
  subroutine ref_grid_sizes(isz, rsz) bind(C, name='ref_grid_sizes')
    use geometry, only: x0, y0, z0, g_nsr
    integer(c_int), intent(out) :: isz(11)
    real(c_double), intent(out) :: rsz(5)
    isz = (/ g_nx, g_ny, g_nz, g_nnx, g_nny, g_nnz, g_npt, nextd, g_nsf, g_nordx, g_nsr /)
    rsz = (/ g_ztop, x0, y0, z0, omega /)
  end subroutine ref_grid_sizes

  subroutine ref_grid_arrays(xp, yp, zp, sig, mu, nzl, xsr, ysr, zsr) bind(C, name='ref_grid_arrays')
    use geometry, only: g_xsr, g_ysr, g_zsr, g_nsr
    real(c_double), intent(out) :: xp(*), yp(*), zp(*), mu(*), xsr(*), ysr(*), zsr(*)
    complex(c_double_complex), intent(out) :: sig(*)
    integer(c_int), intent(out) :: nzl(*)
    xp(1:g_nnx) = g_xp; yp(1:g_nny) = g_yp; zp(1:g_npt) = g_zp
    sig(1:6 * g_npt) = reshape(g_sigma, (/ 6 * g_npt /))
    mu(1:6 * g_npt) = reshape(g_mu, (/ 6 * g_npt /))
    nzl(1:g_nsf) = g_nzl
    xsr(1:g_nsr) = g_xsr; ysr(1:g_nsr) = g_ysr; zsr(1:g_nsr) = g_zsr
  end subroutine ref_grid_arrays
end module mvf_ref_harness
