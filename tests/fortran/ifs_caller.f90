! A Fortran caller of the C ABI (include/ifs.h) through iso_c_binding: the bind(C)
! types and interfaces a maintainer adds to ItoFoggySim, and a wrapper with the
! interface of the reference's GMRES callback calc_vector_b(vector_x, vector_b)
! (lib_solver_GMRES.f90:52-57,72-77).  The reference's own modules are left out
! (lib_mie_type and its dependencies do not build here); the cluster is set up
! directly.
!
!   ifs_caller        host-only calls: struct sizes, argument errors, octree functions
!   ifs_caller gpu    one matvec of a 3-sphere cluster through the calc_vector_b wrapper
module ifs_bind
    use iso_c_binding
    implicit none

    type, bind(C) :: ifs_cluster
        integer(c_int32_t) :: n_spheres
        type(c_ptr) :: centres, radii, m_rel, lmax
    end type

    type, bind(C) :: ifs_params
        real(c_double)     :: k
        integer(c_int32_t) :: n_global, z_selector, ref_row_zeroing, use_fmm, leaf_size
        real(c_double)     :: fmm_d0
        integer(c_int32_t) :: fmm_lmax_cap, workspace_mb
        integer(c_int32_t) :: d_convention, cache_mb
        integer(c_int32_t) :: precision
    end type

    type, bind(C) :: ifs_solve_opts
        integer(c_int32_t) :: restart, max_iter
        real(c_double)     :: tol
        integer(c_int32_t) :: use_x0, verbose, precision, flags
    end type

    type, bind(C) :: ifs_solve_info
        integer(c_int32_t) :: iterations, converged
        real(c_double)     :: rel_residual, seconds, rel_residual_flat
    end type

    interface
        integer(c_int32_t) function ifs_create(ctx, cluster, params) bind(C, name="ifs_create")
            import :: c_ptr, c_int32_t, ifs_cluster, ifs_params
            type(c_ptr), intent(out) :: ctx
            type(ifs_cluster), intent(in) :: cluster
            type(ifs_params), intent(in) :: params
        end function
        integer(c_int32_t) function ifs_destroy(ctx) bind(C, name="ifs_destroy")
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: ctx
        end function
        integer(c_int64_t) function ifs_vector_length(ctx) bind(C, name="ifs_vector_length")
            import :: c_ptr, c_int64_t
            type(c_ptr), value :: ctx
        end function
        integer(c_int32_t) function ifs_matvec(ctx, x, b, n) bind(C, name="ifs_matvec")
            import :: c_ptr, c_int32_t, c_int64_t, c_double_complex
            type(c_ptr), value :: ctx
            complex(c_double_complex), intent(in)  :: x(*)
            complex(c_double_complex), intent(out) :: b(*)
            integer(c_int64_t), value :: n
        end function
        integer(c_int32_t) function ifs_tree_scale(pts, n, scaled, x_min_d) bind(C, name="ifs_tree_scale")
            import :: c_int32_t, c_double
            real(c_double), intent(in) :: pts(*)
            integer(c_int32_t), value :: n
            real(c_double), intent(out) :: scaled(*), x_min_d(*)
        end function
        integer(c_int32_t) function ifs_tree_universal_index(scaled, n, level, out) &
                bind(C, name="ifs_tree_universal_index")
            import :: c_int32_t, c_int64_t, c_double
            real(c_double), intent(in) :: scaled(*)
            integer(c_int32_t), value :: n, level
            integer(c_int64_t), intent(out) :: out(*)
        end function
        integer(c_int32_t) function ifs_tree_domain_i4(box, level, out, cap) bind(C, name="ifs_tree_domain_i4")
            import :: c_int32_t, c_int64_t
            integer(c_int64_t), value :: box
            integer(c_int32_t), value :: level, cap
            integer(c_int64_t), intent(out) :: out(*)
        end function
    end interface

    type(c_ptr), save :: m_ctx = c_null_ptr

contains

    ! The interface of the reference's calc_vector_b callback (lib_solver_GMRES.f90:72-77).
    subroutine lib_ifs_calculate_vector_b(vector_x, vector_b)
        double complex, dimension(:), intent(in) :: vector_x
        double complex, dimension(:), allocatable, intent(inout) :: vector_b
        integer(c_int32_t) :: rc
        if (allocated(vector_b)) then
            if (size(vector_b) /= size(vector_x)) deallocate(vector_b)
        end if
        if (.not. allocated(vector_b)) allocate(vector_b(size(vector_x)))
        rc = ifs_matvec(m_ctx, vector_x, vector_b, int(size(vector_x), c_int64_t))
        if (rc /= 0) then
            print '(a, i0)', "ifs_matvec status ", rc
            error stop 3
        end if
    end subroutine

end module ifs_bind

program ifs_caller
    use ifs_bind
    implicit none
    type(ifs_cluster) :: cl
    type(ifs_params) :: pr
    type(ifs_solve_opts) :: so
    type(ifs_solve_info) :: si
    real(c_double), target :: pts(3), scaled(3), xmd(4)
    real(c_double), target :: centres(9), radii(3), mrel(6)
    integer(c_int32_t), target :: lmax(3)
    integer(c_int64_t) :: u(1), i4(189), n
    integer(c_int32_t) :: rc, ni4
    character(len=16) :: mode
    double complex, allocatable :: x(:), b(:)
    integer :: i

    ! struct layouts as a C compiler lays them out (tests/test_fortran_abi.py compares)
    print '(a, 4(1x, i0))', "sizes", c_sizeof(cl), c_sizeof(pr), c_sizeof(so), c_sizeof(si)

    mode = ""
    if (command_argument_count() > 0) call get_command_argument(1, mode)

    if (trim(mode) /= "gpu") then
        ! argument errors come back as IFS_E_ARG (-1), nothing is printed by the library
        cl%n_spheres = 0
        cl%centres = c_null_ptr
        cl%radii = c_null_ptr
        cl%m_rel = c_null_ptr
        cl%lmax = c_null_ptr
        pr%k = 1d0
        pr%n_global = 1
        rc = ifs_create(m_ctx, cl, pr)
        print '(a, i0)', "ifs_create(n_spheres = 0) ", rc
        print '(a, i0)', "ifs_vector_length(null) ", ifs_vector_length(c_null_ptr)
        allocate(x(1), b(1))
        x = (1d0, 0d0)
        print '(a, i0)', "ifs_matvec(null) ", ifs_matvec(c_null_ptr, x, b, 1_c_int64_t)
        ! reference fixture (lib_tree_helper_functions.f90:2166-2579): the point
        ! (0.75, 0.5, 2^-9 + 2^-8) lies in box 6 at level 1
        scaled = (/ 0.75d0, 0.5d0, 2d0**(-9) + 2d0**(-8) /)
        rc = ifs_tree_universal_index(scaled, 1, 1, u)
        print '(a, i0, 1x, i0)', "universal_index level 1 ", rc, u(1)
        ni4 = ifs_tree_domain_i4(35_c_int64_t, 2, i4, 189)
        print '(a, i0)', "i4(35, level 2) count ", ni4
        pts = (/ 1d0, 2d0, 3d0 /)
        rc = ifs_tree_scale(pts, 1, scaled, xmd)
        print '(a, i0)', "tree_scale ", rc
        stop 0
    end if

    ! one matvec of three spheres through the calc_vector_b-shaped wrapper
    centres = (/ 0d0, 0d0, 0d0, 2.1d0, 0.3d0, -0.4d0, -1.5d0, 1.9d0, 0.8d0 /)
    radii = (/ 0.8d0, 0.6d0, 0.7d0 /)
    mrel = (/ 1.333d0, 1d-9, 1.333d0, 1d-9, 1.333d0, 1d-9 /)
    lmax = (/ 6, 5, 6 /)
    cl = ifs_cluster(3, c_loc(centres), c_loc(radii), c_loc(mrel), c_loc(lmax))
    pr%k = 2d0 * acos(-1d0) / 0.55d0 / 4d0
    pr%n_global = 6
    pr%z_selector = 3
    pr%ref_row_zeroing = 0
    pr%use_fmm = 0
    pr%leaf_size = 10
    pr%fmm_d0 = 6d0
    pr%fmm_lmax_cap = 0
    pr%workspace_mb = 0
    pr%d_convention = 0
    pr%cache_mb = 0
    pr%precision = 0
    rc = ifs_create(m_ctx, cl, pr)
    if (rc /= 0) then
        print '(a, i0)', "ifs_create status ", rc
        error stop 2
    end if
    n = ifs_vector_length(m_ctx)
    allocate(x(n))
    do i = 1, int(n)
        x(i) = cmplx(cos(0.37d0 * i), sin(0.11d0 * i), kind=kind(1d0))
    end do
    call lib_ifs_calculate_vector_b(x, b)
    print '(a, i0, a, es22.15)', "matvec n ", n, " |b - x| ", sqrt(sum(abs(b - x)**2))
    print '(a, 2es24.16)', "b(1) ", b(1)
    rc = ifs_destroy(m_ctx)
end program ifs_caller
