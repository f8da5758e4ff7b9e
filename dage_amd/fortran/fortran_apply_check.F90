!> Drop-in check from Fortran: rebuilds the reference's gauss_test.F90 TEST 1
!! (alpha = 1 Gaussian at the origin on the GaussGenerator default grid,
!! GaussQuad(10, 8, tlin 2, tlog 500), coda on) through the C-ABI and prints
!! the self-interaction energy sum w_x w_y w_z rho V with the LIP weights.
!! With argument "quad" it only checks the host-side quadrature (no GPU);
!! with "gbfmm" it computes the same self-interaction with the GBFMM Coulomb
!! potential (49^3 grid over [-6,6]^3, 4x4x4 boxes, lmax 6) through the
!! dage_gbfmm_* interfaces.
program fortran_apply_check
    use, intrinsic :: iso_c_binding
    use dage_ops_m
    implicit none
    type(DageOperator3D) :: op
    real(8), allocatable :: rho(:,:,:), v(:,:,:), x(:), w(:)
    real(c_double)       :: tp(18), tw(18)
    character(len=16)    :: arg
    integer              :: i, j, k, n
    real(8)              :: pi
    real(8)              :: e, sj, sk, lo, hi, h, ext, cell, dx

    pi = acos(-1.0d0)
    call get_command_argument(1, arg)
    call dage_check(dage_gauss_quad(10_c_int, 8_c_int, 0.0_c_double, 2.0_c_double, &
                                    500.0_c_double, tp, tw), "dage_gauss_quad")
    if (trim(arg) == "quad") then
        write(*, '(a, es24.16, a, es24.16)') "tp(1)=", tp(1), " tw(18)=", tw(18)
        stop 0
    end if
    if (trim(arg) == "gbfmm") then
        call gbfmm_selfint(tp, tw)
        stop 0
    end if

    ! GaussGenerator grid for one Gaussian (gauss_generator.F90:262-281):
    ! Grid1D_init_spheres (grid.F90:603-640) then Grid1D_init_equidistant
    ! (grid.F90:539-567), step 0.2, nlip 7
    ext = sqrt(log(1.0d0 * (1.0d0 / pi)**1.5d0 / 1.0d-12) / 1.0d0)
    lo = -ext
    hi = ext
    cell = 0.2d0 * (7 - 1)
    n = ceiling((hi - lo) / cell)
    dx = n * cell - (hi - lo)
    lo = lo - 0.5d0 * dx
    hi = hi + 0.5d0 * dx
    n = ceiling((hi - lo) / (0.2d0 * (7 - 1)))
    h = (hi - lo) / (n * (7 - 1))
    op = DageOperator3D(0, [lo, lo, lo], [n, n, n], [h, h, h], 7, 10, 8, 2.0d0, 500.0d0, .true.)
    allocate(x(op%dims_in(1)), w(op%dims_in(1)))
    call dage_check(dage_grid3d_coord(op%gridin, 0_c_int, x), "dage_grid3d_coord")
    call dage_check(dage_grid3d_weights(op%gridin, 0_c_int, w), "dage_grid3d_weights")
    allocate(rho(size(x), size(x), size(x)))
    do k = 1, size(x)
        do j = 1, size(x)
            do i = 1, size(x)
                ! GaussGenerator_gen_cubef: q*norm*ex*ey*ez (gauss_generator.F90:291-318)
                rho(i, j, k) = 1.0d0 * (1.0d0 / pi)**1.5d0 * exp(-x(i)**2) * exp(-x(j)**2) &
                               * exp(-x(k)**2)
            end do
        end do
    end do
    v = op%transform_cube(rho)
    ! LIP-weighted integral, summed axis by axis (x innermost)
    e = 0.0d0
    do k = 1, size(x)
        sk = 0.0d0
        do j = 1, size(x)
            sj = 0.0d0
            do i = 1, size(x)
                sj = sj + w(i) * (rho(i, j, k) * v(i, j, k))
            end do
            sk = sk + w(j) * sj
        end do
        e = e + w(k) * sk
    end do
    write(*, '(a, es24.16)') "selfint=", e
    call op%destroy()

contains

    !> GBFMMCoulomb3D_calculate_potential from Fortran: the alpha = 1 Gaussian's
    !! self-interaction with the FMM potential.
    subroutine gbfmm_selfint(tp, tw)
        real(c_double), intent(in) :: tp(18), tw(18)
        type(c_ptr)          :: ctx, grid, fmm
        real(c_double)       :: qmin(3), steps(8), tpc(18), twc(18)
        integer(c_int)       :: ncl(3), nbox
        real(8), allocatable :: xg(:), wg(:), rho(:,:,:), pot(:,:,:)
        real(8)              :: pi, e, sj, sk
        integer              :: i, j, k

        pi = acos(-1.0d0)
        tpc = tp
        twc = tw
        qmin = -6.0d0
        steps = 12.0d0 / 48.0d0
        ncl = 8
        call dage_check(dage_ctx_create(0_c_int, ctx), "dage_ctx_create")
        call dage_check(dage_grid3d_create(ctx, qmin, ncl, 7_c_int, steps, steps, steps, 1_c_int, &
                                           grid), "dage_grid3d_create")
        allocate(xg(49), wg(49), rho(49, 49, 49), pot(49, 49, 49))
        call dage_check(dage_grid3d_coord(grid, 0_c_int, xg), "dage_grid3d_coord")
        call dage_check(dage_grid3d_weights(grid, 0_c_int, wg), "dage_grid3d_weights")
        do k = 1, 49
            do j = 1, 49
                do i = 1, 49
                    rho(i, j, k) = (1.0d0 / pi)**1.5d0 * exp(-(xg(i)**2 + xg(j)**2 + xg(k)**2))
                end do
            end do
        end do
        call dage_check(dage_gbfmm_create(ctx, grid, 2_c_int, 6_c_int, tpc, twc, 18_c_int, fmm), &
                        "dage_gbfmm_create")
        call dage_check(dage_gbfmm_apply(fmm, rho, pot, DAGE_PREC_FP64), "dage_gbfmm_apply")
        call dage_check(dage_gbfmm_box_count(fmm, nbox), "dage_gbfmm_box_count")
        e = 0.0d0
        do k = 1, 49
            sk = 0.0d0
            do j = 1, 49
                sj = 0.0d0
                do i = 1, 49
                    sj = sj + wg(i) * (rho(i, j, k) * pot(i, j, k))
                end do
                sk = sk + wg(j) * sj
            end do
            e = e + wg(k) * sk
        end do
        write(*, '(a, es24.16, a, i0)') "gbfmm_selfint=", e, " boxes=", nbox
        call dage_check(dage_gbfmm_destroy(fmm), "dage_gbfmm_destroy")
        call dage_check(dage_grid3d_destroy(grid), "dage_grid3d_destroy")
        call dage_check(dage_ctx_destroy(ctx), "dage_ctx_destroy")
    end subroutine
end program
