! Writes small .fun and .g3d files with Fortran unformatted stream I/O in the
! layout of DAGE's Function3D_dump / Bubbles_write (src/bubbles/function3d.F90:
! 812-879, src/bubbles/bubbles.F90:1251-1276) and Grid3D_dump
! (src/bubbles/grid.F90:1128-1145), so that the library's C++ reader and
! writer can be checked byte for byte against what a Fortran compiler emits.
! Build and run: tests/golden/make_fun_fixture.sh (amdflang).
program fun_fixture
    use iso_fortran_env, only: int32, int64, real64
    implicit none
    integer, parameter :: fd = 31
    character(len=20) :: label
    integer :: ftype, i, j, k, ib
    integer(int64) :: grid_shape(3), cubestart, bubblestart
    real(real64) :: cube(3, 4, 5)
    integer(int32) :: nbub, lmax, kk
    real(real64) :: z(2), centers(3, 2), bf(13, 4, 2), steps(2, 3), qmin(3)
    integer(int32) :: ncell(3), nlip, grid_type

    label = "fixture"
    ftype = 1
    do k = 1, 5
        do j = 1, 4
            do i = 1, 3
                cube(i, j, k) = 0.25d0 * i - 1.5d0 * j + 0.125d0 * k * k
            end do
        end do
    end do
    nbub = 2
    lmax = 1
    kk = -1
    z = [3.0d0, 1.0d0]
    centers(:, 1) = [0.5d0, -0.25d0, 1.0d0]
    centers(:, 2) = [-1.0d0, 2.0d0, 0.0d0]
    do ib = 1, 2
        do j = 1, 4
            do i = 1, 13
                bf(i, j, ib) = ib * 100.0d0 + j * 10.0d0 + i * 0.5d0
            end do
        end do
    end do

    open(unit=fd, file="fixture.fun", access="stream", status="replace")
    write(fd, pos=25) label
    write(fd) ftype
    inquire(fd, pos=cubestart)
    grid_shape = [3_int64, 4_int64, 5_int64]
    do i = 1, 3
        write(fd) grid_shape(i)
    end do
    write(fd) cube
    inquire(fd, pos=bubblestart)
    write(fd) nbub
    write(fd) nbub
    write(fd) lmax
    write(fd) kk
    do ib = 1, 2
        write(fd) z(ib)
        write(fd) centers(:, ib)
        write(fd) int(ib, int32)
    end do
    do ib = 1, 2
        write(fd) centers(:, ib)
        write(fd) z(ib)
    end do
    do ib = 1, 2
        write(fd) bf(:, :, ib)
    end do
    write(fd, pos=1) cubestart
    write(fd, pos=17) bubblestart
    close(fd)

    nlip = 7
    grid_type = 1
    ncell = [2, 2, 2]
    qmin = [-1.0d0, -2.0d0, -3.0d0]
    steps(:, 1) = [0.1d0, 0.2d0]
    steps(:, 2) = [0.3d0, 0.15d0]
    steps(:, 3) = [0.05d0, 0.5d0]
    open(unit=fd, file="fixture.g3d", access="stream", status="replace")
    write(fd) nlip
    write(fd) grid_type
    do i = 1, 3
        write(fd) ncell(i)
        write(fd) nlip
        write(fd) qmin(i)
        write(fd) steps(:, i)
    end do
    close(fd)
end program
