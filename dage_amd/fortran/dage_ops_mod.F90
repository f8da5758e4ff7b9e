!> ISO_C_BINDING interface to libdage_ops.so (include/dage_ops.h) and a thin
!! Operator3D-style wrapper, so Fortran callers can swap the reference's
!! CUDA path (src/bubbles/cuda.F90, function3d.F90 Operator3D_cuda_init /
!! transform_cuda_cube) for the gfx950 library.
!!
!! Handles are opaque type(c_ptr).  Every C function returns an integer
!! status; the wrappers below stop with the library's message on failure,
!! which is how the reference reports errors (streamcontainer.h:67-73).
!! Cubes are ordinary real(8) arrays cube(nx,ny,nz) in Fortran order.
module dage_ops_m
    use, intrinsic :: iso_c_binding
    implicit none
    private

    integer(c_int), parameter, public :: DAGE_OK = 0
    integer(c_int), parameter, public :: DAGE_PREC_FP64 = 0
    integer(c_int), parameter, public :: DAGE_OP_NO_CODA = 1
    integer(c_int), parameter :: DAGE_PREC_BF16X6 = 1
    integer(c_int), parameter :: DAGE_PREC_BF16X3 = 2

    public :: dage_gauss_quad, dage_ctx_create, dage_ctx_destroy, dage_ctx_synchronize
    public :: dage_grid3d_create, dage_grid3d_destroy, dage_grid3d_shape
    public :: dage_grid3d_coord, dage_grid3d_weights
    public :: dage_coulomb3d_create, dage_operator_set_weights, dage_operator_set_energy
    public :: dage_operator_apply, dage_operator_destroy, dage_last_error_f
    public :: dage_bubbles_create, dage_bubbles_set, dage_bubbles_destroy
    public :: dage_bubbles_create_like, dage_bubbles_get, dage_bubbles_coulomb_transform
    public :: dage_bubbles_helmholtz_transform, dage_bubbles_multiply, dage_bubbles_integrate
    public :: dage_operator_operate_on, dage_fun_write, dage_fun_read_info
    public :: dage_gbfmm_create, dage_gbfmm_apply, dage_gbfmm_box_count
    public :: dage_gbfmm_get_multipoles, dage_gbfmm_destroy, dage_gbfmm_set_energy
    public :: dage_bubbles_inject
    public :: dage_function3d_contaminants, dage_function3d_taylor_bubbles
    public :: dage_function3d_multiply_taylor
    public :: dage_operator_apply_device_strided, dage_bubbles_inject_device_strided
    public :: dage_helmholtz_batch_create, dage_batch_upload, dage_batch_download
    public :: dage_batch_set_energies, dage_batch_apply, dage_batch_tune, dage_batch_destroy
    public :: DAGE_PREC_BF16X6, DAGE_PREC_BF16X3
    public :: DageOperator3D, dage_check

    interface
        integer(c_int) function dage_gauss_quad(nlin, nlog, tstart, tlin, tlog, tp, tw) &
                bind(C, name="dage_gauss_quad")
            import :: c_int, c_double
            integer(c_int), value :: nlin, nlog
            real(c_double), value :: tstart, tlin, tlog
            real(c_double)        :: tp(*), tw(*)
        end function

        integer(c_int) function dage_ctx_create(device, ctx) bind(C, name="dage_ctx_create")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr)           :: ctx
        end function

        integer(c_int) function dage_ctx_destroy(ctx) bind(C, name="dage_ctx_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function dage_ctx_synchronize(ctx) bind(C, name="dage_ctx_synchronize")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function dage_grid3d_create(ctx, qmin, ncell, nlip, sx, sy, sz, &
                                                   grid_type, grid) bind(C, name="dage_grid3d_create")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: ctx
            real(c_double)        :: qmin(3)
            integer(c_int)        :: ncell(3)
            integer(c_int), value :: nlip, grid_type
            real(c_double)        :: sx(*), sy(*), sz(*)
            type(c_ptr)           :: grid
        end function

        integer(c_int) function dage_grid3d_destroy(grid) bind(C, name="dage_grid3d_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: grid
        end function

        integer(c_int) function dage_grid3d_coord(grid, axis, points) bind(C, name="dage_grid3d_coord")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: grid
            integer(c_int), value :: axis
            real(c_double)        :: points(*)
        end function

        integer(c_int) function dage_grid3d_weights(grid, axis, weights) &
                bind(C, name="dage_grid3d_weights")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: grid
            integer(c_int), value :: axis
            real(c_double)        :: weights(*)
        end function

        integer(c_int) function dage_grid3d_shape(grid, shape) bind(C, name="dage_grid3d_shape")
            import :: c_int, c_ptr
            type(c_ptr), value :: grid
            integer(c_int)     :: shape(3)
        end function

        integer(c_int) function dage_coulomb3d_create(ctx, gridin, gridout, tp, tw, ntp, tlog, &
                                                      flags, op) bind(C, name="dage_coulomb3d_create")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: ctx, gridin, gridout
            real(c_double)        :: tp(*), tw(*)
            integer(c_int), value :: ntp, flags
            real(c_double), value :: tlog
            type(c_ptr)           :: op
        end function

        integer(c_int) function dage_operator_set_weights(op, w) &
                bind(C, name="dage_operator_set_weights")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: op
            real(c_double)     :: w(*)
        end function

        integer(c_int) function dage_operator_set_energy(op, energy) &
                bind(C, name="dage_operator_set_energy")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: op
            real(c_double), value :: energy
        end function

        integer(c_int) function dage_operator_apply(op, cubein, cubeout, precision) &
                bind(C, name="dage_operator_apply")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: op
            real(c_double)        :: cubein(*), cubeout(*)
            integer(c_int), value :: precision
        end function

        integer(c_int) function dage_operator_destroy(op) bind(C, name="dage_operator_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: op
        end function

        integer(c_int) function dage_bubbles_create(ctx, nbub, lmax, k, nlip, bubbles) &
                bind(C, name="dage_bubbles_create")
            import :: c_int, c_ptr
            type(c_ptr), value    :: ctx
            integer(c_int), value :: nbub, lmax, k, nlip
            type(c_ptr)           :: bubbles
        end function

        integer(c_int) function dage_bubbles_set(bubbles, ibub, center, charge, ncell, steps, bf) &
                bind(C, name="dage_bubbles_set")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: bubbles
            integer(c_int), value :: ibub, ncell
            real(c_double)        :: center(3)
            real(c_double), value :: charge
            real(c_double)        :: steps(*), bf(*)
        end function

        ! Bubbles_eval_3dgrid on a host cube: cube = cube + multiplier * bubbles (l >= lmin)
        integer(c_int) function dage_bubbles_inject(bubbles, grid, lmin, multiplier, cube) &
                bind(C, name="dage_bubbles_inject")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: bubbles, grid
            integer(c_int), value :: lmin
            real(c_double), value :: multiplier
            real(c_double)        :: cube(*)
        end function

        ! sub-cubes: device boxes with leading dimensions (cube.cu:1042-1177)
        integer(c_int) function dage_operator_apply_device_strided(op, cin, ldx_in, ldy_in, &
                cout, ldx_out, ldy_out, precision) bind(C, name="dage_operator_apply_device_strided")
            import :: c_int, c_ptr
            type(c_ptr), value    :: op, cin, cout
            integer(c_int), value :: ldx_in, ldy_in, ldx_out, ldy_out, precision
        end function

        integer(c_int) function dage_bubbles_inject_device_strided(bubbles, grid, lmin, &
                multiplier, cube, ldx, ldy) bind(C, name="dage_bubbles_inject_device_strided")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: bubbles, grid, cube
            integer(c_int), value :: lmin, ldx, ldy
            real(c_double), value :: multiplier
        end function

        ! Taylor-series terms of cusp-type products (function3d.F90:1310-1470);
        ! cubes are device pointers (dage_cube_device_pointer)
        integer(c_int) function dage_function3d_contaminants(grid, cube, bubbles, order, out) &
                bind(C, name="dage_function3d_contaminants")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: grid, cube, bubbles
            integer(c_int), value :: order
            real(c_double)        :: out(*)
        end function

        integer(c_int) function dage_function3d_taylor_bubbles(grid, cube, bubbles, order, out) &
                bind(C, name="dage_function3d_taylor_bubbles")
            import :: c_int, c_ptr
            type(c_ptr), value    :: grid, cube, bubbles
            integer(c_int), value :: order
            type(c_ptr)           :: out
        end function

        integer(c_int) function dage_function3d_multiply_taylor(grid, cube1, b1, cube2, b2, &
                order, cube_out, bub_out) bind(C, name="dage_function3d_multiply_taylor")
            import :: c_int, c_ptr
            type(c_ptr), value    :: grid, cube1, b1, cube2, b2, cube_out, bub_out
            integer(c_int), value :: order
        end function

        integer(c_int) function dage_bubbles_destroy(bubbles) bind(C, name="dage_bubbles_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: bubbles
        end function

        ! Bubbles(src, lmax=, k=) without content (bubbles.F90:802-830)
        integer(c_int) function dage_bubbles_create_like(src, lmax, k, bubbles) &
                bind(C, name="dage_bubbles_create_like")
            import :: c_int, c_ptr
            type(c_ptr), value    :: src
            integer(c_int), value :: lmax, k
            type(c_ptr)           :: bubbles
        end function

        integer(c_int) function dage_bubbles_get(bubbles, ibub, center, charge, steps, bf) &
                bind(C, name="dage_bubbles_get")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: bubbles
            integer(c_int), value :: ibub
            real(c_double)        :: center(3), charge, steps(*), bf(*)
        end function

        ! Coulomb3D_evaluate_bubbles_potential (coulomb.F90:222-336)
        integer(c_int) function dage_bubbles_coulomb_transform(bin, bout) &
                bind(C, name="dage_bubbles_coulomb_transform")
            import :: c_int, c_ptr
            type(c_ptr), value :: bin, bout
        end function

        ! Helmholtz3D_transform_bubbles_sub (helmholtz3d.F90:519-657)
        integer(c_int) function dage_bubbles_helmholtz_transform(bin, energy, bout) &
                bind(C, name="dage_bubbles_helmholtz_transform")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: bin, bout
            real(c_double), value :: energy
        end function

        ! BubblesMultiplier_multiply, simple product (bubbles.F90:3553-3760)
        integer(c_int) function dage_bubbles_multiply(a, b, bout) bind(C, name="dage_bubbles_multiply")
            import :: c_int, c_ptr
            type(c_ptr), value :: a, b, bout
        end function

        integer(c_int) function dage_bubbles_integrate(bubbles, res) &
                bind(C, name="dage_bubbles_integrate")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: bubbles
            real(c_double)     :: res
        end function

        ! Operator3D_operate_on_Function3D (function3d.F90:3727-3792) on
        ! device cubes (cube_in, cube_out are device addresses)
        integer(c_int) function dage_operator_operate_on(op, cube_in, bin, cube_out, bout, &
                                                         only_cube) bind(C, name="dage_operator_operate_on")
            import :: c_int, c_ptr
            type(c_ptr), value    :: op, cube_in, bin, cube_out, bout
            integer(c_int), value :: only_cube
        end function

        ! Function3D_dump / Function3D_load (function3d.F90:812-935), host arrays
        integer(c_int) function dage_fun_write(path, label, ftype, shape, cube, nbub, lmax, k, &
                                               charges, centers, npts, bf) bind(C, name="dage_fun_write")
            import :: c_int, c_ptr, c_double, c_char
            character(kind=c_char) :: path(*), label(*)
            integer(c_int), value  :: ftype, nbub, lmax, k
            integer(c_int)         :: shape(3), npts(*)
            real(c_double)         :: cube(*), charges(*), centers(*)
            type(c_ptr)            :: bf(*)
        end function

        integer(c_int) function dage_fun_read_info(path, label, ftype, shape, nbub, lmax, k) &
                bind(C, name="dage_fun_read_info")
            import :: c_int, c_char
            character(kind=c_char) :: path(*), label(*)
            integer(c_int)         :: ftype, shape(3), nbub, lmax, k
        end function

        ! GBFMMCoulomb3D (gbfmm_coulomb3d.F90:306-1330)
        integer(c_int) function dage_gbfmm_create(ctx, grid, maxlevel, lmax, tp, tw, ntp, fmm) &
                bind(C, name="dage_gbfmm_create")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: ctx, grid
            integer(c_int), value :: maxlevel, lmax, ntp
            real(c_double)        :: tp(*), tw(*)
            type(c_ptr)           :: fmm
        end function

        integer(c_int) function dage_gbfmm_apply(fmm, rho, pot, precision) &
                bind(C, name="dage_gbfmm_apply")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: fmm
            real(c_double)        :: rho(*), pot(*)
            integer(c_int), value :: precision
        end function

        integer(c_int) function dage_gbfmm_box_count(fmm, nbox) bind(C, name="dage_gbfmm_box_count")
            import :: c_int, c_ptr
            type(c_ptr), value :: fmm
            integer(c_int)     :: nbox
        end function

        integer(c_int) function dage_gbfmm_get_multipoles(fmm, q) &
                bind(C, name="dage_gbfmm_get_multipoles")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: fmm
            real(c_double)     :: q(*)
        end function

        ! GBFMMHelmholtz3D_set_energy: E < -0.005 gives the Helmholtz operator
        integer(c_int) function dage_gbfmm_set_energy(fmm, energy) &
                bind(C, name="dage_gbfmm_set_energy")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: fmm
            real(c_double), value :: energy
        end function

        integer(c_int) function dage_gbfmm_destroy(fmm) bind(C, name="dage_gbfmm_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: fmm
        end function

        ! per-orbital Helmholtz applies as one batch (scf_cycle.F90:2927-3054)
        integer(c_int) function dage_helmholtz_batch_create(ctx, nrhs, grids, energies, tp, tw, &
                ntp, tlog, flags, precision, batch) bind(C, name="dage_helmholtz_batch_create")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: ctx
            integer(c_int), value :: nrhs, ntp, flags, precision
            type(c_ptr)           :: grids(*)
            real(c_double)        :: energies(*), tp(*), tw(*)
            real(c_double), value :: tlog
            type(c_ptr)           :: batch
        end function

        integer(c_int) function dage_batch_upload(batch, i, cube) bind(C, name="dage_batch_upload")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: batch
            integer(c_int), value :: i
            real(c_double)        :: cube(*)
        end function

        integer(c_int) function dage_batch_download(batch, i, cube) &
                bind(C, name="dage_batch_download")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value    :: batch
            integer(c_int), value :: i
            real(c_double)        :: cube(*)
        end function

        integer(c_int) function dage_batch_set_energies(batch, energies) &
                bind(C, name="dage_batch_set_energies")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: batch
            real(c_double)     :: energies(*)
        end function

        integer(c_int) function dage_batch_apply(batch) bind(C, name="dage_batch_apply")
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
        end function

        integer(c_int) function dage_batch_tune(batch, reps, tiles) bind(C, name="dage_batch_tune")
            import :: c_int, c_ptr
            type(c_ptr), value    :: batch
            integer(c_int), value :: reps
            integer(c_int)        :: tiles(*)
        end function

        integer(c_int) function dage_batch_destroy(batch) bind(C, name="dage_batch_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: batch
        end function

        type(c_ptr) function dage_last_error() bind(C, name="dage_last_error")
            import :: c_ptr
        end function
    end interface

    !> Operator3D-shaped handle: gridin -> gridout with a t-quadrature
    !! (coulomb.F90:109-176; function3d.F90:295-336).
    type :: DageOperator3D
        type(c_ptr) :: ctx = c_null_ptr
        type(c_ptr) :: gridin = c_null_ptr
        type(c_ptr) :: gridout = c_null_ptr
        type(c_ptr) :: op = c_null_ptr
        integer     :: dims_in(3) = 0, dims_out(3) = 0
    contains
        procedure :: transform_cube => DageOperator3D_transform_cube
        procedure :: set_energy     => DageOperator3D_set_energy
        procedure :: destroy        => DageOperator3D_destroy
    end type

    interface DageOperator3D
        module procedure :: DageOperator3D_init_coulomb
    end interface

contains

    !> Message of the last failed call on this thread.
    function dage_last_error_f() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char), pointer :: chars(:)
        integer :: n
        type(c_ptr) :: p
        p = dage_last_error()
        call c_f_pointer(p, chars, [1024])
        n = 0
        do while (n < 1024)
            if (chars(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: msg)
        msg = transfer(chars(1:n), msg)
    end function

    subroutine dage_check(status, what)
        integer(c_int), intent(in)   :: status
        character(len=*), intent(in) :: what
        if (status /= DAGE_OK) then
            write(*, '(a, a, a, i0, a, a)') "dage: ", what, " failed (", status, "): ", &
                dage_last_error_f()
            stop 1
        end if
    end subroutine

    !> Coulomb3D on an equidistant-cell grid: qmin, ncell, one step per axis.
    function DageOperator3D_init_coulomb(device, qmin, ncell, step, nlip, nlin, nlog, tlin, &
                                         tlog, coda) result(new)
        integer, intent(in)      :: device, ncell(3), nlip, nlin, nlog
        real(8), intent(in)      :: qmin(3), step(3), tlin, tlog
        logical, intent(in)      :: coda
        type(DageOperator3D)     :: new
        real(c_double), allocatable :: sx(:), sy(:), sz(:), tp(:), tw(:)
        integer(c_int) :: shp(3), ncl(3), flags
        real(c_double) :: qm(3)

        call dage_check(dage_ctx_create(int(device, c_int), new%ctx), "dage_ctx_create")
        allocate(sx(ncell(1)), sy(ncell(2)), sz(ncell(3)))
        sx = step(1)
        sy = step(2)
        sz = step(3)
        qm = qmin
        ncl = ncell
        call dage_check(dage_grid3d_create(new%ctx, qm, ncl, int(nlip, c_int), sx, sy, sz, &
                                           1_c_int, new%gridin), "dage_grid3d_create")
        new%gridout = new%gridin
        call dage_check(dage_grid3d_shape(new%gridin, shp), "dage_grid3d_shape")
        new%dims_in = shp
        new%dims_out = shp
        allocate(tp(nlin + nlog), tw(nlin + nlog))
        call dage_check(dage_gauss_quad(int(nlin, c_int), int(nlog, c_int), 0.0_c_double, &
                                        real(tlin, c_double), real(tlog, c_double), tp, tw), &
                        "dage_gauss_quad")
        flags = 0
        if (.not. coda) flags = DAGE_OP_NO_CODA
        call dage_check(dage_coulomb3d_create(new%ctx, new%gridin, new%gridout, tp, tw, &
                                              int(nlin + nlog, c_int), real(tlog, c_double), &
                                              flags, new%op), "dage_coulomb3d_create")
    end function

    !> Operator3D_transform_cube (function3d.F90:4230-4349) on the GPU.
    function DageOperator3D_transform_cube(self, cubein) result(cubeout)
        class(DageOperator3D), intent(in) :: self
        real(8), intent(in)               :: cubein(:,:,:)
        real(8), allocatable, target      :: cubeout(:,:,:)
        real(c_double), allocatable       :: tmp(:,:,:)
        allocate(cubeout(self%dims_out(1), self%dims_out(2), self%dims_out(3)))
        tmp = cubein
        call dage_check(dage_operator_apply(self%op, tmp, cubeout, DAGE_PREC_FP64), &
                        "dage_operator_apply")
    end function

    !> Helmholtz3D_set_energy (helmholtz3d.F90:411-444).
    subroutine DageOperator3D_set_energy(self, energy)
        class(DageOperator3D), intent(inout) :: self
        real(8), intent(in)                  :: energy
        call dage_check(dage_operator_set_energy(self%op, real(energy, c_double)), &
                        "dage_operator_set_energy")
    end subroutine

    subroutine DageOperator3D_destroy(self)
        class(DageOperator3D), intent(inout) :: self
        integer(c_int) :: rc
        rc = dage_operator_destroy(self%op)
        rc = dage_grid3d_destroy(self%gridin)
        rc = dage_ctx_destroy(self%ctx)
        self%op = c_null_ptr
        self%gridin = c_null_ptr
        self%gridout = c_null_ptr
        self%ctx = c_null_ptr
    end subroutine
end module
