!> Drop-in for the reference's module m_weno_standalone
!! (inputs/weno_standalone.f90:48-1732): the same module variables, the same
!! procedure names and argument lists, with every device region replaced by a
!! call into libweno_mi355 (HIP kernels for gfx950) through m_weno_c.
!!
!! Arrays live on the host as in the reference and are mirrored on the GPU by
!! the library's data environment (weno_target_enter_data & co.), the
!! counterpart of `!$omp target enter data`.  Kernel calls pass host
!! addresses; the library maps them to the device copies.
module m_weno_standalone

    use, intrinsic :: iso_c_binding
    use m_weno_c
    implicit none

    integer, parameter :: wp = kind(1.0d0)

    ! namelist &grid_params (weno_standalone.f90:57-66)
    integer :: m_cells = 199
    integer :: n_cells = 199
    integer :: p_cells = 199
    integer :: v_size = 8
    integer :: buff_size = 6
    real(wp) :: weno_eps = 1.0e-16_wp
    logical :: wenojs = .true.
    real(wp) :: teno_CT = 1.0e-5_wp

    ! fixed stencil sizes (:69-71)
    integer, parameter :: weno_polyn_max = 3
    integer, parameter :: weno_stencils_max = 4
    integer, parameter :: beta_terms_max = 6
    ! The weight branches mapped_weno, wenoz and teno are compile-time
    ! .false. parameters in the reference (:72-74); here runtime switches
    ! (environment WENO_MAPPED=1, WENO_Z=1, WENO_TENO=1).  wenoz_q (:75) is
    ! fixed at 2 as in the reference.
    logical :: mapped_weno = .false.
    logical :: wenoz = .false.
    logical :: teno = .false.
    integer, parameter :: wenoz_q = 2

    logical :: weno5_split_kernels = .false.
    logical :: weno5_specialized_combined = .false.
    ! one pass per sweep: order-1 halo + WENO5 core straight from v_vf
    ! (environment WENO5_FUSED=1)
    logical :: weno5_fused = .false.

    integer :: is1_beg, is1_end
    integer :: is2_beg, is2_end
    integer :: is3_beg, is3_end
    integer :: core1_beg, core1_end

    real(wp), allocatable, target :: v_vf(:,:,:,:)
    real(wp), allocatable, target :: v_rs_ws_x(:,:,:,:), v_rs_ws_y(:,:,:,:), v_rs_ws_z(:,:,:,:)

    real(wp), allocatable, target :: poly_coef_cbL_x(:,:,:), poly_coef_cbR_x(:,:,:)
    real(wp), allocatable, target :: poly_coef_cbL_y(:,:,:), poly_coef_cbR_y(:,:,:)
    real(wp), allocatable, target :: poly_coef_cbL_z(:,:,:), poly_coef_cbR_z(:,:,:)
    real(wp), allocatable, target :: d_cbL_x(:,:), d_cbR_x(:,:)
    real(wp), allocatable, target :: d_cbL_y(:,:), d_cbR_y(:,:)
    real(wp), allocatable, target :: d_cbL_z(:,:), d_cbR_z(:,:)
    real(wp), allocatable, target :: beta_coef_x(:,:,:)
    real(wp), allocatable, target :: beta_coef_y(:,:,:)
    real(wp), allocatable, target :: beta_coef_z(:,:,:)

contains

    !> The library's view of the module's grid and scheme switches.
    function grid() result(g)
        type(weno_grid_t) :: g
        g%m_cells = m_cells
        g%n_cells = n_cells
        g%p_cells = p_cells
        g%v_size = v_size
        g%buff_size = buff_size
        g%weno_eps = weno_eps
        g%wenojs = merge(1, 0, wenojs)
        g%teno = merge(1, 0, teno)
        g%teno_CT = teno_CT
        g%mapped_weno = merge(1, 0, mapped_weno)
        g%wenoz = merge(1, 0, wenoz)
    end function grid

    subroutine s_read_input(input_file)
        character(len=*), intent(in) :: input_file
        integer :: iu, ios
        logical :: found

        namelist /grid_params/ m_cells, n_cells, p_cells, v_size, &
            buff_size, weno_eps, wenojs, teno_CT

        inquire(file=trim(input_file), exist=found)
        if (.not. found) then
            write(*,'(a,a,a)') '  Input file "', trim(input_file), '" not found; using defaults.'
            return
        end if
        open(newunit=iu, file=trim(input_file), status='old', action='read', iostat=ios)
        if (ios /= 0) then
            write(*,'(a,a,a)') '  WARNING: could not open "', trim(input_file), '"; using defaults.'
            return
        end if
        read(iu, nml=grid_params, iostat=ios)
        if (ios /= 0) write(*,'(a)') '  WARNING: error reading namelist grid_params; using defaults.'
        close(iu)

        write(*,'(a,i0)') '  m_cells (from input)     : ', m_cells
        write(*,'(a,i0)') '  n_cells (from input)     : ', n_cells
        write(*,'(a,i0)') '  p_cells (from input)     : ', p_cells
        write(*,'(a,i0)') '  v_size  (from input)     : ', v_size
    end subroutine s_read_input

    !> Maps a host array into the library's data environment and uploads it.
    subroutine s_map_to(p, bytes)
        type(c_ptr), intent(in) :: p
        integer(c_size_t), intent(in) :: bytes
        call weno_check(weno_target_enter_data(p, bytes), 'enter data')
        call weno_check(weno_target_update_to(p, bytes), 'update to')
    end subroutine s_map_to

    subroutine s_initialize_data()
        integer(c_size_t), parameter :: b8 = 8

        is1_beg = -buff_size
        is1_end = m_cells + buff_size
        is2_beg = -buff_size
        is2_end = n_cells + buff_size
        is3_beg = -buff_size
        is3_end = p_cells + buff_size
        core1_beg = is1_beg + weno_polyn_max
        core1_end = is1_end - weno_polyn_max

        allocate(v_vf(is1_beg-weno_polyn_max:is1_end+weno_polyn_max, &
                      is2_beg-weno_polyn_max:is2_end+weno_polyn_max, &
                      is3_beg-weno_polyn_max:is3_end+weno_polyn_max, 1:v_size))
        allocate(v_rs_ws_x(is1_beg-weno_polyn_max:is1_end+weno_polyn_max, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
        allocate(v_rs_ws_y(is1_beg-weno_polyn_max:is1_end+weno_polyn_max, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
        allocate(v_rs_ws_z(is1_beg-weno_polyn_max:is1_end+weno_polyn_max, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
        allocate(poly_coef_cbL_x(core1_beg:core1_end, 0:weno_stencils_max, 0:weno_polyn_max-1))
        allocate(poly_coef_cbR_x(core1_beg:core1_end, 0:weno_stencils_max, 0:weno_polyn_max-1))
        allocate(poly_coef_cbL_y(core1_beg:core1_end, 0:weno_stencils_max, 0:weno_polyn_max-1))
        allocate(poly_coef_cbR_y(core1_beg:core1_end, 0:weno_stencils_max, 0:weno_polyn_max-1))
        allocate(poly_coef_cbL_z(core1_beg:core1_end, 0:weno_stencils_max, 0:weno_polyn_max-1))
        allocate(poly_coef_cbR_z(core1_beg:core1_end, 0:weno_stencils_max, 0:weno_polyn_max-1))
        allocate(d_cbL_x(0:weno_stencils_max, core1_beg:core1_end), d_cbR_x(0:weno_stencils_max, core1_beg:core1_end))
        allocate(d_cbL_y(0:weno_stencils_max, core1_beg:core1_end), d_cbR_y(0:weno_stencils_max, core1_beg:core1_end))
        allocate(d_cbL_z(0:weno_stencils_max, core1_beg:core1_end), d_cbR_z(0:weno_stencils_max, core1_beg:core1_end))
        allocate(beta_coef_x(core1_beg:core1_end, 0:weno_stencils_max, 0:beta_terms_max-1))
        allocate(beta_coef_y(core1_beg:core1_end, 0:weno_stencils_max, 0:beta_terms_max-1))
        allocate(beta_coef_z(core1_beg:core1_end, 0:weno_stencils_max, 0:beta_terms_max-1))

        call s_fill_input_field()
        call s_fill_coefficients(poly_coef_cbL_x, poly_coef_cbR_x, d_cbL_x, d_cbR_x, beta_coef_x, 0.03_wp)
        call s_fill_coefficients(poly_coef_cbL_y, poly_coef_cbR_y, d_cbL_y, d_cbR_y, beta_coef_y, 0.05_wp)
        call s_fill_coefficients(poly_coef_cbL_z, poly_coef_cbR_z, d_cbL_z, d_cbR_z, beta_coef_z, 0.07_wp)

        call weno_check(weno_init(-1_c_int), 'weno_init')
        call s_map_to(c_loc(v_vf), b8*size(v_vf, kind=c_size_t))
        call weno_check(weno_target_enter_data(c_loc(v_rs_ws_x), b8*size(v_rs_ws_x, kind=c_size_t)), 'enter data')
        call weno_check(weno_target_enter_data(c_loc(v_rs_ws_y), b8*size(v_rs_ws_y, kind=c_size_t)), 'enter data')
        call weno_check(weno_target_enter_data(c_loc(v_rs_ws_z), b8*size(v_rs_ws_z, kind=c_size_t)), 'enter data')
        call s_map_to(c_loc(poly_coef_cbL_x), b8*size(poly_coef_cbL_x, kind=c_size_t))
        call s_map_to(c_loc(poly_coef_cbR_x), b8*size(poly_coef_cbR_x, kind=c_size_t))
        call s_map_to(c_loc(poly_coef_cbL_y), b8*size(poly_coef_cbL_y, kind=c_size_t))
        call s_map_to(c_loc(poly_coef_cbR_y), b8*size(poly_coef_cbR_y, kind=c_size_t))
        call s_map_to(c_loc(poly_coef_cbL_z), b8*size(poly_coef_cbL_z, kind=c_size_t))
        call s_map_to(c_loc(poly_coef_cbR_z), b8*size(poly_coef_cbR_z, kind=c_size_t))
        call s_map_to(c_loc(d_cbL_x), b8*size(d_cbL_x, kind=c_size_t))
        call s_map_to(c_loc(d_cbR_x), b8*size(d_cbR_x, kind=c_size_t))
        call s_map_to(c_loc(d_cbL_y), b8*size(d_cbL_y, kind=c_size_t))
        call s_map_to(c_loc(d_cbR_y), b8*size(d_cbR_y, kind=c_size_t))
        call s_map_to(c_loc(d_cbL_z), b8*size(d_cbL_z, kind=c_size_t))
        call s_map_to(c_loc(d_cbR_z), b8*size(d_cbR_z, kind=c_size_t))
        call s_map_to(c_loc(beta_coef_x), b8*size(beta_coef_x, kind=c_size_t))
        call s_map_to(c_loc(beta_coef_y), b8*size(beta_coef_y, kind=c_size_t))
        call s_map_to(c_loc(beta_coef_z), b8*size(beta_coef_z, kind=c_size_t))
    end subroutine s_initialize_data

    subroutine s_finalize_data()
        call weno_check(weno_target_exit_data(c_loc(v_vf)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(v_rs_ws_x)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(v_rs_ws_y)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(v_rs_ws_z)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(poly_coef_cbL_x)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(poly_coef_cbR_x)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(poly_coef_cbL_y)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(poly_coef_cbR_y)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(poly_coef_cbL_z)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(poly_coef_cbR_z)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(d_cbL_x)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(d_cbR_x)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(d_cbL_y)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(d_cbR_y)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(d_cbL_z)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(d_cbR_z)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(beta_coef_x)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(beta_coef_y)), 'exit data')
        call weno_check(weno_target_exit_data(c_loc(beta_coef_z)), 'exit data')
        deallocate(v_vf, v_rs_ws_x, v_rs_ws_y, v_rs_ws_z)
        deallocate(poly_coef_cbL_x, poly_coef_cbR_x, poly_coef_cbL_y, poly_coef_cbR_y)
        deallocate(poly_coef_cbL_z, poly_coef_cbR_z, d_cbL_x, d_cbR_x, d_cbL_y, d_cbR_y)
        deallocate(d_cbL_z, d_cbR_z, beta_coef_x, beta_coef_y, beta_coef_z)
    end subroutine s_finalize_data

    !> The harness's synthetic field (weno_standalone.f90:221-244): the
    !! reference's golden checksums are defined on it.
    subroutine s_fill_input_field()
        integer :: i, j, k, l
        real(wp) :: x, y, z, sx, cy, sz

        do l = 1, v_size
            do k = lbound(v_vf, 3), ubound(v_vf, 3)
                z = real(k - is3_beg + 1, wp)/real(max(1, is3_end - is3_beg + 3), wp)
                sz = 0.25_wp*sin(4.3_wp*z)
                do j = lbound(v_vf, 2), ubound(v_vf, 2)
                    y = real(j - is2_beg + 1, wp)/real(max(1, is2_end - is2_beg + 3), wp)
                    cy = cos(3.7_wp*y)
                    do i = lbound(v_vf, 1), ubound(v_vf, 1)
                        x = real(i - is1_beg + 1, wp)/real(max(1, is1_end - is1_beg + 3), wp)
                        sx = sin(2.1_wp*x*real(l, wp))
                        v_vf(i, j, k, l) = sx + cy + sz + 0.02_wp*real(i + 2*j - k + l, wp)
                    end do
                end do
            end do
        end do
    end subroutine s_fill_input_field

    !> The harness's synthetic, j-constant coefficients (:246-279).
    subroutine s_fill_coefficients(polyL, polyR, dL, dR, beta_coef, offset)
        real(wp), intent(out) :: polyL(core1_beg:, 0:, 0:)
        real(wp), intent(out) :: polyR(core1_beg:, 0:, 0:)
        real(wp), intent(out) :: dL(0:, core1_beg:)
        real(wp), intent(out) :: dR(0:, core1_beg:)
        real(wp), intent(out) :: beta_coef(core1_beg:, 0:, 0:)
        real(wp), intent(in) :: offset
        real(wp), parameter :: wL(0:4) = [0.30_wp, 0.27_wp, 0.20_wp, 0.14_wp, 0.09_wp]
        integer :: j, r, s

        do j = core1_beg, core1_end
            do r = 0, weno_stencils_max
                dL(r, j) = wL(r) + offset*0.10_wp
                dR(r, j) = wL(weno_stencils_max - r) + offset*0.10_wp
                do s = 0, weno_polyn_max - 1
                    polyL(j, r, s) = offset + 0.015_wp*real((r + 1)*(s + 1), wp)
                    polyR(j, r, s) = -offset + 0.012_wp*real(2*r + s + 1, wp)
                end do
                beta_coef(j, r, :) = 0.0_wp
                beta_coef(j, r, 0) = 0.8_wp + offset + 0.05_wp*real(r + 1, wp)
                beta_coef(j, r, 2) = 0.9_wp + offset + 0.04_wp*real(r + 1, wp)
                beta_coef(j, r, 3) = 1.0_wp + offset + 0.03_wp*real(r + 1, wp)
                beta_coef(j, r, 5) = 1.1_wp + offset + 0.02_wp*real(r + 1, wp)
            end do
        end do
    end subroutine s_fill_coefficients

    ! ---- reshape (:281-330): a transpose kernel per direction ---------------
    subroutine s_initialize_weno_x()
        call weno_check(weno_reshape(grid(), WENO_DIR_X, WENO_F64, c_loc(v_vf), c_loc(v_rs_ws_x)), 'reshape x')
    end subroutine s_initialize_weno_x

    subroutine s_initialize_weno_y()
        call weno_check(weno_reshape(grid(), WENO_DIR_Y, WENO_F64, c_loc(v_vf), c_loc(v_rs_ws_y)), 'reshape y')
    end subroutine s_initialize_weno_y

    subroutine s_initialize_weno_z()
        call weno_check(weno_reshape(grid(), WENO_DIR_Z, WENO_F64, c_loc(v_vf), c_loc(v_rs_ws_z)), 'reshape z')
    end subroutine s_initialize_weno_z

    ! ---- order-1 copy (:332-390) --------------------------------------------
    subroutine s_order1_copy_x(vL_x, vR_x)
        real(wp), intent(inout), target, contiguous :: vL_x(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_x(is1_beg:, is2_beg:, is3_beg:, 1:)
        call weno_check(weno_order1_copy(grid(), WENO_DIR_X, WENO_F64, c_loc(v_vf), c_loc(vL_x), c_loc(vR_x)), &
                        'order1 copy x')
    end subroutine s_order1_copy_x

    subroutine s_order1_copy_y(vL_y, vR_y)
        real(wp), intent(inout), target, contiguous :: vL_y(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_y(is1_beg:, is2_beg:, is3_beg:, 1:)
        call weno_check(weno_order1_copy(grid(), WENO_DIR_Y, WENO_F64, c_loc(v_vf), c_loc(vL_y), c_loc(vR_y)), &
                        'order1 copy y')
    end subroutine s_order1_copy_y

    subroutine s_order1_copy_z(vL_z, vR_z)
        real(wp), intent(inout), target, contiguous :: vL_z(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_z(is1_beg:, is2_beg:, is3_beg:, 1:)
        call weno_check(weno_order1_copy(grid(), WENO_DIR_Z, WENO_F64, c_loc(v_vf), c_loc(vL_z), c_loc(vR_z)), &
                        'order1 copy z')
    end subroutine s_order1_copy_z

    !> Direction of a reshaped buffer passed to the 8-argument family.
    integer(c_int) function f_dir_of(v_rs_ws) result(d)
        real(wp), intent(in), target, contiguous :: v_rs_ws(:,:,:,:)
        if (c_associated(c_loc(v_rs_ws), c_loc(v_rs_ws_x))) then
            d = WENO_DIR_X
        else if (c_associated(c_loc(v_rs_ws), c_loc(v_rs_ws_y))) then
            d = WENO_DIR_Y
        else if (c_associated(c_loc(v_rs_ws), c_loc(v_rs_ws_z))) then
            d = WENO_DIR_Z
        else
            write(*,'(a)') '  ERROR: v_rs_ws must be one of v_rs_ws_x/y/z'
            error stop 1
        end if
    end function f_dir_of

    !> One call of the 8-argument family on the reshaped input.
    subroutine s_reconstruct(order, v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR, flags, what)
        integer(c_int), intent(in) :: order, flags
        real(wp), intent(in), target, contiguous :: v_rs_ws(:,:,:,:)
        real(wp), intent(in), target, contiguous :: polyL(:,:,:), polyR(:,:,:), dL(:,:), dR(:,:), beta_coef(:,:,:)
        real(wp), intent(inout), target, contiguous :: vL(:,:,:,:), vR(:,:,:,:)
        character(len=*), intent(in) :: what
        type(weno_coeffs_t) :: c
        c%polyL = c_loc(polyL)
        c%polyR = c_loc(polyR)
        c%dL = c_loc(dL)
        c%dR = c_loc(dR)
        c%beta = c_loc(beta_coef)
        call weno_check(weno_reconstruct(grid(), order, f_dir_of(v_rs_ws), WENO_SRC_RESHAPED, WENO_F64, &
                                         c_loc(v_rs_ws), c, c_loc(vL), c_loc(vR), flags), what)
    end subroutine s_reconstruct

    subroutine s_weno3_kernel(v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR)
        real(wp), intent(in), target, contiguous :: v_rs_ws(is1_beg-weno_polyn_max:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(in), target, contiguous :: polyL(core1_beg:, 0:, 0:), polyR(core1_beg:, 0:, 0:)
        real(wp), intent(in), target, contiguous :: dL(0:, core1_beg:), dR(0:, core1_beg:)
        real(wp), intent(in), target, contiguous :: beta_coef(core1_beg:, 0:, 0:)
        real(wp), intent(inout), target, contiguous :: vL(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR(is1_beg:, is2_beg:, is3_beg:, 1:)
        call s_reconstruct(3, v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR, 0, 'weno3')
    end subroutine s_weno3_kernel

    subroutine s_weno5_kernel(v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR)
        real(wp), intent(in), target, contiguous :: v_rs_ws(is1_beg-weno_polyn_max:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(in), target, contiguous :: polyL(core1_beg:, 0:, 0:), polyR(core1_beg:, 0:, 0:)
        real(wp), intent(in), target, contiguous :: dL(0:, core1_beg:), dR(0:, core1_beg:)
        real(wp), intent(in), target, contiguous :: beta_coef(core1_beg:, 0:, 0:)
        real(wp), intent(inout), target, contiguous :: vL(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR(is1_beg:, is2_beg:, is3_beg:, 1:)
        call s_reconstruct(5, v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR, 0, 'weno5')
    end subroutine s_weno5_kernel

    !> s_weno5_left_kernel (:1145): vL only.
    subroutine s_weno5_left_kernel(v_rs_ws, polyL, dL, beta_coef, vL)
        real(wp), intent(in), target, contiguous :: v_rs_ws(is1_beg-weno_polyn_max:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(in), target, contiguous :: polyL(core1_beg:, 0:, 0:)
        real(wp), intent(in), target, contiguous :: dL(0:, core1_beg:)
        real(wp), intent(in), target, contiguous :: beta_coef(core1_beg:, 0:, 0:)
        real(wp), intent(inout), target, contiguous :: vL(is1_beg:, is2_beg:, is3_beg:, 1:)
        call s_reconstruct(5, v_rs_ws, polyL, polyL, dL, dL, beta_coef, vL, vL, WENO_ONLY_L, 'weno5 left')
    end subroutine s_weno5_left_kernel

    !> s_weno5_right_kernel (:1272): vR only.
    subroutine s_weno5_right_kernel(v_rs_ws, polyR, dR, beta_coef, vR)
        real(wp), intent(in), target, contiguous :: v_rs_ws(is1_beg-weno_polyn_max:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(in), target, contiguous :: polyR(core1_beg:, 0:, 0:)
        real(wp), intent(in), target, contiguous :: dR(0:, core1_beg:)
        real(wp), intent(in), target, contiguous :: beta_coef(core1_beg:, 0:, 0:)
        real(wp), intent(inout), target, contiguous :: vR(is1_beg:, is2_beg:, is3_beg:, 1:)
        call s_reconstruct(5, v_rs_ws, polyR, polyR, dR, dR, beta_coef, vR, vR, WENO_ONLY_R, 'weno5 right')
    end subroutine s_weno5_right_kernel

    subroutine s_weno5_kernel_dispatch(v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR)
        real(wp), intent(in), target, contiguous :: v_rs_ws(is1_beg-weno_polyn_max:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(in), target, contiguous :: polyL(core1_beg:, 0:, 0:), polyR(core1_beg:, 0:, 0:)
        real(wp), intent(in), target, contiguous :: dL(0:, core1_beg:), dR(0:, core1_beg:)
        real(wp), intent(in), target, contiguous :: beta_coef(core1_beg:, 0:, 0:)
        real(wp), intent(inout), target, contiguous :: vL(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR(is1_beg:, is2_beg:, is3_beg:, 1:)
        if (weno5_split_kernels) then
            call s_weno5_left_kernel(v_rs_ws, polyL, dL, beta_coef, vL)
            call s_weno5_right_kernel(v_rs_ws, polyR, dR, beta_coef, vR)
        else
            call s_weno5_kernel(v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR)
        end if
    end subroutine s_weno5_kernel_dispatch

    subroutine s_weno7_kernel(v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR)
        real(wp), intent(in), target, contiguous :: v_rs_ws(is1_beg-weno_polyn_max:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(in), target, contiguous :: polyL(core1_beg:, 0:, 0:), polyR(core1_beg:, 0:, 0:)
        real(wp), intent(in), target, contiguous :: dL(0:, core1_beg:), dR(0:, core1_beg:)
        real(wp), intent(in), target, contiguous :: beta_coef(core1_beg:, 0:, 0:)
        real(wp), intent(inout), target, contiguous :: vL(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR(is1_beg:, is2_beg:, is3_beg:, 1:)
        call s_reconstruct(7, v_rs_ws, polyL, polyR, dL, dR, beta_coef, vL, vR, 0, 'weno7')
    end subroutine s_weno7_kernel

    ! ---- specialised WENO5 (:671-1143): direction-aware loads from v_vf -----
    subroutine s_weno5_field(dir, polyL, polyR, dL, dR, beta_coef, vL, vR, flags, what)
        integer(c_int), intent(in) :: dir, flags
        real(wp), intent(in), target, contiguous :: polyL(:,:,:), polyR(:,:,:), dL(:,:), dR(:,:), beta_coef(:,:,:)
        real(wp), intent(inout), target, contiguous :: vL(:,:,:,:), vR(:,:,:,:)
        character(len=*), intent(in) :: what
        type(weno_coeffs_t) :: c
        c%polyL = c_loc(polyL)
        c%polyR = c_loc(polyR)
        c%dL = c_loc(dL)
        c%dR = c_loc(dR)
        c%beta = c_loc(beta_coef)
        call weno_check(weno_reconstruct(grid(), 5, dir, WENO_SRC_FIELD, WENO_F64, c_loc(v_vf), c, &
                                         c_loc(vL), c_loc(vR), flags), what)
    end subroutine s_weno5_field

    subroutine s_weno5_kernel_x(vL, vR)
        real(wp), intent(inout), target, contiguous :: vL(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR(is1_beg:, is2_beg:, is3_beg:, 1:)
        call s_weno5_field(WENO_DIR_X, poly_coef_cbL_x, poly_coef_cbR_x, d_cbL_x, d_cbR_x, beta_coef_x, vL, vR, &
                           0, 'weno5 x')
    end subroutine s_weno5_kernel_x

    subroutine s_weno5_kernel_y(vL, vR)
        real(wp), intent(inout), target, contiguous :: vL(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR(is1_beg:, is2_beg:, is3_beg:, 1:)
        call s_weno5_field(WENO_DIR_Y, poly_coef_cbL_y, poly_coef_cbR_y, d_cbL_y, d_cbR_y, beta_coef_y, vL, vR, &
                           0, 'weno5 y')
    end subroutine s_weno5_kernel_y

    subroutine s_weno5_kernel_z(vL, vR)
        real(wp), intent(inout), target, contiguous :: vL(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR(is1_beg:, is2_beg:, is3_beg:, 1:)
        call s_weno5_field(WENO_DIR_Z, poly_coef_cbL_z, poly_coef_cbR_z, d_cbL_z, d_cbR_z, beta_coef_z, vL, vR, &
                           0, 'weno5 z')
    end subroutine s_weno5_kernel_z

    ! ---- MP limiter (:1544-1615) ----------------------------------------------
    subroutine s_preserve_monotonicity(v_rs_ws, vL, vR)
        real(wp), intent(in), target, contiguous :: v_rs_ws(is1_beg-weno_polyn_max:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vL(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR(is1_beg:, is2_beg:, is3_beg:, 1:)
        call weno_check(weno_mp_limit(grid(), f_dir_of(v_rs_ws), WENO_SRC_RESHAPED, WENO_F64, c_loc(v_rs_ws), &
                                      c_loc(vL), c_loc(vR)), 'mp')
    end subroutine s_preserve_monotonicity

    ! ---- utilities (:1617-1649) -------------------------------------------------
    real(wp) function f_checksum(arr) result(chk)
        real(wp), intent(in) :: arr(:,:,:,:)
        chk = sum(abs(arr))
    end function f_checksum

    integer function f_get_env_int(name, default_value) result(value)
        character(len=*), intent(in) :: name
        integer, intent(in) :: default_value
        character(len=64) :: buffer
        integer :: length, status, parsed

        value = default_value
        buffer = ''
        call get_environment_variable(name, buffer, length=length, status=status)
        if (status /= 0 .or. length <= 0) return
        read(buffer(1:length), *, iostat=status) parsed
        if (status == 0) value = parsed
    end function f_get_env_int

    subroutine s_get_env_string(name, default_value, value)
        character(len=*), intent(in) :: name
        character(len=*), intent(in) :: default_value
        character(len=*), intent(out) :: value
        integer :: length, status

        value = default_value
        call get_environment_variable(name, value, length=length, status=status)
        if (status /= 0 .or. length <= 0) value = default_value
    end subroutine s_get_env_string

    ! ---- sequences (:1651-1730) ---------------------------------------------
    subroutine s_run_all_kernels(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        real(wp), intent(inout), target, contiguous :: vL_x(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_x(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vL_y(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_y(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vL_z(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_z(is1_beg:, is2_beg:, is3_beg:, 1:)

        call s_prepare_weno5_inputs(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        call s_weno3_kernel(v_rs_ws_x, poly_coef_cbL_x, poly_coef_cbR_x, d_cbL_x, d_cbR_x, beta_coef_x, vL_x, vR_x)
        call s_weno3_kernel(v_rs_ws_y, poly_coef_cbL_y, poly_coef_cbR_y, d_cbL_y, d_cbR_y, beta_coef_y, vL_y, vR_y)
        call s_weno3_kernel(v_rs_ws_z, poly_coef_cbL_z, poly_coef_cbR_z, d_cbL_z, d_cbR_z, beta_coef_z, vL_z, vR_z)
        call s_weno5_kernel_dispatch(v_rs_ws_x, poly_coef_cbL_x, poly_coef_cbR_x, d_cbL_x, d_cbR_x, beta_coef_x, &
                                     vL_x, vR_x)
        call s_weno5_kernel_dispatch(v_rs_ws_y, poly_coef_cbL_y, poly_coef_cbR_y, d_cbL_y, d_cbR_y, beta_coef_y, &
                                     vL_y, vR_y)
        call s_weno5_kernel_dispatch(v_rs_ws_z, poly_coef_cbL_z, poly_coef_cbR_z, d_cbL_z, d_cbR_z, beta_coef_z, &
                                     vL_z, vR_z)
        call s_weno7_kernel(v_rs_ws_x, poly_coef_cbL_x, poly_coef_cbR_x, d_cbL_x, d_cbR_x, beta_coef_x, vL_x, vR_x)
        call s_weno7_kernel(v_rs_ws_y, poly_coef_cbL_y, poly_coef_cbR_y, d_cbL_y, d_cbR_y, beta_coef_y, vL_y, vR_y)
        call s_weno7_kernel(v_rs_ws_z, poly_coef_cbL_z, poly_coef_cbR_z, d_cbL_z, d_cbR_z, beta_coef_z, vL_z, vR_z)
        call s_preserve_monotonicity(v_rs_ws_x, vL_x, vR_x)
        call s_preserve_monotonicity(v_rs_ws_y, vL_y, vR_y)
        call s_preserve_monotonicity(v_rs_ws_z, vL_z, vR_z)
    end subroutine s_run_all_kernels

    subroutine s_run_weno5_kernels(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        real(wp), intent(inout), target, contiguous :: vL_x(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_x(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vL_y(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_y(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vL_z(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_z(is1_beg:, is2_beg:, is3_beg:, 1:)

        if (weno5_fused) then
            ! order-1 halo and WENO5 core of each sweep in one pass from v_vf
            call s_weno5_field(WENO_DIR_X, poly_coef_cbL_x, poly_coef_cbR_x, d_cbL_x, d_cbR_x, beta_coef_x, &
                               vL_x, vR_x, WENO_FUSE_COPY, 'fused x')
            call s_weno5_field(WENO_DIR_Y, poly_coef_cbL_y, poly_coef_cbR_y, d_cbL_y, d_cbR_y, beta_coef_y, &
                               vL_y, vR_y, WENO_FUSE_COPY, 'fused y')
            call s_weno5_field(WENO_DIR_Z, poly_coef_cbL_z, poly_coef_cbR_z, d_cbL_z, d_cbR_z, beta_coef_z, &
                               vL_z, vR_z, WENO_FUSE_COPY, 'fused z')
            return
        end if
        call s_prepare_weno5_inputs(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        call s_run_weno5_reconstruction(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
    end subroutine s_run_weno5_kernels

    subroutine s_prepare_weno5_inputs(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        real(wp), intent(inout), target, contiguous :: vL_x(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_x(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vL_y(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_y(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vL_z(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_z(is1_beg:, is2_beg:, is3_beg:, 1:)

        call s_initialize_weno_x()
        call s_initialize_weno_y()
        call s_initialize_weno_z()
        call s_order1_copy_x(vL_x, vR_x)
        call s_order1_copy_y(vL_y, vR_y)
        call s_order1_copy_z(vL_z, vR_z)
    end subroutine s_prepare_weno5_inputs

    subroutine s_run_weno5_reconstruction(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        real(wp), intent(inout), target, contiguous :: vL_x(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_x(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vL_y(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_y(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vL_z(is1_beg:, is2_beg:, is3_beg:, 1:)
        real(wp), intent(inout), target, contiguous :: vR_z(is1_beg:, is2_beg:, is3_beg:, 1:)

        if (weno5_specialized_combined .and. .not. weno5_split_kernels) then
            call s_weno5_kernel_x(vL_x, vR_x)
            call s_weno5_kernel_y(vL_y, vR_y)
            call s_weno5_kernel_z(vL_z, vR_z)
        else
            call s_weno5_kernel_dispatch(v_rs_ws_x, poly_coef_cbL_x, poly_coef_cbR_x, d_cbL_x, d_cbR_x, &
                                         beta_coef_x, vL_x, vR_x)
            call s_weno5_kernel_dispatch(v_rs_ws_y, poly_coef_cbL_y, poly_coef_cbR_y, d_cbL_y, d_cbR_y, &
                                         beta_coef_y, vL_y, vR_y)
            call s_weno5_kernel_dispatch(v_rs_ws_z, poly_coef_cbL_z, poly_coef_cbR_z, d_cbL_z, d_cbR_z, &
                                         beta_coef_z, vL_z, vR_z)
        end if
    end subroutine s_run_weno5_reconstruction

end module m_weno_standalone
