!> iso_c_binding interfaces of libweno_mi355 (include/weno_c.h).
module m_weno_c

    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: WENO_OK = 0
    integer(c_int), parameter :: WENO_DIR_X = 0, WENO_DIR_Y = 1, WENO_DIR_Z = 2
    integer(c_int), parameter :: WENO_SRC_FIELD = 0, WENO_SRC_RESHAPED = 1
    integer(c_int), parameter :: WENO_F64 = 0, WENO_F32 = 1, WENO_BF16 = 2
    integer(c_int), parameter :: WENO_FUSE_MP = 1, WENO_FUSE_COPY = 2, WENO_STREAMING = 4
    integer(c_int), parameter :: WENO_ONLY_L = 8, WENO_ONLY_R = 16

    type, bind(C) :: weno_grid_t
        integer(c_int) :: m_cells, n_cells, p_cells
        integer(c_int) :: v_size
        integer(c_int) :: buff_size
        real(c_double) :: weno_eps
        integer(c_int) :: wenojs
        integer(c_int) :: teno
        real(c_double) :: teno_CT
        integer(c_int) :: mapped_weno
        integer(c_int) :: wenoz
    end type weno_grid_t

    type, bind(C) :: weno_coeffs_t
        type(c_ptr) :: polyL, polyR, dL, dR, beta
    end type weno_coeffs_t

    interface
        integer(c_int) function weno_init(device) bind(C, name='weno_init')
            import :: c_int
            integer(c_int), value :: device
        end function weno_init

        integer(c_int) function weno_finalize() bind(C, name='weno_finalize')
            import :: c_int
        end function weno_finalize

        type(c_ptr) function weno_last_error() bind(C, name='weno_last_error')
            import :: c_ptr
        end function weno_last_error

        integer(c_int) function weno_sync() bind(C, name='weno_sync')
            import :: c_int
        end function weno_sync

        integer(c_int) function weno_device_count() bind(C, name='weno_device_count')
            import :: c_int
        end function weno_device_count

        integer(c_int) function weno_target_enter_data(host, bytes) bind(C, name='weno_target_enter_data')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: host
            integer(c_size_t), value :: bytes
        end function weno_target_enter_data

        integer(c_int) function weno_target_update_to(host, bytes) bind(C, name='weno_target_update_to')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: host
            integer(c_size_t), value :: bytes
        end function weno_target_update_to

        integer(c_int) function weno_target_update_from(host, bytes) bind(C, name='weno_target_update_from')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: host
            integer(c_size_t), value :: bytes
        end function weno_target_update_from

        integer(c_int) function weno_target_exit_data(host) bind(C, name='weno_target_exit_data')
            import :: c_int, c_ptr
            type(c_ptr), value :: host
        end function weno_target_exit_data

        integer(c_int) function weno_reshape(g, dir, dtype, v_vf, v_rs) bind(C, name='weno_reshape')
            import :: c_int, c_ptr, weno_grid_t
            type(weno_grid_t), intent(in) :: g
            integer(c_int), value :: dir, dtype
            type(c_ptr), value :: v_vf, v_rs
        end function weno_reshape

        integer(c_int) function weno_order1_copy(g, dir, dtype, v_vf, vL, vR) bind(C, name='weno_order1_copy')
            import :: c_int, c_ptr, weno_grid_t
            type(weno_grid_t), intent(in) :: g
            integer(c_int), value :: dir, dtype
            type(c_ptr), value :: v_vf, vL, vR
        end function weno_order1_copy

        integer(c_int) function weno_reconstruct(g, order, dir, src_kind, dtype, src, c, vL, vR, flags) &
            bind(C, name='weno_reconstruct')
            import :: c_int, c_ptr, weno_grid_t, weno_coeffs_t
            type(weno_grid_t), intent(in) :: g
            integer(c_int), value :: order, dir, src_kind, dtype
            type(c_ptr), value :: src
            type(weno_coeffs_t), intent(in) :: c
            type(c_ptr), value :: vL, vR
            integer(c_int), value :: flags
        end function weno_reconstruct

        integer(c_int) function weno_mp_limit(g, dir, src_kind, dtype, src, vL, vR) bind(C, name='weno_mp_limit')
            import :: c_int, c_ptr, weno_grid_t
            type(weno_grid_t), intent(in) :: g
            integer(c_int), value :: dir, src_kind, dtype
            type(c_ptr), value :: src, vL, vR
        end function weno_mp_limit

        integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function c_strlen
    end interface

contains

    !> Stops the program with the library's message when rc /= 0.
    subroutine weno_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer(c_size_t) :: n
        character(len=512) :: text
        integer :: i

        if (rc == WENO_OK) return
        p = weno_last_error()
        text = ''
        if (c_associated(p)) then
            n = min(c_strlen(p), int(len(text), c_size_t))
            call c_f_pointer(p, msg, [n])
            do i = 1, int(n)
                text(i:i) = msg(i)
            end do
        end if
        write(*,'(a,a,a,i0,a,a)') '  ERROR in ', what, ' (status ', rc, '): ', trim(text)
        error stop 1
    end subroutine weno_check

end module m_weno_c
