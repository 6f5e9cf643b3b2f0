!> Oracle driver around the REFERENCE module m_weno_standalone -- TEST
!! INFRASTRUCTURE ONLY.  It is linked against the reference's own
!! inputs/weno_standalone.f90 compiled from source by oracle/Makefile (the
!! reference program's `main` is localised so this program can own it).
!! It runs one reference kernel sequence on a field read from a raw file and
!! dumps the outputs, so tests can compare the oracle restatement with the
!! reference element by element.
!!
!! usage: weno_ref_driver MODE M N P V BUFF EPS WENOJS FIELD OUT
!!   MODE  : reshape | copy | weno3 | weno5 | weno5split | weno5spec | weno7 |
!!           mp5 | mp7 | all
!!   FIELD : raw float64 file holding v_vf in Fortran order, or 'harness'
!!           (keep the field of s_fill_input_field)
!!   OUT   : raw float64 file; x,y,z outputs concatenated (vL then vR per
!!           direction, or v_rs_ws_{x,y,z} for MODE=reshape)
program weno_ref_driver

    use m_weno_standalone
    implicit none

    character(len=256) :: mode, fin, fout, arg
    real(wp), allocatable, target :: vL_x(:,:,:,:), vR_x(:,:,:,:)
    real(wp), allocatable, target :: vL_y(:,:,:,:), vR_y(:,:,:,:)
    real(wp), allocatable, target :: vL_z(:,:,:,:), vR_z(:,:,:,:)
    integer :: iu

    call get_command_argument(1, mode)
    call get_command_argument(2, arg); read(arg, *) m_cells
    call get_command_argument(3, arg); read(arg, *) n_cells
    call get_command_argument(4, arg); read(arg, *) p_cells
    call get_command_argument(5, arg); read(arg, *) v_size
    call get_command_argument(6, arg); read(arg, *) buff_size
    call get_command_argument(7, arg); read(arg, *) weno_eps
    call get_command_argument(8, arg); read(arg, *) wenojs
    call get_command_argument(9, fin)
    call get_command_argument(10, fout)

    call s_initialize_data()

    if (trim(fin) /= 'harness') then
        open(newunit=iu, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
        read(iu) v_vf
        close(iu)
    end if

    allocate(vL_x(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    allocate(vR_x(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    allocate(vL_y(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    allocate(vR_y(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    allocate(vL_z(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    allocate(vR_z(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    vL_x = 0.0_wp; vR_x = 0.0_wp
    vL_y = 0.0_wp; vR_y = 0.0_wp
    vL_z = 0.0_wp; vR_z = 0.0_wp

    select case (trim(mode))
    case ('reshape')
        call s_initialize_weno_x()
        call s_initialize_weno_y()
        call s_initialize_weno_z()
    case ('copy')
        call s_order1_copy_x(vL_x, vR_x)
        call s_order1_copy_y(vL_y, vR_y)
        call s_order1_copy_z(vL_z, vR_z)
    case ('weno3')
        call s_prepare_weno5_inputs(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        call s_weno3_kernel(v_rs_ws_x, poly_coef_cbL_x, poly_coef_cbR_x, d_cbL_x, d_cbR_x, beta_coef_x, vL_x, vR_x)
        call s_weno3_kernel(v_rs_ws_y, poly_coef_cbL_y, poly_coef_cbR_y, d_cbL_y, d_cbR_y, beta_coef_y, vL_y, vR_y)
        call s_weno3_kernel(v_rs_ws_z, poly_coef_cbL_z, poly_coef_cbR_z, d_cbL_z, d_cbR_z, beta_coef_z, vL_z, vR_z)
    case ('weno5', 'weno5split')
        weno5_split_kernels = (trim(mode) == 'weno5split')
        call s_prepare_weno5_inputs(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        call s_run_weno5_reconstruction(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
    case ('weno5spec')
        weno5_specialized_combined = .true.
        call s_prepare_weno5_inputs(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        call s_run_weno5_reconstruction(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
    case ('weno7', 'mp7')
        call s_prepare_weno5_inputs(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        call s_weno7_kernel(v_rs_ws_x, poly_coef_cbL_x, poly_coef_cbR_x, d_cbL_x, d_cbR_x, beta_coef_x, vL_x, vR_x)
        call s_weno7_kernel(v_rs_ws_y, poly_coef_cbL_y, poly_coef_cbR_y, d_cbL_y, d_cbR_y, beta_coef_y, vL_y, vR_y)
        call s_weno7_kernel(v_rs_ws_z, poly_coef_cbL_z, poly_coef_cbR_z, d_cbL_z, d_cbR_z, beta_coef_z, vL_z, vR_z)
        if (trim(mode) == 'mp7') then
            call s_preserve_monotonicity(v_rs_ws_x, vL_x, vR_x)
            call s_preserve_monotonicity(v_rs_ws_y, vL_y, vR_y)
            call s_preserve_monotonicity(v_rs_ws_z, vL_z, vR_z)
        end if
    case ('mp5')
        call s_prepare_weno5_inputs(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        call s_run_weno5_reconstruction(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        call s_preserve_monotonicity(v_rs_ws_x, vL_x, vR_x)
        call s_preserve_monotonicity(v_rs_ws_y, vL_y, vR_y)
        call s_preserve_monotonicity(v_rs_ws_z, vL_z, vR_z)
    case ('all')
        call s_run_all_kernels(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
    case default
        write(*,'(a,a)') 'unknown mode ', trim(mode)
        error stop 2
    end select

    open(newunit=iu, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
    if (trim(mode) == 'reshape') then
        write(iu) v_rs_ws_x, v_rs_ws_y, v_rs_ws_z
    else
        write(iu) vL_x, vR_x, vL_y, vR_y, vL_z, vR_z
    end if
    close(iu)

    write(*,'(a,es18.10)') '  Checksum X : ', f_checksum(vL_x) + f_checksum(vR_x)
    write(*,'(a,es18.10)') '  Checksum Y : ', f_checksum(vL_y) + f_checksum(vR_y)
    write(*,'(a,es18.10)') '  Checksum Z : ', f_checksum(vL_z) + f_checksum(vR_z)

    call s_finalize_data()

end program weno_ref_driver
