!> Harness driver over the MI355X library: same namelist file argument,
!! WENO_* environment switches, kernel sequences and stdout lines as the
!! reference's program weno_test_driver (inputs/weno_standalone.f90:1734-1861),
!! so `wenoss.sh test`-style checksum parsing works unchanged.
!!
!! Extra switches: WENO_MAPPED=1, WENO_Z=1, WENO_TENO=1 (mapped-WENO, WENO-Z
!! or TENO weights when wenojs=F, in that precedence), WENO5_FUSED=1
!! (weno5 mode: one pass per sweep from v_vf, order-1 halo included),
!! WENO_METRICS_JSON=<path> (the run's configuration, checksums and times
!! as one JSON object in that file; non-finite checksums are written as
!! null, and "status" says which of the three closing lines was printed).
program weno_test_driver

    use, intrinsic :: iso_c_binding
    use, intrinsic :: ieee_arithmetic
    use m_weno_c
    use m_weno_standalone
    implicit none

    real(wp), allocatable, target :: vL_x(:,:,:,:), vR_x(:,:,:,:)
    real(wp), allocatable, target :: vL_y(:,:,:,:), vR_y(:,:,:,:)
    real(wp), allocatable, target :: vL_z(:,:,:,:), vR_z(:,:,:,:)
    real(wp) :: chk_x, chk_y, chk_z, chk_total
    real(wp) :: elapsed_total, elapsed_average
    integer(8) :: c_start, c_end, c_rate
    integer :: warmup_iters, benchmark_iters, iter
    character(len=16) :: kernel_mode
    character(len=16) :: weno5_bench_scope
    character(len=256) :: input_file
    character(len=256) :: metrics_path
    integer(c_size_t) :: nb

    input_file = 'inputs/weno_input.nml'
    if (command_argument_count() >= 1) call get_command_argument(1, input_file)

    warmup_iters = max(0, f_get_env_int('WENO_WARMUP_ITERS', 1))
    benchmark_iters = max(1, f_get_env_int('WENO_BENCH_ITERS', 5))
    call s_get_env_string('WENO_KERNEL_MODE', 'all', kernel_mode)
    call s_get_env_string('WENO5_BENCH_SCOPE', 'full', weno5_bench_scope)
    weno5_split_kernels = f_get_env_int('WENO5_SPLIT_KERNELS', 0) /= 0
    weno5_specialized_combined = f_get_env_int('WENO5_SPECIALIZED_COMBINED', 0) /= 0
    weno5_fused = f_get_env_int('WENO5_FUSED', 0) /= 0
    teno = f_get_env_int('WENO_TENO', 0) /= 0
    mapped_weno = f_get_env_int('WENO_MAPPED', 0) /= 0
    wenoz = f_get_env_int('WENO_Z', 0) /= 0

    write(*,'(a)') '==============================================='
    write(*,'(a)') '  Standalone MFC WENO harness (MI355X HIP path)'
    write(*,'(a)') '==============================================='
    write(*,'(a,i0)') '  HIP devices available    : ', weno_device_count()
    write(*,'(a,i0)') '  Default device           : ', 0
    write(*,'(a,a)')  '  Kernel mode              : ', trim(kernel_mode)
    write(*,'(a,a)')  '  WENO5 benchmark scope    : ', trim(weno5_bench_scope)
    write(*,'(a,l1)') '  WENO5 split kernels      : ', weno5_split_kernels
    write(*,'(a,l1)') '  WENO5 specialized comb.  : ', weno5_specialized_combined
    write(*,'(a,l1)') '  WENO5 fused sweeps       : ', weno5_fused
    write(*,'(a,i0)') '  Warmup iterations        : ', warmup_iters
    write(*,'(a,i0)') '  Benchmark iterations     : ', benchmark_iters

    call s_read_input(trim(input_file))
    call s_initialize_data()

    allocate(vL_x(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    allocate(vR_x(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    allocate(vL_y(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    allocate(vR_y(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    allocate(vL_z(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    allocate(vR_z(is1_beg:is1_end, is2_beg:is2_end, is3_beg:is3_end, 1:v_size))
    vL_x = 0.0_wp; vR_x = 0.0_wp
    vL_y = 0.0_wp; vR_y = 0.0_wp
    vL_z = 0.0_wp; vR_z = 0.0_wp

    nb = 8_c_size_t*size(vL_x, kind=c_size_t)
    call s_map_to(c_loc(vL_x), nb)
    call s_map_to(c_loc(vR_x), nb)
    call s_map_to(c_loc(vL_y), nb)
    call s_map_to(c_loc(vR_y), nb)
    call s_map_to(c_loc(vL_z), nb)
    call s_map_to(c_loc(vR_z), nb)

    select case (trim(kernel_mode))
    case ('weno5')
        if (trim(weno5_bench_scope) == 'kernel') then
            call s_prepare_weno5_inputs(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
            do iter = 1, warmup_iters
                call s_run_weno5_reconstruction(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
            end do
        else
            do iter = 1, warmup_iters
                call s_run_weno5_kernels(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
            end do
        end if
    case default
        do iter = 1, warmup_iters
            call s_run_all_kernels(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        end do
    end select

    call weno_check(weno_sync(), 'sync')
    call system_clock(c_start, c_rate)
    select case (trim(kernel_mode))
    case ('weno5')
        if (trim(weno5_bench_scope) == 'kernel') then
            do iter = 1, benchmark_iters
                call s_run_weno5_reconstruction(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
            end do
        else
            do iter = 1, benchmark_iters
                call s_run_weno5_kernels(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
            end do
        end if
    case default
        do iter = 1, benchmark_iters
            call s_run_all_kernels(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
        end do
    end select
    call weno_check(weno_sync(), 'sync')
    call system_clock(c_end)

    call weno_check(weno_target_update_from(c_loc(vL_x), nb), 'update from')
    call weno_check(weno_target_update_from(c_loc(vR_x), nb), 'update from')
    call weno_check(weno_target_update_from(c_loc(vL_y), nb), 'update from')
    call weno_check(weno_target_update_from(c_loc(vR_y), nb), 'update from')
    call weno_check(weno_target_update_from(c_loc(vL_z), nb), 'update from')
    call weno_check(weno_target_update_from(c_loc(vR_z), nb), 'update from')

    elapsed_total = real(c_end - c_start, wp)/real(c_rate, wp)
    elapsed_average = elapsed_total/real(benchmark_iters, wp)

    chk_x = f_checksum(vL_x) + f_checksum(vR_x)
    chk_y = f_checksum(vL_y) + f_checksum(vR_y)
    chk_z = f_checksum(vL_z) + f_checksum(vR_z)
    chk_total = chk_x + chk_y + chk_z

    write(*,'(a,es18.10)') '  Checksum X : ', chk_x
    write(*,'(a,es18.10)') '  Checksum Y : ', chk_y
    write(*,'(a,es18.10)') '  Checksum Z : ', chk_z
    write(*,'(a,es18.10)') '  Checksum T : ', chk_total
    write(*,'(a,f12.6,a)') '  Kernel time total        : ', elapsed_total, ' s'
    write(*,'(a,f12.6,a)') '  Kernel time per iter     : ', elapsed_average, ' s'

    if (.not. ieee_is_finite(chk_total)) then
        write(*,'(a)') '  WARNING: non-finite values detected in standalone output.'
    else if (chk_total == 0.0_wp) then
        write(*,'(a)') '  WARNING: all-zero output; offload may not have executed.'
    else
        write(*,'(a)') '  All kernel families executed with finite non-zero output.'
    end if

    call s_get_env_string('WENO_METRICS_JSON', '', metrics_path)
    if (len_trim(metrics_path) > 0) call s_write_metrics(trim(metrics_path))

    call weno_check(weno_target_exit_data(c_loc(vL_x)), 'exit data')
    call weno_check(weno_target_exit_data(c_loc(vR_x)), 'exit data')
    call weno_check(weno_target_exit_data(c_loc(vL_y)), 'exit data')
    call weno_check(weno_target_exit_data(c_loc(vR_y)), 'exit data')
    call weno_check(weno_target_exit_data(c_loc(vL_z)), 'exit data')
    call weno_check(weno_target_exit_data(c_loc(vR_z)), 'exit data')
    deallocate(vL_x, vR_x, vL_y, vR_y, vL_z, vR_z)
    call s_finalize_data()
    call weno_check(weno_finalize(), 'finalize')

contains

    !> JSON value of a real: the number, or null when it is not finite.
    function json_real(x) result(txt)
        real(wp), intent(in) :: x
        character(len=32) :: txt
        if (ieee_is_finite(x)) then
            write(txt, '(es24.16)') x
            txt = adjustl(txt)
        else
            txt = 'null'
        end if
    end function json_real

    function json_bool(b) result(txt)
        logical, intent(in) :: b
        character(len=5) :: txt
        txt = merge('true ', 'false', b)
    end function json_bool

    subroutine s_write_metrics(path)
        character(len=*), intent(in) :: path
        integer :: iu, ios
        character(len=16) :: status
        if (.not. ieee_is_finite(chk_total)) then
            status = 'non-finite'
        else if (chk_total == 0.0_wp) then
            status = 'all-zero'
        else
            status = 'ok'
        end if
        open(newunit=iu, file=path, status='replace', action='write', iostat=ios)
        if (ios /= 0) then
            write(*,'(3a)') '  WARNING: could not write metrics to "', path, '"'
            return
        end if
        write(iu,'(a)') '{'
        write(iu,'(3a)') '  "kernel_mode": "', trim(kernel_mode), '",'
        write(iu,'(3a)') '  "weno5_bench_scope": "', trim(weno5_bench_scope), '",'
        write(iu,'(3a)') '  "weno5_split_kernels": ', trim(json_bool(weno5_split_kernels)), ','
        write(iu,'(3a)') '  "weno5_specialized_combined": ', trim(json_bool(weno5_specialized_combined)), ','
        write(iu,'(3a)') '  "weno5_fused": ', trim(json_bool(weno5_fused)), ','
        write(iu,'(a,i0,a)') '  "m_cells": ', m_cells, ','
        write(iu,'(a,i0,a)') '  "n_cells": ', n_cells, ','
        write(iu,'(a,i0,a)') '  "p_cells": ', p_cells, ','
        write(iu,'(a,i0,a)') '  "v_size": ', v_size, ','
        write(iu,'(a,i0,a)') '  "buff_size": ', buff_size, ','
        write(iu,'(3a)') '  "weno_eps": ', trim(json_real(weno_eps)), ','
        write(iu,'(3a)') '  "wenojs": ', trim(json_bool(wenojs)), ','
        write(iu,'(3a)') '  "teno_CT": ', trim(json_real(teno_CT)), ','
        write(iu,'(a,i0,a)') '  "warmup_iters": ', warmup_iters, ','
        write(iu,'(a,i0,a)') '  "bench_iters": ', benchmark_iters, ','
        write(iu,'(3a)') '  "checksum_x": ', trim(json_real(chk_x)), ','
        write(iu,'(3a)') '  "checksum_y": ', trim(json_real(chk_y)), ','
        write(iu,'(3a)') '  "checksum_z": ', trim(json_real(chk_z)), ','
        write(iu,'(3a)') '  "checksum_t": ', trim(json_real(chk_total)), ','
        write(iu,'(3a)') '  "kernel_time_total_s": ', trim(json_real(elapsed_total)), ','
        write(iu,'(3a)') '  "kernel_time_per_iter_s": ', trim(json_real(elapsed_average)), ','
        write(iu,'(3a)') '  "status": "', trim(status), '"'
        write(iu,'(a)') '}'
        close(iu)
    end subroutine s_write_metrics

end program weno_test_driver
