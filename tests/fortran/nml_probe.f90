!> Test helper (no GPU): reads a namelist file with the harness module's
!! s_read_input and prints the resulting grid parameters on one line.
program nml_probe
    use m_weno_standalone
    implicit none
    character(len=256) :: input_file
    input_file = 'missing.nml'
    if (command_argument_count() >= 1) call get_command_argument(1, input_file)
    call s_read_input(trim(input_file))
    write(*,'(a,5(1x,i0),1x,es12.5,1x,l1,1x,es12.5)') 'VALUES', m_cells, n_cells, p_cells, v_size, buff_size, &
        weno_eps, wenojs, teno_CT
end program nml_probe
