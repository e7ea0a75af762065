! main.f90 -- config.h-driven driver of the GPU LS-SPH solver.  The sequence
! is the reference's main program (source/main.f90:53-166): parameters and
! checks, particles, markers, cell list, initial output (with the initial
! ghost fill), then the time loop of RK + PST (or the steady check of the
! diffusion test) with periodic output.  With USE_GRAPH = 1 each loop
! iteration's RK + PST is one replay of a captured HIP graph.
!
! LSSPH_SETUP_ONLY=<file> in the environment writes the host setup (after
! init_particle) to <file> in the write_data layout and stops before any GPU
! call.

#include "config.h"
#include "lssph_defaults.h"

program main
    use iso_c_binding
    use iso_fortran_env, only: int64
    use lssph_binding
    use global_types            , only: Param_type, SP_type, VM_type, Cell_type
    use init_param_mod          , only: init_param
    use check_param_mod         , only: check_param
    use init_particle_mod       , only: init_particle
    use init_virtual_markers_mod, only: init_virtual_markers
    use init_cell_mod           , only: init_cell
    use update_cell_mod         , only: update_cell
    use RK2_mod                 , only: RK2
    use RK4_mod                 , only: RK4
    use PST_mod                 , only: PST
    use check_steady_state_mod  , only: check_steady_state
    use write_param_mod         , only: write_param
    use write_data_mod          , only: write_data
    use write_progress_mod      , only: write_init_info, write_progress
    implicit none

    type(Param_type) :: param
    type(SP_type)    :: SP
    type(VM_type)    :: VM
    type(Cell_type)  :: cell
    integer          :: step, is_steady, loop_start, log_interval, n, un
    integer(int64)   :: t_start, t_now, rate
    character(len=1024) :: setup_file

    call init_param(param)
    call check_param(param)
    call init_particle(param, SP)

    call get_environment_variable("LSSPH_SETUP_ONLY", setup_file, n)
    if (n > 0) then
        open(newunit=un, file=trim(setup_file), status="replace", form="unformatted", access="stream")
        write(un) SP%ptype(1:SP%num_total)
        write(un) SP%x(1:SP%num_total)
        write(un) SP%y(1:SP%num_total)
        write(un) SP%u(1:SP%num_total)
        write(un) SP%v(1:SP%num_total)
        write(un) SP%rho(1:SP%num_total)
        write(un) SP%pre(1:SP%num_total)
        write(un) SP%tem(1:SP%num_total)
        close(un)
        open(newunit=un, file=trim(setup_file) // ".param", status="replace")
        write(un, '(a, 1x, es26.17e3)') "dx", param%dx
        write(un, '(a, 1x, es26.17e3)') "h", param%h
        write(un, '(a, 1x, es26.17e3)') "h_eff", param%h_eff
        write(un, '(a, 1x, es26.17e3)') "W_ave", param%W_ave
        write(un, '(a, 1x, es26.17e3)') "dt", param%dt
        write(un, '(a, 1x, es26.17e3)') "WL_thick", param%WL_thick
        write(un, '(a, 1x, i0)') "num_WL", param%num_WL
        write(un, '(a, 1x, i0)') "num_int", SP%num_int
        write(un, '(a, 1x, i0)') "num_ext", SP%num_ext
        close(un)
        stop
    endif

    call init_virtual_markers(param, SP, VM)
    call init_cell(param, SP, VM, cell)
    call update_cell(param, SP, cell)

    if (trim(adjustl(param%read_name)) == "new") then
        call write_param(param, SP, VM, cell)
        call write_data(0, param, SP)
        loop_start = param%start_step
    else
        call write_param_dirs()
        loop_start = param%start_step + 1
    endif
    call write_init_info(param, SP)

    log_interval = max(param%total_step / 100, 1)
    call system_clock(t_start, rate)

    do step = loop_start, param%end_step
#if (USE_GRAPH == 1)
        call lssph_check(lssph_step(dev, 1_c_int32_t), "lssph_step")
#else
#if (RK == 2)
        call RK2(param, SP, VM, cell)
#else
        call RK4(param, SP, VM, cell)
#endif
#if (TARGET_PROBLEM != 1)
        call PST(param, SP, VM, cell)
#endif
#endif

#if (TARGET_PROBLEM == 1)
        call check_steady_state(is_steady, param, SP)
        if (is_steady == 1) then
            write(*, *) "  Reached a steady state at step: ", step
            call write_data(step, param, SP)
            call system_clock(t_now)
            call write_progress(step, param, dble(t_now - t_start) / dble(rate))
            exit
        endif
#endif

        if (mod(step, param%write_step) == 0) call write_data(step, param, SP)

        if (mod(step, log_interval) == 0 .or. step == param%end_step) then
            call lssph_check(lssph_synchronize(dev), "lssph_synchronize")
            call system_clock(t_now)
            call write_progress(step, param, dble(t_now - t_start) / dble(rate))
        endif
    enddo

    call lssph_check(lssph_destroy(dev), "lssph_destroy")
    is_steady = 0

contains

    ! A restart writes into results/SAVE_NAME/data as well.
    subroutine write_param_dirs()
        call execute_command_line("mkdir -p results/" // trim(adjustl(param%save_name)) // "/data")
        call execute_command_line("mkdir -p results/" // trim(adjustl(param%save_name)) // "/config")
    end subroutine write_param_dirs

end program main
