! host_io.f90 -- output in the reference's formats, so its Python analysis
! (analysis/common/read_utils.py:22-107) reads the results unchanged:
!   results/SAVE_NAME/data/<step>.dat   write_data_mod.f90:43-52
!   results/SAVE_NAME/config/parameters.csv, data/cell_idx.dat, data/VM.dat
!                                       write_param_mod.f90:39-161
! plus terminal progress lines.

#include "config.h"
#include "lssph_defaults.h"

module write_data_mod
    use global_types, only: Param_type, SP_type
    use device_state_mod, only: download_state
    implicit none
contains
    ! Stream-binary snapshot: ptype(i4), x, y, u, v, rho, pre [, tem] (f8).
    subroutine write_data(step, param, SP)
        integer         , intent(in)    :: step
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        character(len=1024) :: step_str, file_name
        integer :: un
        call download_state(SP)
        write(step_str, *) step
        file_name = "results/" // trim(adjustl(param%save_name)) // "/data/" // trim(adjustl(step_str)) // ".dat"
        open(newunit=un, file=trim(file_name), status="replace", form="unformatted", access="stream")
        write(un) SP%ptype(1:SP%num_total)
        write(un) SP%x(1:SP%num_total)
        write(un) SP%y(1:SP%num_total)
        write(un) SP%u(1:SP%num_total)
        write(un) SP%v(1:SP%num_total)
        write(un) SP%rho(1:SP%num_total)
        write(un) SP%pre(1:SP%num_total)
#if (TARGET_PROBLEM == 4)
        write(un) SP%tem(1:SP%num_total)
#endif
        close(un)
    end subroutine write_data
end module write_data_mod


module write_param_mod
    use iso_c_binding
    use lssph_binding
    use global_types, only: Param_type, SP_type, VM_type, Cell_type
    use ghost_mod, only: ghost
    implicit none
contains
    subroutine mkdir(path)
        character(len=*), intent(in) :: path
        call execute_command_line("mkdir -p " // trim(adjustl(path)))
    end subroutine mkdir

    subroutine copy_config(config_dir)
        character(len=*), intent(in) :: config_dir
        logical :: present
        inquire(file="config.h", exist=present)
        if (present) call execute_command_line("cp -f config.h " // trim(adjustl(config_dir)) // "/")
    end subroutine copy_config

    ! Directories, parameters.csv, cell_idx.dat, the initial ghost fill and
    ! VM.dat (write_param_mod.f90:39-161).
    subroutine write_param(param, SP, VM, cell)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        type(VM_type)   , intent(inout) :: VM
        type(Cell_type) , intent(inout) :: cell
        character(len=1024) :: base_dir, config_dir, data_dir
        integer :: un
        base_dir = "results/" // trim(adjustl(param%save_name))
        config_dir = trim(base_dir) // "/config"
        data_dir = trim(base_dir) // "/data"
        call mkdir(base_dir)
        call mkdir(config_dir)
        call mkdir(data_dir)
        ! the run's config.h beside its parameters (write_param_mod.f90:50-53,
        ! which copies it from the working directory)
        call copy_config(config_dir)

        open(newunit=un, file=trim(config_dir) // "/parameters.csv", status="replace", form="formatted")
        write(un, "(a)") "Parameter, Value"
        write(un, "(a, ',', i0    )") "target_problem", param%target_problem
        write(un, "(a, ',', e21.14)") "u_top"         , param%u_top
        write(un, "(a, ',', e21.14)") "tem_top"       , param%tem_top
        write(un, "(a, ',', e21.14)") "tem_bottom"    , param%tem_bottom
        write(un, "(a, ',', e21.14)") "tem_left"      , param%tem_left
        write(un, "(a, ',', e21.14)") "tem_right"     , param%tem_right
        write(un, "(a, ',', e21.14)") "len_x"         , param%len_x
        write(un, "(a, ',', e21.14)") "len_y"         , param%len_y
        write(un, "(a, ',', i0    )") "num_x"         , param%num_x
        write(un, "(a, ',', i0    )") "num_y"         , param%num_y
        write(un, "(a, ',', i0    )") "start_step"    , param%start_step
        write(un, "(a, ',', i0    )") "end_step"      , param%end_step
        write(un, "(a, ',', i0    )") "write_step"    , param%write_step
        write(un, "(a, ',', e21.14)") "threshold"     , param%threshold
        write(un, "(a, ',', e21.14)") "coe_CFL"       , param%coe_CFL
        write(un, "(a, ',', e21.14)") "coe_dif"       , param%coe_dif
        write(un, "(a, ',', e21.14)") "zeta_RSST"     , param%zeta_RSST
        write(un, "(a, ',', e21.14)") "xi_VIM"        , param%xi_VIM
        write(un, "(a, ',', e21.14)") "rho_ref"       , param%rho_ref
        write(un, "(a, ',', e21.14)") "vis_ref"       , param%vis_ref
        write(un, "(a, ',', e21.14)") "K_ref"         , param%K_ref
        write(un, "(a, ',', e21.14)") "k_th_ref"      , param%k_th_ref
        write(un, "(a, ',', e21.14)") "cp_ref"        , param%cp_ref
        write(un, "(a, ',', e21.14)") "alpha_ref"     , param%alpha_ref
        write(un, "(a, ',', e21.14)") "gravity"       , param%gravity
        write(un, "(a, ',', e21.14)") "pos_pert"      , param%pos_pert
        write(un, "(a, ',', i0    )") "tg_a"          , param%tg_a
        write(un, "(a, ',', i0    )") "tg_b"          , param%tg_b
        write(un, "(a, ',', e21.14)") "coe_h"         , param%coe_h
        write(un, "(a, ',', e21.14)") "PST_c"         , param%PST_c
        write(un, "(a, ',', e21.14)") "coe_delta_sph" , param%coe_delta_sph
        write(un, "(a, ',', i0    )") "kernel_type"   , param%kernel_type
        write(un, "(a, ',', i0    )") "sph_model"     , param%sph_model
        write(un, "(a, ',', i0    )") "wall_model"    , param%wall_model
        write(un, "(a, ',', i0    )") "total_step"    , param%total_step
        write(un, "(a, ',', e21.14)") "dx"            , param%dx
        write(un, "(a, ',', e21.14)") "SP_mass"       , param%SP_mass
        write(un, "(a, ',', e21.14)") "h"             , param%h
        write(un, "(a, ',', e21.14)") "h_eff"         , param%h_eff
        write(un, "(a, ',', e21.14)") "W_ave"         , param%W_ave
        write(un, "(a, ',', i0    )") "num_WL"        , param%num_WL
        write(un, "(a, ',', e21.14)") "WL_thick"      , param%WL_thick
        write(un, "(a, ',', e21.14)") "delta_tem"     , param%delta_tem
        write(un, "(a, ',', e21.14)") "tem_ave"       , param%tem_ave
        write(un, "(a, ',', e21.14)") "sound_speed"   , param%sound_speed
        write(un, "(a, ',', e21.14)") "kinematic_vis" , param%kinematic_vis
        write(un, "(a, ',', e21.14)") "thermal_dif"   , param%thermal_dif
        write(un, "(a, ',', e21.14)") "Re"            , param%Re
        write(un, "(a, ',', e21.14)") "Ra"            , param%Ra
        write(un, "(a, ',', e21.14)") "Pr"            , param%Pr
        write(un, "(a, ',', e21.14)") "dt_CFL"        , param%dt_CFL
        write(un, "(a, ',', e21.14)") "dt_vis"        , param%dt_vis
        write(un, "(a, ',', e21.14)") "dt_th"         , param%dt_th
        write(un, "(a, ',', e21.14)") "dt_CFL_relax"  , param%dt_CFL_relax
        write(un, "(a, ',', e21.14)") "dt_vis_relax"  , param%dt_vis_relax
        write(un, "(a, ',', e21.14)") "dt"            , param%dt
        write(un, "(a, ',', i0    )") "num_total"     , SP%num_total
        write(un, "(a, ',', i0    )") "num_int"       , SP%num_int
        write(un, "(a, ',', i0    )") "num_ext"       , SP%num_ext
        close(un)

        call lssph_check(lssph_download_cell_index(dev, cell%idx_SP), "lssph_download_cell_index")
        open(newunit=un, file=trim(data_dir) // "/cell_idx.dat", status="replace", form="unformatted", access="stream")
        write(un) cell%idx_SP(1:SP%num_total)
        close(un)

        if (param%periodic == 0) then
            call ghost(param, SP, VM, cell)
            call lssph_check(lssph_download_vm(dev, VM%ptype, VM%x, VM%y, VM%u, VM%v, VM%pre, VM%tem), &
                             "lssph_download_vm")
            open(newunit=un, file=trim(data_dir) // "/VM.dat", status="replace", form="unformatted", access="stream")
            write(un) VM%ptype(1:SP%num_ext)
            write(un) VM%x(1:SP%num_ext)
            write(un) VM%y(1:SP%num_ext)
            write(un) VM%u(1:SP%num_ext)
            write(un) VM%v(1:SP%num_ext)
            write(un) VM%pre(1:SP%num_ext)
#if (TARGET_PROBLEM == 4)
            write(un) VM%tem(1:SP%num_ext)
#endif
            close(un)
        endif
    end subroutine write_param
end module write_param_mod


module write_progress_mod
    use global_types, only: Param_type, SP_type
    implicit none
contains
    subroutine write_init_info(param, SP)
        type(Param_type), intent(in) :: param
        type(SP_type)   , intent(in) :: SP
        write(*, '(a)') " LS-SPH on MI355X (gfx950)"
        write(*, '(a, i0, a, i0, a, i0)') "   particles: total ", SP%num_total, ", fluid ", SP%num_int, &
                                          ", walls ", SP%num_ext
        write(*, '(a, i0, a, i0, a, i0, a, i0)') "   problem ", param%target_problem, ", kernel ", &
            param%kernel_type, ", wall model ", param%wall_model, ", RK", param%rk
        write(*, '(a, es12.5, a, es12.5, a, i0)') "   dt ", param%dt, ", h_eff ", param%h_eff, &
            ", precision ", param%precision
    end subroutine write_init_info

    ! Progress bar on the terminal and appended to results/<SAVE_NAME>/config/
    ! run.log, in the reference's format (write_progress_mod.f90:22-69): the
    ! share of the run's steps done (counted from the restart step for a
    ! restart) and the wall time since the loop began.
    subroutine write_progress(step, param, seconds)
        integer         , intent(in) :: step
        type(Param_type), intent(in) :: param
        double precision, intent(in) :: seconds
        character(len=20)   :: bar
        character(len=1024) :: log_file
        character(len=96)   :: line
        integer :: done, filled, un
        double precision :: frac
        done = step
        if (trim(adjustl(param%read_name)) /= "new") done = step - param%start_step
        frac = dble(done) / dble(max(param%total_step, 1))
        filled = max(0, min(20, int(frac * 20.0d0)))
        bar = repeat("=", filled) // repeat(" ", 20 - filled)
        write(line, "(a, a, a, f6.2, a, es11.4, a)") "   Progress: [", bar, "] ", frac * 100.0d0, &
            "%  |  Elapsed Time: ", seconds, " [s]"
        write(*, "(a)") trim(line)
        log_file = "results/" // trim(adjustl(param%save_name)) // "/config/run.log"
        open(newunit=un, file=trim(log_file), status="unknown", position="append", form="formatted")
        write(un, "(a)") trim(line)
        close(un)
    end subroutine write_progress
end module write_progress_mod
