! host_types.f90 -- derived types, kernel function, parameters and checks of
! the config.h-driven host.  Module names and interfaces follow the
! reference (source/core/global_types.f90, kernel_functions_mod.f90,
! source/setup/init_param_mod.f90, source/check/check_param_mod.f90) so the
! driver reads like the reference's main program.

#include "config.h"
#include "lssph_defaults.h"

module global_types
    implicit none

    ! Param_type of global_types.f90:22-73 plus the options of this host.
    type :: Param_type
        integer            :: target_problem
        double precision   :: u_top
        double precision   :: tem_top, tem_bottom, tem_left, tem_right
        character(len=256) :: save_name
        character(len=256) :: read_name
        integer            :: omp_threads
        double precision   :: len_x, len_y
        integer            :: num_x, num_y
        integer            :: start_step, end_step, write_step
        double precision   :: threshold
        double precision   :: coe_CFL, coe_dif
        double precision   :: zeta_RSST, xi_VIM
        double precision   :: rho_ref, vis_ref, K_ref
        double precision   :: k_th_ref, cp_ref, alpha_ref
        double precision   :: gravity
        double precision   :: pos_pert
        integer            :: tg_a, tg_b
        double precision   :: coe_h
        double precision   :: PST_c
        double precision   :: coe_delta_sph
        integer            :: kernel_type, sph_model, wall_model
        integer            :: total_step, num_WL
        double precision   :: dx, SP_mass, h, h_eff, W_ave, WL_thick
        double precision   :: delta_tem, tem_ave
        double precision   :: sound_speed, kinematic_vis, thermal_dif
        double precision   :: Re, Ra, Pr
        double precision   :: dt, dt_CFL, dt_vis, dt_th
        double precision   :: dt_CFL_relax, dt_vis_relax
        ! options of this host
        integer            :: rk, periodic, precision, cell_split
        double precision   :: diff_coef, diff_source, cond_max, cell_skin
    end type Param_type

    ! SP_type of global_types.f90:78-103: host mirror of the device state.
    type :: SP_type
        integer                       :: num_total, num_int, num_ext
        integer         , allocatable :: ptype(:)
        double precision, allocatable :: x(:), y(:)
        double precision, allocatable :: u(:), v(:)
        double precision, allocatable :: rho(:), pre(:), tem(:)
        double precision, allocatable :: RHS_u(:), RHS_v(:)
        double precision, allocatable :: RHS_rho(:), RHS_tem(:)
    end type SP_type

    ! VM_type of global_types.f90:108-114 (host copy for VM.dat)
    type :: VM_type
        integer         , allocatable :: ptype(:)
        double precision, allocatable :: x(:), y(:)
        double precision, allocatable :: u(:), v(:)
        double precision, allocatable :: pre(:), tem(:)
    end type VM_type

    ! Cell_type: the cell list lives on the device; the host keeps idx_SP
    ! for cell_idx.dat.
    type :: Cell_type
        integer, allocatable :: idx_SP(:)
    end type Cell_type

end module global_types


module kernel_functions_mod
    implicit none
    double precision, parameter, private :: pi = acos(-1.0d0)
contains
    ! W(r, h) in 2-D for KERNEL_TYPE 1..5 (kernel_functions_mod.f90:26-81)
    pure subroutine cal_W(r, h, W)
        double precision, intent(in)  :: r, h
        double precision, intent(out) :: W
        double precision :: q, t
        q = r / h
#if (KERNEL_TYPE == 1)
        if (q <= 1.0d0) then
            W = (2.0d0 - q)**3 - 4.0d0 * (1.0d0 - q)**3
        elseif (q <= 2.0d0) then
            W = (2.0d0 - q)**3
        else
            W = 0.0d0
        endif
        W = 5.0d0 / (14.0d0 * pi * (h * h)) * W
#elif (KERNEL_TYPE == 2)
        if (q <= 1.0d0) then
            W = (3.0d0 - q)**5 - 6.0d0 * (2.0d0 - q)**5 + 15.0d0 * (1.0d0 - q)**5
        elseif (q <= 2.0d0) then
            W = (3.0d0 - q)**5 - 6.0d0 * (2.0d0 - q)**5
        elseif (q <= 3.0d0) then
            W = (3.0d0 - q)**5
        else
            W = 0.0d0
        endif
        W = 7.0d0 / (478.0d0 * pi * (h * h)) * W
#elif (KERNEL_TYPE == 3)
        t = 1.0d0 - 0.5d0 * q
        W = 0.0d0
        if (q <= 2.0d0) W = t**4 * (1.0d0 + 2.0d0 * q)
        W = 7.0d0 / (4.0d0 * pi * (h * h)) * W
#elif (KERNEL_TYPE == 4)
        t = 1.0d0 - 0.5d0 * q
        W = 0.0d0
        if (q <= 2.0d0) W = t**6 * (1.0d0 + 3.0d0 * q + 35.0d0 * (q * q) / 12.0d0)
        W = 9.0d0 / (4.0d0 * pi * (h * h)) * W
#else
        t = 1.0d0 - 0.5d0 * q
        W = 0.0d0
        if (q <= 2.0d0) W = t**8 * (1.0d0 + 4.0d0 * q + 25.0d0 * (q * q) / 4.0d0 + 4.0d0 * (q * q * q))
        W = 39.0d0 / (14.0d0 * pi * (h * h)) * W
#endif
    end subroutine cal_W
end module kernel_functions_mod


module init_param_mod
    use global_types        , only: Param_type
    use kernel_functions_mod, only: cal_W
    implicit none
contains
    ! config.h macros -> param, with the derivations of init_param_mod.f90:92-199
    subroutine init_param(param)
        type(Param_type), intent(out) :: param

        param%target_problem = TARGET_PROBLEM
        param%u_top          = U_TOP
        param%tem_top        = TEM_TOP
        param%tem_bottom     = TEM_BOTTOM
        param%tem_left       = TEM_LEFT
        param%tem_right      = TEM_RIGHT
        param%save_name      = SAVE_NAME
        param%read_name      = READ_NAME
        param%omp_threads    = OMP_THREADS
        param%len_x          = LEN_X
        param%len_y          = LEN_Y
        param%num_x          = NUM_X
        param%num_y          = NUM_Y
        param%start_step     = START_STEP
        param%end_step       = END_STEP
        param%write_step     = WRITE_STEP
        param%threshold      = THRESHOLD
        param%coe_CFL        = COE_CFL
        param%coe_dif        = COE_DIF
        param%zeta_RSST      = ZETA_RSST
        param%xi_VIM         = XI_VIM
        param%rho_ref        = RHO_REF
        param%vis_ref        = VIS_REF
        param%K_ref          = K_REF
        param%k_th_ref       = K_TH_REF
        param%cp_ref         = CP_REF
        param%alpha_ref      = ALPHA_REF
        param%gravity        = GRAVITY
        param%pos_pert       = POS_PERT
        param%tg_a           = TG_A
        param%tg_b           = TG_B
        param%coe_h          = COE_H
        param%PST_c          = PST_C
        param%coe_delta_sph  = COE_DELTA_SPH
        param%kernel_type    = KERNEL_TYPE
        param%sph_model      = SPH_MODEL
        param%wall_model     = WALL_MODEL
        param%rk             = RK
        param%periodic       = PERIODIC
        param%precision      = PRECISION
        param%diff_coef      = DIFF_COEF
        param%diff_source    = DIFF_SOURCE
        param%cond_max       = COND_MAX
        param%cell_skin      = CELL_SKIN
        param%cell_split     = CELL_SPLIT

        if (trim(adjustl(param%read_name)) == "new") then
            param%total_step = param%end_step
        else
            param%total_step = param%end_step - param%start_step
        endif
        param%dx      = param%len_x / dble(param%num_x)
        param%SP_mass = param%rho_ref * (param%dx * param%dx)
        param%h       = param%coe_h * param%dx
#if (KERNEL_TYPE == 2)
        param%h_eff   = 3.0d0 * param%h
#else
        param%h_eff   = 2.0d0 * param%h
#endif
        call cal_W(param%dx, param%h, param%W_ave)
        if (param%periodic == 1) then
            param%num_WL = 0
        else
            param%num_WL = ceiling(param%h_eff / param%dx)
        endif
        param%WL_thick = dble(param%num_WL) * param%dx
#if (TARGET_PROBLEM == 4)
        param%delta_tem = param%tem_bottom - param%tem_top
        param%tem_ave   = 0.5d0 * (param%tem_bottom + param%tem_top)
#else
        param%delta_tem = 0.0d0
        param%tem_ave   = 0.0d0
#endif
        param%sound_speed = sqrt(param%K_ref / param%rho_ref)
#if (TARGET_PROBLEM == 1)
        param%kinematic_vis = 1.0d0
#else
        param%kinematic_vis = param%vis_ref / param%rho_ref
#endif
        param%thermal_dif = param%k_th_ref / (param%rho_ref * param%cp_ref)
#if (TARGET_PROBLEM == 2)
        param%Re = 1.0d0 * param%len_x / param%kinematic_vis
#else
        param%Re = param%u_top * param%len_x / param%kinematic_vis
#endif
        param%Ra = param%alpha_ref * param%delta_tem * param%gravity * (param%len_y * param%len_y * param%len_y) &
                   / (param%kinematic_vis * param%thermal_dif)
        param%Pr = param%kinematic_vis / param%thermal_dif
        param%dt_CFL = param%coe_CFL * (param%dx / param%sound_speed)
        param%dt_vis = param%coe_dif * ((param%dx * param%dx) / param%kinematic_vis)
        param%dt_th  = param%coe_dif * ((param%dx * param%dx) / param%thermal_dif)
        param%dt_CFL_relax = param%dt_CFL * param%zeta_RSST * param%xi_VIM
        param%dt_vis_relax = param%dt_vis * (param%xi_VIM * param%xi_VIM)
#if (TARGET_PROBLEM == 1)
        param%dt = param%dt_vis
#elif (TARGET_PROBLEM == 4)
        param%dt = min(param%dt_CFL_relax, param%dt_vis_relax, param%dt_th)
#else
        param%dt = min(param%dt_CFL_relax, param%dt_vis_relax)
#endif
        if (DT > 0.0d0) param%dt = DT
    end subroutine init_param
end module init_param_mod


module check_param_mod
    use global_types, only: Param_type
    implicit none
contains
    subroutine fatal(msg)
        character(len=*), intent(in) :: msg
        write(*,*) "+ ======================================================== +"
        write(*,*) "|   Fatal Error in setup:                                  |"
        write(*,*) "|      - ", msg
        write(*,*) "|      - Please check 'config.h'."
        write(*,*) "+ ======================================================== +"
        error stop
    end subroutine fatal

    ! The conditions of check_param_mod.f90:22-284, plus the options of the
    ! GPU path (RK 2/4, PRECISION 0/1).
    subroutine check_param(param)
        type(Param_type), intent(in) :: param
        if (param%target_problem < 1 .or. param%target_problem > 4) call fatal("TARGET_PROBLEM must be between 1 and 4.")
        if (param%kernel_type < 1 .or. param%kernel_type > 5) call fatal("KERNEL_TYPE must be between 1 and 5.")
        if (param%sph_model < 1 .or. param%sph_model > 6) call fatal("SPH_MODEL must be between 1 and 6.")
        if (param%wall_model < 1 .or. param%wall_model > 3) call fatal("WALL_MODEL must be between 1 and 3.")
        if (param%rk /= 2 .and. param%rk /= 4) call fatal("RK must be 2 or 4.")
        if (param%precision /= 0 .and. param%precision /= 1) call fatal("PRECISION must be 0 or 1.")
#if (PERIODIC == 0)
#if (TARGET_PROBLEM == 1)
#if (U_BOUNDARY_TOP != 1 && U_BOUNDARY_TOP != 2) || (U_BOUNDARY_BOTTOM != 1 && U_BOUNDARY_BOTTOM != 2) || \
    (U_BOUNDARY_LEFT != 1 && U_BOUNDARY_LEFT != 2) || (U_BOUNDARY_RIGHT != 1 && U_BOUNDARY_RIGHT != 2)
        call fatal("Diffusion Equation Test: U_BOUNDARY_* must be 1 or 2.")
#endif
        if (param%len_x /= 1.0d0 .or. param%len_y /= 1.0d0) &
            call fatal("For Diffusion Equation Test, LEN_X and LEN_Y must be exactly 1.0.")
#endif
#if (TARGET_PROBLEM == 2)
#if (U_BOUNDARY_TOP != 1) || (U_BOUNDARY_BOTTOM != 1) || (U_BOUNDARY_LEFT != 1) || (U_BOUNDARY_RIGHT != 1)
        call fatal("Taylor-Green Vortex: U_BOUNDARY_* must be 1.")
#endif
#endif
#if (TARGET_PROBLEM == 3)
#if (U_BOUNDARY_TOP != 3) || (U_BOUNDARY_BOTTOM != 2) || (U_BOUNDARY_LEFT != 2) || (U_BOUNDARY_RIGHT != 2)
        call fatal("Lid-driven Cavity Flow: U_BOUNDARY_TOP must be 3, all other U_BOUNDARY_* 2.")
#endif
#endif
#if (TARGET_PROBLEM == 4)
#if (TEM_BOUNDARY_TOP != 1) || (TEM_BOUNDARY_BOTTOM != 1) || (TEM_BOUNDARY_LEFT != 2) || (TEM_BOUNDARY_RIGHT != 2)
        call fatal("Boussinesq Convection: TEM_BOUNDARY_TOP/BOTTOM must be 1, LEFT/RIGHT 2.")
#endif
#endif
#endif
        if (param%target_problem == 2 .and. (param%tg_a <= 0 .or. param%tg_b <= 0)) &
            call fatal("TG_A and TG_B must be positive.")
        if (param%len_x <= 0.0d0 .or. param%len_y <= 0.0d0) call fatal("LEN_X and LEN_Y must be positive.")
        if (param%num_x <= 0 .or. param%num_y <= 0) call fatal("NUM_X and NUM_Y must be positive.")
        if (param%start_step <= 0 .or. param%end_step <= 0 .or. param%write_step <= 0) &
            call fatal("START_STEP, END_STEP and WRITE_STEP must be positive.")
        if (param%start_step >= param%end_step) call fatal("START_STEP must be less than END_STEP.")
        if (param%coe_CFL <= 0.0d0 .or. param%coe_dif <= 0.0d0 .or. param%coe_CFL >= 1.0d0 .or. &
            param%coe_dif >= 1.0d0) call fatal("COE_CFL and COE_DIF must be between 0 and 1.")
        if (param%zeta_RSST < 1.0d0 .or. param%xi_VIM < 1.0d0) call fatal("ZETA_RSST and XI_VIM must be 1.0 or greater.")
        if (param%rho_ref <= 0.0d0 .or. param%vis_ref <= 0.0d0 .or. param%K_ref <= 0.0d0 .or. &
            param%k_th_ref <= 0.0d0 .or. param%cp_ref <= 0.0d0 .or. param%alpha_ref <= 0.0d0) &
            call fatal("Physical properties must be positive.")
        if (param%gravity < 0.0d0) call fatal("GRAVITY must be non-negative.")
        if (param%target_problem == 1 .and. (param%pos_pert < 0.0d0 .or. param%pos_pert >= 0.5d0)) &
            call fatal("POS_PERT must be between 0.0 and 0.5")
        if (param%coe_h <= 0.0d0) call fatal("COE_H must be positive.")
        if (param%PST_c < 0.0d0 .or. param%coe_delta_sph < 0.0d0) call fatal("PST_C and COE_DELTA_SPH must be non-negative.")
        if (abs(param%len_x / dble(param%num_x) - param%len_y / dble(param%num_y)) > 1.0d-8) &
            call fatal("Particle spacing must be isotropic (dx == dy).")
    end subroutine check_param
end module check_param_mod
