! host_solver.f90 -- the hot-path subroutines of the reference, with their
! module names and signatures, as thin wrappers over the GPU library:
!   init_virtual_markers  (init_virtual_markers_mod.f90:23-141)
!   init_cell, update_cell (init_cell_mod.f90:23-184, update_cell_mod.f90:21-97)
!   ghost                 (ghost_mod.f90:43-78)
!   calculate_RHS         (calculate_RHS_mod.f90:29-52)
!   RK2, RK4              (RK2_mod.f90:30-128, RK4_mod.f90:30-179)
!   PST                   (PST_mod.f90:28-116)
!   check_steady_state    (check_steady_state_mod.f90:23-69)
! Particle state lives on the device after init_cell; the host arrays of SP
! are refreshed by download_state (write_data does it before writing).

#include "config.h"
#include "lssph_defaults.h"

module device_state_mod
    use iso_c_binding
    use lssph_binding
    use global_types, only: Param_type, SP_type
    implicit none
contains
    ! Param_type -> lssph_params_t (the numeric fields of the C struct)
    function to_c_params(param) result(p)
        type(Param_type), intent(in) :: param
        type(lssph_params_t) :: p
        p%len_x = param%len_x
        p%len_y = param%len_y
        p%dx = param%dx
        p%sp_mass = param%SP_mass
        p%h = param%h
        p%h_eff = param%h_eff
        p%w_ave = param%W_ave
        p%wl_thick = param%WL_thick
        p%rho_ref = param%rho_ref
        p%vis_ref = param%vis_ref
        p%sound_speed = param%sound_speed
        p%thermal_dif = param%thermal_dif
        p%alpha_ref = param%alpha_ref
        p%gravity = param%gravity
        p%tem_ave = param%tem_ave
        p%tem_top = param%tem_top
        p%tem_bottom = param%tem_bottom
        p%tem_left = param%tem_left
        p%tem_right = param%tem_right
        p%u_top = param%u_top
        p%zeta_rsst = param%zeta_RSST
        p%xi_vim = param%xi_VIM
        p%coe_delta_sph = param%coe_delta_sph
        p%pst_c = param%PST_c
        p%dt = param%dt
        p%threshold = param%threshold
        p%diff_coef = param%diff_coef
        p%diff_source = param%diff_source
        p%cond_max = param%cond_max
        p%target_problem = param%target_problem
        p%kernel_type = param%kernel_type
        p%sph_model = param%sph_model
        p%wall_model = param%wall_model
        p%rk = param%rk
        p%u_bc_top = U_BOUNDARY_TOP
        p%u_bc_bottom = U_BOUNDARY_BOTTOM
        p%u_bc_left = U_BOUNDARY_LEFT
        p%u_bc_right = U_BOUNDARY_RIGHT
        p%tem_bc_top = TEM_BOUNDARY_TOP
        p%tem_bc_bottom = TEM_BOUNDARY_BOTTOM
        p%tem_bc_left = TEM_BOUNDARY_LEFT
        p%tem_bc_right = TEM_BOUNDARY_RIGHT
        p%num_x = param%num_x
        p%num_y = param%num_y
        p%num_wl = param%num_WL
        p%periodic = param%periodic
        p%precision = param%precision
    end function to_c_params

    ! device -> host copy of SP (x, y, u, v, rho, pre, tem)
    subroutine download_state(SP)
        type(SP_type), intent(inout) :: SP
        call lssph_check(lssph_download(dev, SP%ptype, SP%x, SP%y, SP%u, SP%v, SP%rho, SP%pre, SP%tem), &
                         "lssph_download")
    end subroutine download_state
end module device_state_mod


module init_virtual_markers_mod
    use global_types, only: Param_type, SP_type, VM_type
    implicit none
contains
    ! Host copy of the markers (for VM.dat); the device derives the marker
    ! positions and normals itself from the wall particles when the state is
    ! uploaded.
    subroutine init_virtual_markers(param, SP, VM)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(in)    :: SP
        type(VM_type)   , intent(inout) :: VM
        integer :: m
        m = max(SP%num_ext, 1)
        allocate(VM%ptype(m), VM%x(m), VM%y(m), VM%u(m), VM%v(m), VM%pre(m), VM%tem(m))
        VM%ptype = 0
        VM%x = 0.0d0
        VM%y = 0.0d0
        VM%u = 0.0d0
        VM%v = 0.0d0
        VM%pre = 0.0d0
        VM%tem = 0.0d0
    end subroutine init_virtual_markers
end module init_virtual_markers_mod


module init_cell_mod
    use iso_c_binding
    use lssph_binding
    use global_types, only: Param_type, SP_type, VM_type, Cell_type
    use device_state_mod, only: to_c_params
    implicit none
contains
    ! Creates the device handle, uploads SP (which also places the virtual
    ! markers) and builds the cell list.
    subroutine init_cell(param, SP, VM, cell)
        type(Param_type), intent(in)  :: param
        type(SP_type)   , intent(in)  :: SP
        type(VM_type)   , intent(in)  :: VM
        type(Cell_type) , intent(out) :: cell
        type(lssph_params_t) :: p(1)
        integer(c_int32_t) :: n_int(1), n_ext(1)
        p(1) = to_c_params(param)
        n_int(1) = SP%num_int
        n_ext(1) = SP%num_ext
        if (c_associated(dev)) call lssph_check(lssph_destroy(dev), "lssph_destroy")
        call lssph_check(lssph_create(p, 1_c_int32_t, n_int, n_ext, dev), "lssph_create")
        call lssph_check(lssph_set_skin(dev, param%cell_skin), "lssph_set_skin")
        call lssph_check(lssph_set_cell_split(dev, int(param%cell_split, c_int32_t)), "lssph_set_cell_split")
        call lssph_check(lssph_upload(dev, SP%ptype, SP%x, SP%y, SP%u, SP%v, SP%rho, SP%pre, SP%tem), &
                         "lssph_upload")
        call lssph_check(lssph_init_cell(dev), "lssph_init_cell")
        allocate(cell%idx_SP(SP%num_total))
        cell%idx_SP = 0
    end subroutine init_cell
end module init_cell_mod


module update_cell_mod
    use lssph_binding
    use global_types, only: Param_type, SP_type, Cell_type
    implicit none
contains
    subroutine update_cell(param, SP, cell)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(in)    :: SP
        type(Cell_type) , intent(inout) :: cell
        call lssph_check(lssph_update_cell(dev), "lssph_update_cell")
    end subroutine update_cell
end module update_cell_mod


module ghost_mod
    use lssph_binding
    use global_types, only: Param_type, SP_type, VM_type, Cell_type
    implicit none
contains
    subroutine ghost(param, SP, VM, cell)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        type(VM_type)   , intent(inout) :: VM
        type(Cell_type) , intent(in)    :: cell
        call lssph_check(lssph_ghost(dev), "lssph_ghost")
    end subroutine ghost
end module ghost_mod


module calculate_RHS_mod
    use lssph_binding
    use global_types, only: Param_type, SP_type, Cell_type
    implicit none
contains
    ! RHS of the fluid particles into SP%RHS_* (downloaded).
    subroutine calculate_RHS(param, SP, cell)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        type(Cell_type) , intent(in)    :: cell
        call lssph_check(lssph_calculate_rhs(dev), "lssph_calculate_rhs")
        call lssph_check(lssph_download_rhs(dev, SP%RHS_u, SP%RHS_v, SP%RHS_rho, SP%RHS_tem), "lssph_download_rhs")
    end subroutine calculate_RHS
end module calculate_RHS_mod


module RK2_mod
    use lssph_binding
    use global_types, only: Param_type, SP_type, VM_type, Cell_type
    implicit none
contains
    subroutine RK2(param, SP, VM, cell)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        type(VM_type)   , intent(inout) :: VM
        type(Cell_type) , intent(inout) :: cell
        call lssph_check(lssph_rk2(dev), "lssph_rk2")
    end subroutine RK2
end module RK2_mod


module RK4_mod
    use lssph_binding
    use global_types, only: Param_type, SP_type, VM_type, Cell_type
    implicit none
contains
    subroutine RK4(param, SP, VM, cell)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        type(VM_type)   , intent(inout) :: VM
        type(Cell_type) , intent(inout) :: cell
        call lssph_check(lssph_rk4(dev), "lssph_rk4")
    end subroutine RK4
end module RK4_mod


module PST_mod
    use lssph_binding
    use global_types, only: Param_type, SP_type, VM_type, Cell_type
    implicit none
contains
    subroutine PST(param, SP, VM, cell)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        type(VM_type)   , intent(inout) :: VM
        type(Cell_type) , intent(inout) :: cell
        call lssph_check(lssph_pst(dev), "lssph_pst")
    end subroutine PST
end module PST_mod


module check_steady_state_mod
    use iso_c_binding
    use lssph_binding
    use global_types, only: Param_type, SP_type
    implicit none
contains
    subroutine check_steady_state(is_steady, param, SP)
        integer         , intent(out)   :: is_steady
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        integer(c_int32_t) :: flag(1)
        real(c_double) :: df(1)
        call lssph_check(lssph_check_steady_state(dev, flag, df), "lssph_check_steady_state")
        is_steady = flag(1)
    end subroutine check_steady_state
end module check_steady_state_mod
