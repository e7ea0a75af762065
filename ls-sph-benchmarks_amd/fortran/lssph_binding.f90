! lssph_binding.f90 -- iso_c_binding interface to liblssph_hip.so
! (include/lssph.h).  The derived type lssph_params_t mirrors the C struct
! field by field; every C entry point has an interface here.  `dev` holds the
! device handle of the run (created by init_cell), so the host modules keep
! the reference's subroutine signatures.
module lssph_binding
    use iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: LSSPH_FP64 = 0, LSSPH_MIXED = 1

    type, bind(C) :: lssph_params_t
        real(c_double) :: len_x, len_y, dx, sp_mass, h, h_eff, w_ave, wl_thick
        real(c_double) :: rho_ref, vis_ref, sound_speed, thermal_dif
        real(c_double) :: alpha_ref, gravity, tem_ave
        real(c_double) :: tem_top, tem_bottom, tem_left, tem_right
        real(c_double) :: u_top
        real(c_double) :: zeta_rsst, xi_vim, coe_delta_sph, pst_c
        real(c_double) :: dt, threshold
        real(c_double) :: diff_coef, diff_source, cond_max
        integer(c_int32_t) :: target_problem, kernel_type, sph_model, wall_model, rk
        integer(c_int32_t) :: u_bc_top, u_bc_bottom, u_bc_left, u_bc_right
        integer(c_int32_t) :: tem_bc_top, tem_bc_bottom, tem_bc_left, tem_bc_right
        integer(c_int32_t) :: num_x, num_y, num_wl
        integer(c_int32_t) :: periodic, precision
    end type lssph_params_t

    type, bind(C) :: lssph_stats_t
        integer(c_int64_t) :: fp64_fallbacks, list_overflows
        integer(c_int32_t) :: max_neighbors, max_cell_count, escaped, pad
    end type lssph_stats_t

    type, bind(C) :: lssph_kernel_times_t
        real(c_double) :: lsa_ms, pst_ms, cells_ms, ghost_ms, other_ms
        integer(c_int64_t) :: lsa_launches, pst_launches, cells_launches, ghost_launches, other_launches
    end type lssph_kernel_times_t

    ! device handle of this run
    type(c_ptr), save :: dev = c_null_ptr

    interface
        integer(c_int32_t) function lssph_create(params, n_inst, n_int, n_ext, out) bind(C, name="lssph_create")
            import :: c_int32_t, c_ptr, lssph_params_t
            type(lssph_params_t), intent(in) :: params(*)
            integer(c_int32_t), value :: n_inst
            integer(c_int32_t), intent(in) :: n_int(*), n_ext(*)
            type(c_ptr), intent(out) :: out
        end function
        integer(c_int32_t) function lssph_destroy(h) bind(C, name="lssph_destroy")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        type(c_ptr) function lssph_last_error(h) bind(C, name="lssph_last_error")
            import :: c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_upload(h, ptype, x, y, u, v, rho, pre, tem) bind(C, name="lssph_upload")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), intent(in) :: ptype(*)
            real(c_double), intent(in) :: x(*), y(*), u(*), v(*), rho(*), pre(*), tem(*)
        end function
        integer(c_int32_t) function lssph_download(h, ptype, x, y, u, v, rho, pre, tem) bind(C, name="lssph_download")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: ptype(*)
            real(c_double), intent(out) :: x(*), y(*), u(*), v(*), rho(*), pre(*), tem(*)
        end function
        integer(c_int32_t) function lssph_init_cell(h) bind(C, name="lssph_init_cell")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_update_cell(h) bind(C, name="lssph_update_cell")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_ghost(h) bind(C, name="lssph_ghost")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_calculate_rhs(h) bind(C, name="lssph_calculate_rhs")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_download_rhs(h, ru, rv, rrho, rtem) bind(C, name="lssph_download_rhs")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: ru(*), rv(*), rrho(*), rtem(*)
        end function
        integer(c_int32_t) function lssph_rk2(h) bind(C, name="lssph_rk2")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_rk4(h) bind(C, name="lssph_rk4")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_pst(h) bind(C, name="lssph_pst")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_step(h, nsteps) bind(C, name="lssph_step")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: nsteps
        end function
        integer(c_int32_t) function lssph_check_steady_state(h, is_steady, df_max) &
                bind(C, name="lssph_check_steady_state")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: is_steady(*)
            real(c_double), intent(out) :: df_max(*)
        end function
        integer(c_int32_t) function lssph_download_vm(h, ptype, x, y, u, v, pre, tem) bind(C, name="lssph_download_vm")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: ptype(*)
            real(c_double), intent(out) :: x(*), y(*), u(*), v(*), pre(*), tem(*)
        end function
        integer(c_int32_t) function lssph_download_cell_index(h, idx_sp) bind(C, name="lssph_download_cell_index")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: idx_sp(*)
        end function
        integer(c_int32_t) function lssph_neighbors(h, counts, list, capacity, total) bind(C, name="lssph_neighbors")
            import :: c_int32_t, c_int64_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: counts(*), list(*)
            integer(c_int64_t), value :: capacity
            integer(c_int64_t), intent(out) :: total
        end function
        integer(c_int32_t) function lssph_stats(h, out) bind(C, name="lssph_stats")
            import :: c_int32_t, c_ptr, lssph_stats_t
            type(c_ptr), value :: h
            type(lssph_stats_t), intent(out) :: out
        end function
        integer(c_int32_t) function lssph_reset_stats(h) bind(C, name="lssph_reset_stats")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_synchronize(h) bind(C, name="lssph_synchronize")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_set_tuning(h, block, list_len, reorder_interval) &
                bind(C, name="lssph_set_tuning")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: block, list_len, reorder_interval
        end function
        integer(c_int32_t) function lssph_set_skin(h, skin) bind(C, name="lssph_set_skin")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: skin
        end function
        integer(c_int32_t) function lssph_set_cell_split(h, columns) bind(C, name="lssph_set_cell_split")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: columns
        end function
        integer(c_int32_t) function lssph_set_scan(h, staged) bind(C, name="lssph_set_scan")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: staged
        end function
        type(c_ptr) function lssph_stream(h) bind(C, name="lssph_stream")
            import :: c_ptr
            type(c_ptr), value :: h
        end function
        integer(c_int32_t) function lssph_profile(h, enable) bind(C, name="lssph_profile")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: enable
        end function
        integer(c_int32_t) function lssph_kernel_times(h, out) bind(C, name="lssph_kernel_times")
            import :: c_int32_t, c_ptr, lssph_kernel_times_t
            type(c_ptr), value :: h
            type(lssph_kernel_times_t), intent(out) :: out
        end function
        integer(c_int32_t) function lssph_pair_lists(h, cap, counts, list, path) bind(C, name="lssph_pair_lists")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: cap
            integer(c_int32_t), intent(out) :: counts(*), list(*), path(*)
        end function
        integer(c_int32_t) function lssph_probe(h, inst, field, q, n, x, y, out) bind(C, name="lssph_probe")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: inst, field, q, n
            real(c_double), intent(in) :: x(*), y(*)
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int32_t) function lssph_probe_static(h, inst, field, q, n, x, y, out) &
                bind(C, name="lssph_probe_static")
            import :: c_int32_t, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), value :: inst, field, q, n
            real(c_double), intent(in) :: x(*), y(*)
            real(c_double), intent(out) :: out(*)
        end function
        integer(c_int32_t) function lssph_set_wave_particles(h, ppw) bind(C, name="lssph_set_wave_particles")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: ppw
        end function
        integer(c_int32_t) function lssph_set_isolation(h, on) bind(C, name="lssph_set_isolation")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), value :: on
        end function
        integer(c_int32_t) function lssph_instance_status(h, retired, reason) bind(C, name="lssph_instance_status")
            import :: c_int32_t, c_ptr
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: retired(*), reason(*)
        end function
    end interface

contains

    ! Stops the run with the library's message when a call failed (the
    ! reference stops on fatal setup errors the same way, `error stop`).
    subroutine lssph_check(rc, where)
        integer(c_int32_t), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer :: n
        if (rc == 0) return
        p = lssph_last_error(dev)
        call c_f_pointer(p, msg, [4096])
        n = 0
        do while (n < 4096)
            if (msg(n + 1) == c_null_char) exit
            n = n + 1
        enddo
        write(*, '(a, a, a, i0, a)') " Fatal Error in ", where, " (code ", rc, "):"
        write(*, *) "  ", msg(1:n)
        error stop
    end subroutine lssph_check

end module lssph_binding
