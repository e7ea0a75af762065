! ref_harness.f90 -- drives the compiled reference (shobuzako-kensuke/
! LS-SPH-Benchmarks, source/*.f90) to write golden vectors and to time its
! CPU loop.  TEST INFRASTRUCTURE ONLY: built by oracle/build_ref.sh against
! the reference's own modules; the reference sources are never copied.
!
! usage: ref_harness <outdir> <nsteps> <dump_every> <threads> [inject.dat]
!        ref_harness <outdir> <nsteps> 0 <threads> [warmup]   (timing mode)
!   Sequence (mirrors source/main.f90:53-166 without its file output):
!   init_param -> init_particle -> [inject state] -> init_virtual_markers
!   -> init_cell -> update_cell -> ghost (the initial fill that
!   write_param_mod.f90:144 performs) -> dumps -> nsteps x (RK, PST or
!   check_steady_state).  dump_every = 0 selects timing mode: no dumps,
!   jittered lattice, `warmup` untimed steps first (default 0), then the
!   wall-clock of nsteps steps printed as "particle_steps_per_s <value>".

#include "config.h"

program ref_harness
    use omp_lib
    use global_types            , only: Param_type, SP_type, VM_type, Cell_type
    use init_param_mod          , only: init_param
    use init_particle_mod       , only: init_particle
    use init_virtual_markers_mod, only: init_virtual_markers
    use init_cell_mod           , only: init_cell
    use update_cell_mod         , only: update_cell
    use ghost_mod               , only: ghost
    use calculate_RHS_mod       , only: calculate_RHS
    use RK2_mod                 , only: RK2
    use RK4_mod                 , only: RK4
#if (TARGET_PROBLEM != 1)
    use PST_mod                 , only: PST
#endif
    use check_steady_state_mod  , only: check_steady_state
    implicit none

    type(Param_type) :: param
    type(SP_type)    :: SP, SP2
    type(VM_type)    :: VM, VM2
    type(Cell_type)  :: cell, cell2
    character(len=1024) :: outdir, arg, inject
    integer :: nsteps, dump_every, nthreads, step, is_steady, un, i, warmup
    double precision :: t0, t1, rnd, df

    call get_command_argument(1, outdir)
    call get_command_argument(2, arg); read(arg, *) nsteps
    call get_command_argument(3, arg); read(arg, *) dump_every
    call get_command_argument(4, arg); read(arg, *) nthreads
    inject = ""
    warmup = 0
    if (command_argument_count() >= 5) then
        call get_command_argument(5, arg)
        if (dump_every == 0) then
            read(arg, *) warmup
        else
            inject = arg
        endif
    endif

    call init_param(param)
    call omp_set_dynamic(.false.)
    call omp_set_num_threads(nthreads)
    call init_particle(param, SP)

    if (dump_every > 0) then
        call write_params(trim(outdir) // "/param.txt")
        call write_state(trim(outdir) // "/setup.dat")
    else
        ! timing mode: seeded +-0.2 dx jitter of the fluid lattice
        call random_seed(put=[(2029 + 17*i, i = 1, 64)])
        do i = 1, SP%num_int
            call random_number(rnd)
            SP%x(i) = SP%x(i) + 0.2d0 * param%dx * (2.0d0*rnd - 1.0d0)
            call random_number(rnd)
            SP%y(i) = SP%y(i) + 0.2d0 * param%dx * (2.0d0*rnd - 1.0d0)
        enddo
    endif

    if (len_trim(inject) > 0) call read_state(trim(inject))

    call init_virtual_markers(param, SP, VM)
    call init_cell(param, SP, VM, cell)
    call update_cell(param, SP, cell)
    call ghost(param, SP, VM, cell)

    if (dump_every > 0) then
        call write_state(trim(outdir) // "/state0.dat")
        call write_vm(trim(outdir) // "/vm0.dat")
        open(newunit=un, file=trim(outdir)//"/cell0.dat", status="replace", &
             form="unformatted", access="stream")
        write(un) cell%idx_SP(1:SP%num_total)
        close(un)
        call calculate_RHS(param, SP, cell)
        call write_rhs(trim(outdir) // "/rhs0.dat")
        call write_neighbors(trim(outdir) // "/nbr.dat")
#if (TARGET_PROBLEM != 1)
        SP2 = SP
        VM2 = VM
        cell2 = cell
        call PST(param, SP2, VM2, cell2)
        call write_state_of(SP2, trim(outdir) // "/pst0.dat")
#endif
    endif

    t0 = omp_get_wtime()
    do step = 1, warmup + nsteps
        if (step == warmup + 1) t0 = omp_get_wtime()
#if (RK == 2)
        call RK2(param, SP, VM, cell)
#else
        call RK4(param, SP, VM, cell)
#endif
#if (TARGET_PROBLEM != 1)
        call PST(param, SP, VM, cell)
#else
        call check_steady_state(is_steady, param, SP)
        if (dump_every > 0) then
            df = maxval(abs(SP%f_previous - SP%u))
            write(arg, '(a,i0,a)') "/steady_", step, ".txt"
            open(newunit=un, file=trim(outdir)//trim(arg), status="replace")
            write(un, '(i0)') is_steady
            close(un)
        endif
#endif
        if (dump_every > 0) then
            if (mod(step, dump_every) == 0 .or. step == nsteps .or. step <= 2) then
                write(arg, '(a,i0,a)') "/step_", step, ".dat"
                call write_state(trim(outdir) // trim(arg))
            endif
        endif
    enddo
    t1 = omp_get_wtime()

    if (dump_every == 0) then
        write(*, '(a, es16.8)') "particle_steps_per_s ", dble(SP%num_int) * dble(nsteps) / (t1 - t0)
        write(*, '(a, es16.8)') "seconds ", t1 - t0
        write(*, '(a, i0)') "num_int ", SP%num_int
        write(*, '(a, i0)') "threads ", nthreads
    endif

contains

    subroutine write_params(fname)
        character(len=*), intent(in) :: fname
        integer :: u
        open(newunit=u, file=fname, status="replace")
        write(u, '(a, 1x, es26.17e3)') "dx", param%dx
        write(u, '(a, 1x, es26.17e3)') "SP_mass", param%SP_mass
        write(u, '(a, 1x, es26.17e3)') "h", param%h
        write(u, '(a, 1x, es26.17e3)') "h_eff", param%h_eff
        write(u, '(a, 1x, es26.17e3)') "W_ave", param%W_ave
        write(u, '(a, 1x, i0)') "num_WL", param%num_WL
        write(u, '(a, 1x, es26.17e3)') "WL_thick", param%WL_thick
        write(u, '(a, 1x, es26.17e3)') "delta_tem", param%delta_tem
        write(u, '(a, 1x, es26.17e3)') "tem_ave", param%tem_ave
        write(u, '(a, 1x, es26.17e3)') "sound_speed", param%sound_speed
        write(u, '(a, 1x, es26.17e3)') "kinematic_vis", param%kinematic_vis
        write(u, '(a, 1x, es26.17e3)') "thermal_dif", param%thermal_dif
        write(u, '(a, 1x, es26.17e3)') "Re", param%Re
        write(u, '(a, 1x, es26.17e3)') "Ra", param%Ra
        write(u, '(a, 1x, es26.17e3)') "Pr", param%Pr
        write(u, '(a, 1x, es26.17e3)') "dt_CFL", param%dt_CFL
        write(u, '(a, 1x, es26.17e3)') "dt_vis", param%dt_vis
        write(u, '(a, 1x, es26.17e3)') "dt_th", param%dt_th
        write(u, '(a, 1x, es26.17e3)') "dt", param%dt
        write(u, '(a, 1x, i0)') "num_total", SP%num_total
        write(u, '(a, 1x, i0)') "num_int", SP%num_int
        write(u, '(a, 1x, i0)') "num_ext", SP%num_ext
        close(u)
    end subroutine write_params

    subroutine write_state_of(S, fname)
        type(SP_type), intent(in) :: S
        character(len=*), intent(in) :: fname
        integer :: u
        open(newunit=u, file=fname, status="replace", form="unformatted", access="stream")
        write(u) S%ptype(1:S%num_total)
        write(u) S%x(1:S%num_total)
        write(u) S%y(1:S%num_total)
        write(u) S%u(1:S%num_total)
        write(u) S%v(1:S%num_total)
        write(u) S%rho(1:S%num_total)
        write(u) S%pre(1:S%num_total)
#if (TARGET_PROBLEM == 4)
        write(u) S%tem(1:S%num_total)
#endif
        close(u)
    end subroutine write_state_of

    subroutine write_state(fname)
        character(len=*), intent(in) :: fname
        call write_state_of(SP, fname)
    end subroutine write_state

    subroutine read_state(fname)
        character(len=*), intent(in) :: fname
        integer :: u
        open(newunit=u, file=fname, status="old", form="unformatted", access="stream")
        read(u) SP%ptype(1:SP%num_total)
        read(u) SP%x(1:SP%num_total)
        read(u) SP%y(1:SP%num_total)
        read(u) SP%u(1:SP%num_total)
        read(u) SP%v(1:SP%num_total)
        read(u) SP%rho(1:SP%num_total)
        read(u) SP%pre(1:SP%num_total)
#if (TARGET_PROBLEM == 4)
        read(u) SP%tem(1:SP%num_total)
#endif
        close(u)
    end subroutine read_state

    subroutine write_vm(fname)
        character(len=*), intent(in) :: fname
        integer :: u
        open(newunit=u, file=fname, status="replace", form="unformatted", access="stream")
        write(u) VM%ptype(1:SP%num_ext)
        write(u) VM%x(1:SP%num_ext)
        write(u) VM%y(1:SP%num_ext)
        write(u) VM%u(1:SP%num_ext)
        write(u) VM%v(1:SP%num_ext)
        write(u) VM%pre(1:SP%num_ext)
#if (TARGET_PROBLEM == 4)
        write(u) VM%tem(1:SP%num_ext)
#endif
        close(u)
    end subroutine write_vm

    subroutine write_rhs(fname)
        character(len=*), intent(in) :: fname
        integer :: u
        open(newunit=u, file=fname, status="replace", form="unformatted", access="stream")
        write(u) SP%RHS_u(1:SP%num_int)
#if (TARGET_PROBLEM != 1)
        write(u) SP%RHS_v(1:SP%num_int)
        write(u) SP%RHS_rho(1:SP%num_int)
#endif
#if (TARGET_PROBLEM == 4)
        write(u) SP%RHS_tem(1:SP%num_int)
#endif
        close(u)
    end subroutine write_rhs

    ! Neighbor sets as LSSPH_A visits them (LSSPH_A_mod.f90:102-124):
    ! per fluid particle its count, then the 1-based indices in visit order.
    subroutine write_neighbors(fname)
        character(len=*), intent(in) :: fname
        integer :: u, me, nc, c, k, you, cnt
        integer, allocatable :: counts(:), lst(:)
        integer :: total
        double precision :: xj, yj
        allocate(counts(SP%num_int))
        allocate(lst(SP%num_int * 200))
        total = 0
        do me = 1, SP%num_int
            cnt = 0
            do nc = 1, 9
                c = cell%neighbors(nc, cell%idx_SP(me))
                if (c == 0) cycle
                do k = 1, cell%num_SP(c)
                    you = cell%list(k, c)
                    if (you == me) cycle
                    xj = SP%x(you) - SP%x(me)
                    yj = SP%y(you) - SP%y(me)
                    if (xj**2 + yj**2 > param%h_eff**2) cycle
                    cnt = cnt + 1
                    total = total + 1
                    lst(total) = you
                enddo
            enddo
            counts(me) = cnt
        enddo
        open(newunit=u, file=fname, status="replace", form="unformatted", access="stream")
        write(u) counts(1:SP%num_int)
        write(u) lst(1:total)
        close(u)
    end subroutine write_neighbors

end program ref_harness
