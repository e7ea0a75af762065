! host_setup.f90 -- one-time host setup: particle lattice and types, initial
! conditions, restart.  Follows source/setup/setup_closed_box_mod.f90,
! impose_initial_conditions_mod.f90, restart_mod.f90 and
! init_particle_mod.f90 (fluid particles first, then the ghost walls), plus
! the periodic box and the seeded SplitMix64 jitter of this host.

#include "config.h"
#include "lssph_defaults.h"

module setup_closed_box_mod
    use global_types, only: Param_type, SP_type
    implicit none
contains
    subroutine alloc_sp(SP, n_int, n_ext)
        type(SP_type), intent(inout) :: SP
        integer, intent(in) :: n_int, n_ext
        integer :: n
        n = n_int + n_ext
        SP%num_int = n_int
        SP%num_ext = n_ext
        SP%num_total = n
        allocate(SP%ptype(n), SP%x(n), SP%y(n), SP%u(n), SP%v(n), SP%rho(n), SP%pre(n))
        allocate(SP%tem(n))
        allocate(SP%RHS_u(n_int), SP%RHS_v(n_int), SP%RHS_rho(n_int), SP%RHS_tem(n_int))
        SP%u = 0.0d0
        SP%v = 0.0d0
        SP%pre = 0.0d0
        SP%tem = 0.0d0
        SP%rho = 0.0d0
    end subroutine alloc_sp

    ! Lattice offset by dx/2, built by repeated addition (setup_closed_box_mod
    ! .f90:72-82): column positions of a row and row positions.
    subroutine lattice_1d(n, dx, c)
        integer, intent(in) :: n
        double precision, intent(in) :: dx
        double precision, intent(out) :: c(n)
        integer :: i
        c(1) = dx / 2.0d0
        do i = 2, n
            c(i) = c(i - 1) + dx
        enddo
    end subroutine lattice_1d

    ! Closed box with num_WL wall layers (setup_closed_box_mod.f90:23-237).
    subroutine setup_closed_box(param, SP)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        integer :: nxf, nyf, i, j, ki, ke, n_int, n_ext
        double precision :: x_min, x_max, y_min, y_max, xx, yy
        double precision, allocatable :: cx(:), cy(:)
        logical :: fluid
        nxf = param%num_x + 2 * param%num_WL
        nyf = param%num_y + 2 * param%num_WL
        allocate(cx(nxf), cy(nyf))
        call lattice_1d(nxf, param%dx, cx)
        call lattice_1d(nyf, param%dx, cy)
        x_min = 0.0d0 + param%WL_thick
        x_max = param%len_x + param%WL_thick
        y_min = 0.0d0 + param%WL_thick
        y_max = param%len_y + param%WL_thick
        n_int = 0
        do j = 1, nyf
            do i = 1, nxf
                if (x_min < cx(i) .and. x_max > cx(i) .and. y_min < cy(j) .and. y_max > cy(j)) n_int = n_int + 1
            enddo
        enddo
        n_ext = nxf * nyf - n_int
        call alloc_sp(SP, n_int, n_ext)
        ki = 0
        ke = n_int
        do j = 1, nyf
            do i = 1, nxf
                xx = cx(i)
                yy = cy(j)
                fluid = (x_min < xx .and. x_max > xx .and. y_min < yy .and. y_max > yy)
                if (fluid) then
                    ki = ki + 1
                    SP%x(ki) = xx
                    SP%y(ki) = yy
                    SP%ptype(ki) = 0
                else
                    ke = ke + 1
                    SP%x(ke) = xx
                    SP%y(ke) = yy
                    SP%rho(ke) = param%rho_ref
                    if (yy < y_min) then
                        SP%ptype(ke) = 1
                        if (xx < x_min) SP%ptype(ke) = 5
                        if (xx > x_max) SP%ptype(ke) = 6
                    elseif (yy > y_max) then
                        SP%ptype(ke) = 2
                        if (xx < x_min) SP%ptype(ke) = 7
                        if (xx > x_max) SP%ptype(ke) = 8
                    elseif (xx < x_min) then
                        SP%ptype(ke) = 3
                    else
                        SP%ptype(ke) = 4
                    endif
                endif
            enddo
        enddo
    end subroutine setup_closed_box

    ! Doubly periodic box: num_x x num_y fluid particles, no walls.
    subroutine setup_periodic_box(param, SP)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        integer :: i, j, k
        double precision, allocatable :: cx(:), cy(:)
        allocate(cx(param%num_x), cy(param%num_y))
        call lattice_1d(param%num_x, param%dx, cx)
        call lattice_1d(param%num_y, param%dx, cy)
        call alloc_sp(SP, param%num_x * param%num_y, 0)
        k = 0
        do j = 1, param%num_y
            do i = 1, param%num_x
                k = k + 1
                SP%x(k) = cx(i)
                SP%y(k) = cy(j)
                SP%ptype(k) = 0
                SP%rho(k) = param%rho_ref
            enddo
        enddo
    end subroutine setup_periodic_box
end module setup_closed_box_mod


module impose_initial_conditions_mod
    use iso_fortran_env, only: int64
    use global_types, only: Param_type, SP_type
    implicit none
    double precision, parameter, private :: pi = acos(-1.0d0)
contains
    ! k-th uniform in [0, 1) of the SplitMix64 stream of seed (Steele, Lea &
    ! Flood 2014); 64-bit integer arithmetic wraps modulo 2**64.
    function splitmix_uniform(seed, k) result(r)
        integer(int64), intent(in) :: seed, k
        double precision :: r
        integer(int64) :: z
        integer(int64), parameter :: golden = -7046029254386353131_int64   ! 0x9E3779B97F4A7C15
        integer(int64), parameter :: m1 = -4658895280553007687_int64       ! 0xBF58476D1CE4E5B9
        integer(int64), parameter :: m2 = -7723592293110705685_int64       ! 0x94D049BB133111EB
        z = seed + k * golden
        z = ieor(z, ishft(z, -30)) * m1
        z = ieor(z, ishft(z, -27)) * m2
        z = ieor(z, ishft(z, -31))
        r = dble(ishft(z, -11)) * (1.0d0 / 9007199254740992.0d0)
    end function splitmix_uniform

    ! +-pos_pert*dx jitter of the fluid positions: x from draw 2i-1, y from
    ! draw 2i of the stream (1-based draws), wrapped into a periodic box.
    subroutine jitter(param, SP)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        integer :: i
        integer(int64) :: seed
        seed = int(JITTER_SEED, int64)
        do i = 1, SP%num_int
            SP%x(i) = SP%x(i) + param%pos_pert * param%dx * (2.0d0 * splitmix_uniform(seed, 2_int64*i - 1) - 1.0d0)
            SP%y(i) = SP%y(i) + param%pos_pert * param%dx * (2.0d0 * splitmix_uniform(seed, 2_int64*i) - 1.0d0)
            if (param%periodic == 1) then
                SP%x(i) = modulo(SP%x(i), param%len_x)
                SP%y(i) = modulo(SP%y(i), param%len_y)
            endif
        enddo
    end subroutine jitter

    subroutine impose_diffusion_equation(param, SP)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        integer :: i
        SP%rho(:) = param%rho_ref
        SP%u(:) = 0.0d0
        SP%pre(:) = 0.0d0
        SP%v(:) = 0.0d0
#if (HEAT_MODE == 1)
        do i = 1, SP%num_int
            SP%u(i) = sin(2.0d0 * pi * SP%x(i) / param%len_x) * sin(2.0d0 * pi * SP%y(i) / param%len_y)
        enddo
#else
        i = 0
#endif
    end subroutine impose_diffusion_equation

    ! impose_initial_conditions_mod.f90:61-96
    subroutine impose_TG_vortex(param, SP)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        integer :: i
        double precision :: x, y, a, b, Lx, Ly, ratio
        a = dble(param%tg_a)
        b = dble(param%tg_b)
        Lx = param%len_x
        Ly = param%len_y
        ratio = (a * Ly) / (b * Lx)
        do i = 1, SP%num_int
            x = SP%x(i) - param%WL_thick
            y = SP%y(i) - param%WL_thick
            SP%u(i) = + 1.0d0 * sin(a*pi*x / Lx) * cos(b*pi*y / Ly)
            SP%v(i) = - 1.0d0 * ratio * cos(a*pi*x / Lx) * sin(b*pi*y / Ly)
            SP%pre(i) = 0.25d0 * param%rho_ref * 1.0d0 * &
                        (cos(2.0d0*a*pi*x / Lx) + (ratio*ratio) * cos(2.0d0*b*pi*y / Ly))
            SP%rho(i) = param%rho_ref + SP%pre(i) / (param%sound_speed * param%sound_speed)
        enddo
    end subroutine impose_TG_vortex

    subroutine impose_cavity(param, SP)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        SP%rho(:) = param%rho_ref
        SP%pre(:) = 0.0d0
        SP%u(:) = 0.0d0
        SP%v(:) = 0.0d0
    end subroutine impose_cavity

    ! impose_initial_conditions_mod.f90:119-143
    subroutine impose_boussinesq_convection(param, SP)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        integer :: i
        double precision :: x, y
        call impose_cavity(param, SP)
        do i = 1, SP%num_int
            x = SP%x(i) - param%WL_thick
            y = SP%y(i) - param%WL_thick
            SP%tem(i) = - (param%delta_tem / param%len_y) * y + param%tem_bottom
            if (y <= 0.1d0 * param%len_y) then
                SP%tem(i) = SP%tem(i) + 1.0d-2 * param%delta_tem * cos(pi*x / param%len_x)
            endif
        enddo
    end subroutine impose_boussinesq_convection
end module impose_initial_conditions_mod


module restart_mod
    use global_types, only: Param_type, SP_type
    implicit none
contains
    ! Reads results/READ_NAME/data/START_STEP.dat (restart_mod.f90:32-53).
    subroutine restart_closed_box(param, SP)
        type(Param_type), intent(in)    :: param
        type(SP_type)   , intent(inout) :: SP
        character(len=1024) :: restart_id, file_name
        integer :: un
        write(restart_id, *) param%start_step
        file_name = "results/" // trim(adjustl(param%read_name)) // "/data/" // trim(adjustl(restart_id)) // ".dat"
        open(newunit=un, file=trim(adjustl(file_name)), status="old", form="unformatted", access="stream")
        read(un) SP%ptype(1:SP%num_total)
        read(un) SP%x(1:SP%num_total)
        read(un) SP%y(1:SP%num_total)
        read(un) SP%u(1:SP%num_total)
        read(un) SP%v(1:SP%num_total)
        read(un) SP%rho(1:SP%num_total)
        read(un) SP%pre(1:SP%num_total)
#if (TARGET_PROBLEM == 4)
        read(un) SP%tem(1:SP%num_total)
#endif
        close(un)
    end subroutine restart_closed_box
end module restart_mod


module init_particle_mod
    use global_types, only: Param_type, SP_type
    use setup_closed_box_mod, only: setup_closed_box, setup_periodic_box
    use impose_initial_conditions_mod
    use restart_mod, only: restart_closed_box
    implicit none
contains
    ! init_particle_mod.f90:46-86 (+ seeded jitter before the initial field)
    subroutine init_particle(param, SP)
        type(Param_type), intent(in)  :: param
        type(SP_type)   , intent(out) :: SP
        if (param%periodic == 1) then
            call setup_periodic_box(param, SP)
        else
            call setup_closed_box(param, SP)
        endif
        if (trim(adjustl(param%read_name)) == "new") then
            if (param%pos_pert > 0.0d0) call jitter(param, SP)
#if (TARGET_PROBLEM == 1)
            call impose_diffusion_equation(param, SP)
#elif (TARGET_PROBLEM == 2)
            call impose_TG_vortex(param, SP)
#elif (TARGET_PROBLEM == 3)
            call impose_cavity(param, SP)
#else
            call impose_boussinesq_convection(param, SP)
#endif
        else
            call restart_closed_box(param, SP)
        endif
    end subroutine init_particle
end module init_particle_mod
