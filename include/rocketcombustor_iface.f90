!! rocketcombustor_iface.f90 -- ISO_C_BINDING interface of the C-ABI in
!! rocketcombustor.h, for a Fortran host that keeps the reference's call
!! sequence (src/fvm2d.f90 main loop -> lu_sgs_stage).  Arrays are passed in
!! the reference's own layouts: Coor(2,NNT), N(4,NE) (the first three rows
!! are read), PA/WA/hWA(Nvar,NE), GradC(2,Nvar,NE), alpaF(NF), step(NE).
!! Every function returns 0 or an RC_ERR_* code; rc_check turns a nonzero
!! code into the reference's `stop` with the library's message.
module rocketcombustor_iface
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: RC_MAXSP = 8
    integer(c_int), parameter :: RC_OK = 0, RC_ERR_ARGUMENT = 1, RC_ERR_TEMPERATURE = 2, &
                                 RC_ERR_NEGATIVE = 3, RC_ERR_NAN = 4, RC_ERR_MESH = 5, &
                                 RC_ERR_DEVICE = 6, RC_ERR_MEMORY = 7, RC_ERR_SINGULAR = 8

    type, bind(c) :: rc_mechanism
        integer(c_int) :: nsp
        real(c_double) :: mole_weight(RC_MAXSP)
        real(c_double) :: rconstant
        real(c_double) :: t_low, t_crit, t_high
        real(c_double) :: low_coef(7, RC_MAXSP)      ! C low_coef[k][j] = low_coef(j+1,k+1)
        real(c_double) :: high_coef(7, RC_MAXSP)
        real(c_double) :: low_b1(RC_MAXSP), high_b1(RC_MAXSP)
        integer(c_int) :: model
        real(c_double) :: coef_f(RC_MAXSP), coef_r(RC_MAXSP)
        real(c_double) :: a_f, beta, e_act
    end type

    type, bind(c) :: rc_params
        real(c_double) :: cfl
        integer(c_int) :: limit_id
        integer(c_int) :: combustion
        real(c_double) :: p_out
        real(c_double) :: ox_y(RC_MAXSP), ox_t, ox_mass_flux
        real(c_double) :: fu_y(RC_MAXSP), fu_t, fu_mass_flux
    end type

    !! transport.in / diffusion.in tables (C mu_low[k][j] = mu_low(j+1,k+1))
    type, bind(c) :: rc_transport
        real(c_double) :: t_low, t_crit, t_high
        real(c_double) :: mu_low(4, RC_MAXSP), mu_high(4, RC_MAXSP)
        real(c_double) :: k_low(4, RC_MAXSP), k_high(4, RC_MAXSP)
        real(c_double) :: k_b
        real(c_double) :: sigma(RC_MAXSP), eps(RC_MAXSP)
    end type

    interface
        integer(c_int) function rc_case_create(mech, par, nnt, coor, ne, cellnodes, nb, bnodes, btype, &
                                               device, handle) bind(c, name="rc_case_create")
            import :: c_int, c_double, c_ptr, rc_mechanism, rc_params
            type(rc_mechanism), intent(in) :: mech
            type(rc_params), intent(in) :: par
            integer(c_int), value :: nnt, ne, nb, device
            real(c_double), intent(in) :: coor(*)
            integer(c_int), intent(in) :: cellnodes(*), bnodes(*), btype(*)
            type(c_ptr), intent(out) :: handle
        end function

        integer(c_int) function rc_case_create_periodic(mech, par, nnt, coor, ne, cellnodes, nb, bnodes, btype, &
                                                        ncon, con, dxyz, device, handle) &
                                                        bind(c, name="rc_case_create_periodic")
            import :: c_int, c_double, c_ptr, rc_mechanism, rc_params
            type(rc_mechanism), intent(in) :: mech
            type(rc_params), intent(in) :: par
            integer(c_int), value :: nnt, ne, nb, ncon, device
            real(c_double), intent(in) :: coor(*), dxyz(*)
            integer(c_int), intent(in) :: cellnodes(*), bnodes(*), btype(*), con(*)
            type(c_ptr), intent(out) :: handle
        end function

        subroutine rc_case_destroy(h) bind(c, name="rc_case_destroy")
            import :: c_ptr
            type(c_ptr), value :: h
        end subroutine

        type(c_ptr) function rc_last_error(h) bind(c, name="rc_last_error")
            import :: c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function rc_case_info(h, sizes6) bind(c, name="rc_case_info")
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: h
            integer(c_long_long), intent(out) :: sizes6(6)
        end function

        integer(c_int) function rc_case_levels(h, cells, n) bind(c, name="rc_case_levels")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), intent(out) :: cells(*)
            integer(c_int), value :: n
        end function

        integer(c_int) function rc_set_primitive(h, pa) bind(c, name="rc_set_primitive")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: pa(*)
        end function

        integer(c_int) function rc_timestep(h) bind(c, name="rc_timestep")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function rc_gradient(h) bind(c, name="rc_gradient")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function rc_limiter_barth(h) bind(c, name="rc_limiter_barth")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function rc_reconstruction(h) bind(c, name="rc_reconstruction")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function rc_flux_sourceterm(h) bind(c, name="rc_flux_sourceterm")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function rc_lusgs_steady(h) bind(c, name="rc_lusgs_steady")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function rc_lusgs_stage(h) bind(c, name="rc_lusgs_stage")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function

        integer(c_int) function rc_run_steady(h, nsteps, every, res_out, done) bind(c, name="rc_run_steady")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            integer(c_int), value :: nsteps, every
            real(c_double), intent(out) :: res_out(*)
            integer(c_int), intent(out) :: done
        end function

        !! viscous path (IfViscous): the transport and diffusion tables
        integer(c_int) function rc_set_transport(h, tr) bind(c, name="rc_set_transport")
            import :: c_int, c_ptr, rc_transport
            type(c_ptr), value :: h
            type(rc_transport), intent(in) :: tr
        end function

        !! dual-time backward Euler (implicit_time_method = 3): compute_p_out,
        !! backwardEuler_LU_SGS(Inner_rk), backwardEuler_stage(istep)
        integer(c_int) function rc_set_outlet_pressure(h, p_out) bind(c, name="rc_set_outlet_pressure")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: p_out
        end function

        integer(c_int) function rc_backward_euler_lusgs(h, dt, inner) bind(c, name="rc_backward_euler_lusgs")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: dt
            integer(c_int), value :: inner
        end function

        integer(c_int) function rc_backward_euler_stage(h, dt, innerloop, reltol, ninner, res_out) &
                bind(c, name="rc_backward_euler_stage")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: dt, reltol
            integer(c_int), value :: innerloop
            integer(c_int), intent(out) :: ninner
            real(c_double), intent(out) :: res_out(*)
        end function

        !! dual-time BDF2 (implicit_time_method = 5): BDF2_stage(istep) with
        !! first = 1 at istep == 1 (Time_BDF2.f90:4-69)
        integer(c_int) function rc_bdf2_stage(h, dt, innerloop, reltol, first, ninner, res_out) &
                bind(c, name="rc_bdf2_stage")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), value :: dt, reltol
            integer(c_int), value :: innerloop, first
            integer(c_int), intent(out) :: ninner
            real(c_double), intent(out) :: res_out(*)
        end function

        integer(c_int) function rc_residual(h, res) bind(c, name="rc_residual")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: res(*)
        end function

        integer(c_int) function rc_get_primitive(h, pa) bind(c, name="rc_get_primitive")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: pa(*)
        end function

        integer(c_int) function rc_get_conserved(h, wa) bind(c, name="rc_get_conserved")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: wa(*)
        end function

        integer(c_int) function rc_get_increment(h, hwa) bind(c, name="rc_get_increment")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: hwa(*)
        end function

        integer(c_int) function rc_get_gradient(h, gradc) bind(c, name="rc_get_gradient")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: gradc(*)
        end function

        integer(c_int) function rc_get_timestep(h, step, alpaf) bind(c, name="rc_get_timestep")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: step(*), alpaf(*)
        end function

        integer(c_int) function rc_get_ghost(h, gpa) bind(c, name="rc_get_ghost")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: gpa(*)
        end function

        integer(c_int) function rc_get_derived(h, rho, mach, prod) bind(c, name="rc_get_derived")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: rho(*), mach(*), prod(*)
        end function

        integer(c_int) function rc_set_timing(h, enable) bind(c, name="rc_set_timing")
            import :: c_int, c_ptr
            type(c_ptr), value :: h
            integer(c_int), value :: enable
        end function

        integer(c_int) function rc_get_timing(h, t8) bind(c, name="rc_get_timing")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: t8(8)
        end function

        integer(c_int) function rc_mesh_order_multicolor(ne, cellnodes, block, perm, ncolors) &
                bind(c, name="rc_mesh_order_multicolor")
            import :: c_int
            integer(c_int), value :: ne, block
            integer(c_int), intent(in) :: cellnodes(*)
            integer(c_int), intent(out) :: perm(*)
            integer(c_int), intent(out) :: ncolors
        end function

        integer(c_int) function rc_batch_create(device, b) bind(c, name="rc_batch_create")
            import :: c_int, c_ptr
            integer(c_int), value :: device
            type(c_ptr), intent(out) :: b
        end function

        subroutine rc_batch_destroy(b) bind(c, name="rc_batch_destroy")
            import :: c_ptr
            type(c_ptr), value :: b
        end subroutine

        integer(c_int) function rc_batch_add(b, h) bind(c, name="rc_batch_add")
            import :: c_int, c_ptr
            type(c_ptr), value :: b, h
        end function

        integer(c_int) function rc_batch_remove(b, h) bind(c, name="rc_batch_remove")
            import :: c_int, c_ptr
            type(c_ptr), value :: b, h
        end function

        integer(c_int) function rc_batch_info(b, info3) bind(c, name="rc_batch_info")
            import :: c_int, c_ptr, c_long_long
            type(c_ptr), value :: b
            integer(c_long_long), intent(out) :: info3(3)
        end function

        integer(c_int) function rc_batch_set_launch(b, cell_threads, run_threads, xcd) &
                bind(c, name="rc_batch_set_launch")
            import :: c_int, c_ptr
            type(c_ptr), value :: b
            integer(c_int), value :: cell_threads, run_threads, xcd
        end function

        integer(c_int) function rc_batch_run_steady(b, nsteps, every, res_out, done, err_case) &
                bind(c, name="rc_batch_run_steady")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: b
            integer(c_int), value :: nsteps, every
            real(c_double), intent(out) :: res_out(*)
            integer(c_int), intent(out) :: done, err_case
        end function

        integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

contains

    !! the library's last message as a Fortran string
    function rc_message(h) result(msg)
        type(c_ptr), intent(in) :: h
        character(len=:), allocatable :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: chars(:)
        integer :: n, i
        p = rc_last_error(h)
        n = int(c_strlen(p))
        call c_f_pointer(p, chars, [n])
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = chars(i)
        enddo
    end function

    !! turns a nonzero return code into the reference's `stop`
    subroutine rc_check(code, h, where)
        integer(c_int), intent(in) :: code
        type(c_ptr), intent(in) :: h
        character(len=*), intent(in) :: where
        if (code /= RC_OK) then
            write(*,*) 'error in ', where, ': ', rc_message(h)
            stop
        endif
    end subroutine

end module
