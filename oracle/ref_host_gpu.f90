!! Test infrastructure: the drop-in demonstration.  The reference solver's
!! own Fortran host -- start-up, mesh reading, initial field, restart and
!! reconstruction set-up (src/fvm2d.f90:15-24), its residual monitor
!! `step_residual` (src/residual.f90:61) and its backup writer `BackupData`
!! (src/output.f90:256) -- runs unchanged, linked from the reference's
!! objects; the only change is the steady step of the main loop
!! (src/fvm2d.f90:76), where `call lu_sgs_stage(istep)` becomes a call of
!! the MI355X library through include/rocketcombustor_iface.f90, and, for
!! unsteady decks with implicit_time_method=3 or 5, the dual-time stage
!! (src/fvm2d.f90:93 and :95), where `call backwardEuler_stage(istep)` and
!! `call BDF2_stage(istep)` become rc_backward_euler_stage and rc_bdf2_stage
!! after the reference's own compute_p_out hands
!! its exit pressure to the library.  Unsteady runs write the inner-iteration
!! rows (istep, inner, Res/total_volume, Romax) to OUTDIR/resid.bin and
!! converge_uns.dat, and, at the main loop's screen-output steps, the
!! reference's step_residual (converge.dat) of the residual the stage leaves
!! in hWA, with its rows (istep, Res/total_volume, Romax) in
!! OUTDIR/resid_step.bin.  Viscous decks hand the reference's transport and
!! diffusion tables to the library (rc_set_transport).
!!
!! usage (in a case directory): rc_ref_host_gpu OUTDIR
!! writes converge.dat and fvp1.sav as the reference does, and OUTDIR/resid.bin,
!! OUTDIR/final_PA.bin in the layout of oracle/ref_dump.f90.
program rc_ref_host_gpu
    use mainvar
    use Parallel_Var
    use ConnectFace
    use species
    use thermo_property
    use chemical_kinetics
    use oxidizer
    use fuel
    use transport_property
    use diffusion_velocity
    use rocketcombustor_iface
    implicit none
    include 'mpif.h'
    type(rc_mechanism) :: mech
    type(rc_params) :: par
    type(rc_transport) :: trp
    type(c_ptr) :: h
    integer(c_int), allocatable :: cells(:), bn(:), bt(:)
    integer :: i, ib, np, nb, istep, ierr, k, istop
    integer(c_int) :: ninner
    double precision :: time_all, RoRes, res(64), romax, ttime_n, totaltime_num
    double precision, allocatable :: rows(:, :)
    character(len=256) :: outdir

    call get_command_argument(1, outdir)

    !! set-up exactly as the reference's main program (fvm2d.f90:15-24)
    call startup
    call ReadMesh
    call initialization
    if (IfReadBackup) call Read_Backup
    call compute_reconstruction_matrix

    !! module state -> the library's tables
    mech%nsp = Nsp
    mech%mole_weight = 0.; mech%mole_weight(1:Nsp) = mole_weight(1:Nsp)
    mech%rconstant = Rconstant
    mech%t_low = thermo_lowest_tem; mech%t_crit = thermo_critical_tem; mech%t_high = thermo_highest_tem
    mech%low_coef = 0.; mech%high_coef = 0.
    mech%low_coef(1:7, 1:Nsp) = low_coef_Cp_species(1:7, 1:Nsp)
    mech%high_coef(1:7, 1:Nsp) = high_coef_Cp_species(1:7, 1:Nsp)
    mech%low_b1 = 0.; mech%high_b1 = 0.
    mech%low_b1(1:Nsp) = low_enthalpy_b1(1:Nsp); mech%high_b1(1:Nsp) = high_enthalpy_b1(1:Nsp)
    mech%model = code_chemical_model
    mech%coef_f = 0.; mech%coef_r = 0.
    mech%coef_f(1:Nsp) = coef_f_react(1:Nsp, 1); mech%coef_r(1:Nsp) = coef_r_react(1:Nsp, 1)
    mech%a_f = Af_react(1); mech%beta = beta_react(1); mech%e_act = E_react(1)

    par%cfl = cfl
    par%limit_id = limit_ID
    par%combustion = merge(1, 0, If_turn_on_combustion)
    par%p_out = p_out
    par%ox_y = 0.; par%ox_y(1:Nsp) = oxidizer_mass_fraction(1:Nsp)
    par%ox_t = oxidizer_temp_in; par%ox_mass_flux = oxidizer_mass_flow     ! per unit height after initialization.f90:22
    par%fu_y = 0.; par%fu_y(1:Nsp) = fuel_mass_fraction(1:Nsp)
    par%fu_t = fuel_temp_in; par%fu_mass_flux = fuel_mass_flow             ! initialization.f90:23

    allocate(cells(3*NE))
    do i = 1, NE
        cells(3*i-2:3*i) = N(1:3, i)
    enddo
    nb = sum(iBound(1:NBoundary))
    allocate(bn(2*nb), bt(nb))
    nb = 0
    do ib = 1, NBoundary
        do np = 1, iBound(ib)
            nb = nb + 1
            bn(2*nb-1:2*nb) = BoundN(1:2, np, ib)
            bt(nb) = iType(ib)
        enddo
    enddo

    call rc_check(rc_case_create(mech, par, NNT, Coor, NE, cells, nb, bn, bt, 0, h), c_null_ptr, 'rc_case_create')
    if (IfViscous) then
        !! module transport_property / diffusion_velocity -> the library's tables
        trp%t_low = transport_lowest_tem; trp%t_crit = transport_critical_tem
        trp%t_high = transport_highest_tem
        trp%mu_low = 0.; trp%mu_high = 0.; trp%k_low = 0.; trp%k_high = 0.
        trp%mu_low(1, 1:Nsp) = low_mu_A(1:Nsp); trp%mu_low(2, 1:Nsp) = low_mu_B(1:Nsp)
        trp%mu_low(3, 1:Nsp) = low_mu_C(1:Nsp); trp%mu_low(4, 1:Nsp) = low_mu_D(1:Nsp)
        trp%mu_high(1, 1:Nsp) = high_mu_A(1:Nsp); trp%mu_high(2, 1:Nsp) = high_mu_B(1:Nsp)
        trp%mu_high(3, 1:Nsp) = high_mu_C(1:Nsp); trp%mu_high(4, 1:Nsp) = high_mu_D(1:Nsp)
        trp%k_low(1, 1:Nsp) = low_K_A(1:Nsp); trp%k_low(2, 1:Nsp) = low_K_B(1:Nsp)
        trp%k_low(3, 1:Nsp) = low_K_C(1:Nsp); trp%k_low(4, 1:Nsp) = low_K_D(1:Nsp)
        trp%k_high(1, 1:Nsp) = high_K_A(1:Nsp); trp%k_high(2, 1:Nsp) = high_K_B(1:Nsp)
        trp%k_high(3, 1:Nsp) = high_K_C(1:Nsp); trp%k_high(4, 1:Nsp) = high_K_D(1:Nsp)
        trp%k_b = k_B
        trp%sigma = 0.; trp%eps = 0.
        trp%sigma(1:Nsp) = collision_diameter(1:Nsp); trp%eps(1:Nsp) = LJ_energy(1:Nsp)
        call rc_check(rc_set_transport(h, trp), h, 'rc_set_transport')
    endif
    call rc_check(rc_set_primitive(h, PA), h, 'rc_set_primitive')

    open(91, file=trim(outdir)//'/resid.bin', access='stream', form='unformatted', status='replace')
    time_all = 0.
    if (.not. IfSteady) then
        !! the main loop, unsteady implicit branch with backward Euler
        !! (fvm2d.f90:48-120)
        if (IfExplicit .or. (implicit_time_method /= 3 .and. implicit_time_method /= 5)) then
            write(*,*) 'rc_ref_host_gpu: unsteady runs need implicit_time_method=3 or 5'
            stop 2
        endif
        allocate(rows(Nvar + 1, InnerLoop))
        open(94, file=trim(outdir)//'/resid_step.bin', access='stream', form='unformatted', status='replace')
        open(93, file='converge_uns.dat')
        write(93,*) 'VARIABLES = "Time step","Inner iteration step", "L1(Res)","Romax"'
        totaltime_num = totaltime*(1.-1.e-10)
        istop = 0
        do istep = 1, nstep_max
            if (ttime + dt > totaltime_num) then
                dt = totaltime - ttime
                istop = 1
            endif
            ttime_n = ttime
            ttime = ttime_n + cs(1)*dt
            call compute_p_out
            call rc_check(rc_set_outlet_pressure(h, p_out), h, 'compute_p_out')
            if (implicit_time_method == 5) then
                call rc_check(rc_bdf2_stage(h, dt, InnerLoop, relTol, merge(1, 0, istep == 1), ninner, rows), &
                              h, 'BDF2_stage')                          ! was: call BDF2_stage(istep)
            else
                call rc_check(rc_backward_euler_stage(h, dt, InnerLoop, relTol, ninner, rows), h, &
                              'backwardEuler_stage')                    ! was: call backwardEuler_stage(istep)
            endif
            do k = 1, ninner
                write(91) dble(istep), dble(k), rows(:, k)
                write(93,*) istep, k, rows(1:Nvar, k), rows(Nvar + 1, k)
            enddo
            ttime = ttime_n + dt
            call compute_p_out
            if (istep == 1 .or. mod(istep, NScreenOutput) == 0) then
                ! hWA is the residual the stage's closing flux_sourceterm left
                call rc_check(rc_get_increment(h, hWA), h, 'rc_get_increment')
                call rc_check(rc_get_conserved(h, WA), h, 'rc_get_conserved')
                call step_residual(istep, time_all, RoRes)             ! the reference's converge.dat
                call Residual(res(1:Nvar), romax)
                res(1:Nvar) = res(1:Nvar) / total_volume
                write(94) dble(istep), res(1:Nvar), romax
            endif
            if (istop == 1) exit
        enddo
        close(93)
        close(94)
    else
    !! the main loop, steady implicit branch (fvm2d.f90:48-150)
    do istep = 1, nstep_max
        call rc_check(rc_lusgs_stage(h), h, 'lu_sgs_stage')            ! was: call lu_sgs_stage(istep)
        if (istep == 1 .or. mod(istep, NScreenOutput) == 0) then
            call rc_check(rc_get_increment(h, hWA), h, 'rc_get_increment')
            call rc_check(rc_get_conserved(h, WA), h, 'rc_get_conserved')
            call step_residual(istep, time_all, RoRes)                 ! the reference's converge.dat
            call Residual(res(1:Nvar), romax)
            res(1:Nvar) = res(1:Nvar) / total_volume
            write(91) res(1:Nvar), romax
        endif
    enddo
    endif
    close(91)

    call rc_check(rc_get_primitive(h, PA), h, 'rc_get_primitive')
    call rc_check(rc_get_conserved(h, WA), h, 'rc_get_conserved')
    call BackupData
    open(92, file=trim(outdir)//'/final_PA.bin', access='stream', form='unformatted', status='replace')
    write(92) PA(:, 1:NE)
    close(92)
    call rc_case_destroy(h)
    write(*,*) 'rc_ref_host_gpu finished', min(istep, nstep_max), 'steps'
    call MPI_Finalize(ierr)
end program
