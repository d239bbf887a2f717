!! Test infrastructure only: a driver program linked against the reference
!! solver's own objects (built from the sources under the reference tree by
!! oracle/ref_build.sh into oracle/_ref/).  It replaces the reference's
!! `program main` (src/fvm2d.f90:7-177) and runs the same set-up sequence
!! (src/fvm2d.f90:15-24) and the same steady step as `lu_sgs_stage`
!! (src/Time_LUSGS_steady.f90:4-30), calling each stage separately so the
!! per-call arrays can be written out as golden vectors.
!!
!! usage (in a case directory holding grid1.in, fvp1.sav, flow/, combustion/):
!!   rc_ref_dump NSTEPS NDUMP OUTDIR [TIMING]
!! Steps 1..NDUMP write every per-call array; every step appends the
!! volume-normalised L1 norms of dW and max(rho) (src/residual.f90:61-79) to
!! OUTDIR/resid.bin; the final PA/WA are written to OUTDIR/final_*.bin.
!! With TIMING=1 nothing but resid.bin/final is written and per-stage wall
!! times are printed.
!!
!! Unsteady decks with implicit_time_method=3 run the dual-time backward
!! Euler steps instead (fvm2d.f90:60-120, backwardEuler_stage of
!! Time_BackwardEuler.f90:4-56 with its closing reconstruction/flux_sourceterm,
!! then step_residual's norms into OUTDIR/resid_step.bin), and decks with
!! implicit_time_method=5 the BDF2 steps the same way (BDF2_stage of
!! Time_BDF2.f90:4-69 with its WP/WP2 prologue and BDF2_LU_SGS): inner
!! iteration k overall (counted across steps) dumps s<k>_res, s<k>_dPA,
!! s<k>_PA, s<k>_WA for k <= NDUMP, and resid.bin holds rows
!! (istep, inner, Res(1:Nvar)/total_volume, Romax).
program rc_ref_dump
    use mainvar
    use Parallel_Var
    implicit none
    include 'mpif.h'
    integer :: istep, nsteps, ndump, ierr, timing, k, inner, kin
    character(len=256) :: arg, outdir
    double precision :: res(64), romax, t0, t1, tstage(6), roress, roress1, ttime_n, phi
    integer(8) :: c0, c1, crate

    call get_command_argument(1, arg); read(arg,*) nsteps
    call get_command_argument(2, arg); read(arg,*) ndump
    call get_command_argument(3, outdir)
    timing = 0
    if (command_argument_count() >= 4) then
        call get_command_argument(4, arg); read(arg,*) timing
    endif

    call startup
    call ReadMesh
    call initialization
    if (IfReadBackup) call Read_Backup
    call compute_reconstruction_matrix

    if (timing == 0) call dump_setup(trim(outdir))
    call dump_d(trim(outdir)//'/init_PA.bin', PA(:,1:NE))
    call dump_d(trim(outdir)//'/init_WA.bin', WA(:,1:NE))

    open(91, file=trim(outdir)//'/resid.bin', access='stream', form='unformatted', status='replace')
    tstage = 0.
    call system_clock(c0, crate)
    if (.not. IfSteady) then
        if (IfExplicit .or. (implicit_time_method /= 3 .and. implicit_time_method /= 5)) then
            write(*,*) 'rc_ref_dump: unsteady runs are dumped for implicit_time_method=3 or 5 only'
            stop 2
        endif
        kin = 0
        open(94, file=trim(outdir)//'/resid_step.bin', access='stream', form='unformatted', status='replace')
        do istep = 1, nsteps
            ttime_n = ttime
            ttime = ttime_n + cs(1)*dt
            call compute_p_out
            if (implicit_time_method == 5) then
                ! the prologue of BDF2_stage (Time_BDF2.f90:12-20)
                if (istep == 1) then
                    WP = WA
                    WP2 = WP
                    phi = 0.
                else
                    WP2 = WP
                    WP = WA
                    phi = 0.5
                endif
            endif
            do inner = 1, InnerLoop
                kin = kin + 1
                call TimeStep
                call reconstruction
                call flux_sourceterm
                if (kin <= ndump) call dump_step_d(outdir, kin, 'res', hWA(:,1:NE))
                if (implicit_time_method == 5) then
                    call BDF2_LU_SGS(inner, phi)
                else
                    call backwardEuler_LU_SGS(inner)
                endif
                if (kin <= ndump) then
                    call dump_step_d(outdir, kin, 'dPA', hWA(:,1:NE))
                    call dump_step_d(outdir, kin, 'PA', PA(:,1:NE))
                    call dump_step_d(outdir, kin, 'WA', WA(:,1:NE))
                endif
                call Residual(res(1:Nvar), romax)
                res(1:Nvar) = res(1:Nvar) / total_volume
                write(91) dble(istep), dble(inner), res(1:Nvar), romax
                roress = sum(res(Nflow+1:Nvar))
                if (inner == 1) roress1 = roress
                if (roress < roress1*relTol .or. inner == InnerLoop) then
                    ! the closing reconstruction + flux_sourceterm of
                    ! backwardEuler_stage (Time_BackwardEuler.f90:36-38)
                    call reconstruction
                    call flux_sourceterm
                    exit
                endif
                if (isnan(WA(1,1))) then
                    write(*,*) 'NaN at step', istep, 'inner', inner
                    stop 3
                endif
            enddo
            ttime = ttime_n + dt
            call compute_p_out
            ! step_residual after the stage (fvm2d.f90:137): norms of the
            ! residual the closing flux_sourceterm left in hWA
            call Residual(res(1:Nvar), romax)
            res(1:Nvar) = res(1:Nvar) / total_volume
            write(94) dble(istep), res(1:Nvar), romax
        enddo
        close(94)
    endif
    do istep = 1, merge(nsteps, 0, IfSteady)
        if (istep <= ndump .and. timing == 0) then
            call TimeStep
            call dump_step_d(outdir, istep, 'step', step(1:NE))
            call dump_step_d(outdir, istep, 'alpaF', alpaF(1:NF))
            call gradient
            call dump_step_d(outdir, istep, 'GPA', GPA)
            call dump_step_d(outdir, istep, 'gradC0', GradC(:,:,1:NE))
            if (limit_ID > 0) call limiter_Barth
            call dump_step_d(outdir, istep, 'gradC', GradC(:,:,1:NE))
            call flux_sourceterm
            call dump_step_d(outdir, istep, 'res', hWA(:,1:NE))
            call LU_SGS_steady
            call dump_step_d(outdir, istep, 'dW', hWA(:,1:NE))
            call dump_step_d(outdir, istep, 'PA', PA(:,1:NE))
            call dump_step_d(outdir, istep, 'WA', WA(:,1:NE))
            if (isnan(WA(1,1))) then
                write(*,*) 'NaN at step', istep
                stop 3
            endif
        else if (timing == 1) then
            call system_clock(c1); t0 = dble(c1)
            call TimeStep
            call system_clock(c1); tstage(1) = tstage(1) + (dble(c1)-t0)/crate; t0 = dble(c1)
            call gradient
            if (limit_ID > 0) call limiter_Barth
            call system_clock(c1); tstage(2) = tstage(2) + (dble(c1)-t0)/crate; t0 = dble(c1)
            call flux_sourceterm
            call system_clock(c1); tstage(3) = tstage(3) + (dble(c1)-t0)/crate; t0 = dble(c1)
            call LU_SGS_steady
            call system_clock(c1); tstage(4) = tstage(4) + (dble(c1)-t0)/crate; t0 = dble(c1)
            if (isnan(WA(1,1))) then
                write(*,*) 'NaN at step', istep
                stop 3
            endif
        else
            call lu_sgs_stage(istep)
        endif
        call Residual(res(1:Nvar), romax)
        res(1:Nvar) = res(1:Nvar) / total_volume
        write(91) res(1:Nvar), romax
    enddo
    call system_clock(c1)
    close(91)
    call dump_d(trim(outdir)//'/final_PA.bin', PA(:,1:NE))
    call dump_d(trim(outdir)//'/final_WA.bin', WA(:,1:NE))
    write(*,'(a,i10,a,es14.6,a,i10)') 'steps', nsteps, ' wall_s', dble(c1-c0)/crate, ' NE', NE
    if (timing == 1) then
        write(*,'(a,4es14.6)') 'stage_s timestep,reconstruction,flux_sourceterm,lusgs', (tstage(k), k=1,4)
    endif
    call MPI_Finalize(ierr)

contains

    subroutine dump_d(fname, a)
        character(len=*), intent(in) :: fname
        double precision, intent(in) :: a(..)
        select rank(a)
        rank(1)
            open(92, file=fname, access='stream', form='unformatted', status='replace'); write(92) a; close(92)
        rank(2)
            open(92, file=fname, access='stream', form='unformatted', status='replace'); write(92) a; close(92)
        rank(3)
            open(92, file=fname, access='stream', form='unformatted', status='replace'); write(92) a; close(92)
        end select
    end subroutine

    subroutine dump_i(fname, a)
        character(len=*), intent(in) :: fname
        integer, intent(in) :: a(..)
        select rank(a)
        rank(1)
            open(92, file=fname, access='stream', form='unformatted', status='replace'); write(92) a; close(92)
        rank(2)
            open(92, file=fname, access='stream', form='unformatted', status='replace'); write(92) a; close(92)
        end select
    end subroutine

    subroutine dump_step_d(dir, istep, name, a)
        character(len=*), intent(in) :: dir, name
        integer, intent(in) :: istep
        double precision, intent(in) :: a(..)
        character(len=512) :: fname
        write(fname, '(a,a,i0,a,a,a)') trim(dir), '/s', istep, '_', name, '.bin'
        call dump_d(trim(fname), a)
    end subroutine

    subroutine dump_setup(dir)
        character(len=*), intent(in) :: dir
        call dump_i(dir//'/Neighbor.bin', Neighbor(:,1:NF))
        call dump_i(dir//'/FNode.bin', FNode(:,1:NF))
        call dump_d(dir//'/VECX.bin', VECX(:,1:NF))
        call dump_i(dir//'/NeighF.bin', NeighF(1:3,1:NE))
        call dump_i(dir//'/NeighC.bin', NeighC(0:3,1:NE))
        call dump_i(dir//'/Neigh.bin', Neigh(1:NE))
        call dump_i(dir//'/Num_ghostcell.bin', Num_ghostcell(1:NE))
        call dump_i(dir//'/Ghostcell.bin', Ghostcell(:,1:NE))
        call dump_i(dir//'/Hostcell.bin', Hostcell)
        call dump_i(dir//'/BCtype_Ghostcell.bin', BCtype_Ghostcell)
        call dump_d(dir//'/vol.bin', vol(1:NE))
        call dump_d(dir//'/cellxy.bin', cellxy(:,1:NE))
        call dump_d(dir//'/reflen.bin', reflen(:,1:NE))
        call dump_d(dir//'/wdis.bin', wdis(:,1:NE))
        call dump_d(dir//'/AMLS.bin', AMLS(:,:,1:NE))
        call dump_d(dir//'/vol_Ghostcell.bin', vol_Ghostcell)
        call dump_d(dir//'/CellXY_Ghostcell.bin', CellXY_Ghostcell)
        call dump_d(dir//'/total_volume.bin', (/ total_volume, p_out /))
    end subroutine

end program
