! Test-infrastructure harness around the UNMODIFIED reference Fortran sources.
!
! Not part of the product.  It replaces only the reference's program unit
! (main.f90) so the reference routines can be driven without the Tecplot
! output path, which reads the unallocated muAV array when ARTIV=0
! (tecplotter.f90:37-38).  The set-up chain mirrors main.f90:16-44 and the
! time loop mirrors the body of iterations.f90:19-57 (stepper, CT correction,
! diagnostic flux + gradient, residual/divB norms), calling the reference's own
! subroutines.  Outputs are raw little-endian stream files in the current
! directory:
!
!   mode "setup": topo.i32, geom.f64, ops.f64, q0.f64
!   mode "resid": q0.f64, resid.f64 (+ av.f64 when ARTIV=1)
!   mode "run"  : q0.f64, qend.f64, diag.txt, timing.txt
!   mode "runnd": as "run" without the diagnostic flux and gradient that
!                 iterations.f90:35-36 evaluate every iteration (used only
!                 every 10th): timing of the loop without that work
!   mode "mapder8": mapder8.f64, the 8-node shape-function derivative tables
!                 of MAPDER_8_NODE (mapder.f90:117-253) for the run's N:
!                 dmdxs(8,2,N,N), dmdxf1(8,2,N+1,N), dmdxf2(8,2,N,N+1) (the
!                 reference's INIT_SETUP stops right after computing them,
!                 initsetup.f90:71-78, so they are taken here after a
!                 CURVE_WALL=0 set-up)
!   mode "plot" : as "run", then the reference's own Tecplot output chain
!                 (tecplotter2dsetup, tecplotter2d, tecpost_parallel, the
!                 calls of iterations.f90:82-86) writes ./TEC/TEC*.plt and the
!                 chain's timing to plot_timing.txt; needs ARTIV=1
!                 (tecplotter2d reads muAV) and an existing ./TEC
!
! Under mpirun (NPROC > 1, METIS_CELL in QUAD.INP) only rank 0 writes files,
! and the loop is timed between barriers: timing.txt holds the steps, the
! seconds, NCELLGLOB and N.
!
! usage:  ref_harness setup|resid|run|runnd|mapder8|plot   (reads QUAD.INP like the reference)
program ref_harness
    use setup2d
    implicit none
    include 'mpif.h'

    character(16) :: mode
    integer :: fh, k, ic, is, js, nsteps
    double precision :: t0, t1, residnorm, divB, residglob, divBglob

    call MPI_INIT(ierror)
    call MPI_COMM_SIZE(MPI_COMM_WORLD, size, ierror)
    call MPI_COMM_RANK(MPI_COMM_WORLD, rank, ierror)
    NPROC = size

    call get_command_argument(1, mode)

    ! set-up chain of main.f90:16-44
    call readinput
    call READ_CELL_DATA
    call READ_VRT_DATA
    call READ_METIS
    call GLOB2LOCAL
    call READ_BND_DATA
    call CHECK_MESH
    call CONNECTIVITY
    call CONNECT_BDRY
    call MATCHPROC
    call MATCHCYCREM
    call GETIVCELL_PROC
    call GETIVCELL_CYC
    call MAPCYCLOC
    call INIT_SETUP

    if (rank == 0) then
        open(newunit=fh, file='q0.f64', access='stream', form='unformatted', status='replace')
        write(fh) Q
        close(fh)
    end if

    if (trim(mode) == 'setup') then
        open(newunit=fh, file='topo.i32', access='stream', form='unformatted', status='replace')
        write(fh) NCELL, NVERT, NFACE, NINTER, NPAIR, NFREE
        write(fh) IVCELL(1:NCELL,1:4)
        write(fh) IF2C(1:NFACE,1:4)
        write(fh) IBFINTER(1:NINTER)
        if (NPAIR > 0) write(fh) CYC2FPAIR(1:NPAIR,1:2)
        if (NFREE > 0) write(fh) IBFREE(1:NFREE)
        close(fh)
        open(newunit=fh, file='geom.f64', access='stream', form='unformatted', status='replace')
        write(fh) XV(1:NVERT), YV(1:NVERT)
        write(fh) Jac, S1, S2, XXsolu
        close(fh)
        open(newunit=fh, file='ops.f64', access='stream', form='unformatted', status='replace')
        write(fh) Xs, Xf, Lmat, Mmat
        close(fh)

    else if (trim(mode) == 'mapder8') then
        deallocate(dmdxs, dmdxf1, dmdxf2)
        allocate(dmdxs(8,2,N,N), dmdxf1(8,2,N+1,N), dmdxf2(8,2,N,N+1))
        call MAPDER_8_NODE
        open(newunit=fh, file='mapder8.f64', access='stream', form='unformatted', status='replace')
        write(fh) dmdxs, dmdxf1, dmdxf2
        close(fh)

    else if (trim(mode) == 'resid') then
        call compresid
        open(newunit=fh, file='resid.f64', access='stream', form='unformatted', status='replace')
        write(fh) resid
        close(fh)
        if (ARTIV == 1) then
            open(newunit=fh, file='av.f64', access='stream', form='unformatted', status='replace')
            write(fh) eigvmax, eigvmax_g, epsilon0
            write(fh) SM_IND, muAV, muAVfi, muAVfj
            close(fh)
        end if

    else if (trim(mode) == 'run' .or. trim(mode) == 'runnd' .or. trim(mode) == 'plot') then
        nsteps = MAXITER - iter
        if (rank == 0) open(newunit=fh, file='diag.txt', status='replace')
        call MPI_BARRIER(MPI_COMM_WORLD, ierror)
        t0 = MPI_WTIME()
        do while (iter < MAXITER)
            ! loop body of iterations.f90:20-57
            iter = iter + 1
            if (k_stage == 5) then
                call FiveStageRk
            else if (k_stage == 3) then
                call RK_S5P3
            else
                stop 'k_stage must be 3 or 5'
            end if
            call correct_B_via_curl_Az
            if (trim(mode) /= 'runnd') then
                call comp_inviscid_flux
                call nablaQsp(1,9)
            end if
            if (mod(iter,10) == 1 .and. trim(mode) /= 'runnd' .and. rank == 0) then
                residnorm = 0.d0
                divB = 0.d0
                do ic = 1,NCELL
                    do js = 1,N
                    do is = 1,N
                    do k = 1,9
                        residnorm = residnorm + (resid(k,is,js,ic))**2
                    end do
                    divB = divB + (nablaQs(6,1,is,js,ic)+nablaQs(7,2,is,js,ic))**2
                    end do
                    end do
                end do
                residglob = sqrt(residnorm/(NCELLGLOB*N*N*5))
                divBglob = sqrt(divB/(NCELLGLOB*N*N))
                write(fh, '(I8,3ES26.17E3)') iter, ctime, residglob, divBglob
            end if
        end do
        call MPI_BARRIER(MPI_COMM_WORLD, ierror)
        t1 = MPI_WTIME()
        if (rank == 0) then
            close(fh)
            open(newunit=fh, file='qend.f64', access='stream', form='unformatted', status='replace')
            write(fh) Q
            close(fh)
            open(newunit=fh, file='timing.txt', status='replace')
            write(fh, '(I8,ES26.17E3,I12,I4)') nsteps, t1 - t0, NCELLGLOB, N
            close(fh)
        end if
        if (trim(mode) == 'plot') then
            t0 = MPI_WTIME()
            call tecplotter2dsetup
            t1 = MPI_WTIME()
            call tecplotter2d
            call tecpost_parallel
            ! plot_timing.txt: seconds of the plot-node set-up, seconds of one output
            open(newunit=fh, file='plot_timing.txt', status='replace')
            write(fh, '(2ES26.17E3)') t1 - t0, MPI_WTIME() - t1
            close(fh)
        end if
    else
        stop 'mode must be setup, resid, run, runnd, mapder8 or plot'
    end if

    call MPI_FINALIZE(ierror)
end program ref_harness
