! Drop-in driver: the reference's own set-up chain and output routines with the
! time-loop body on the MI355X through chorus_hip (iso_c_binding).
!
! Program flow follows main.f90:8-65 and iterations.f90:1-89: read QUAD.INP and
! the mesh, build connectivity, INIT_SETUP (initial condition or restart), then
! per iteration one GPU step (stepper + CT), the residual / div B norms every 10
! iterations into resid.out, and at mod(iter,nwrite)==1 and at MAXITER the
! Tecplot file (the library's device-side replacement of tecplotter2dsetup,
! tecplotter2d and tecpost_parallel, into ./TEC/, which must exist as for the
! reference; with ARTIV=0 its muAV/etaAV columns are zeros, where the
! reference reads the unallocated muAV, tecplotter.f90:37-38) and the restart
! dump (the library's device-side WRITE_ALL_DATA_PARALLEL; with "refio" the
! state is copied back and the reference's own writer runs).
!
! Curved cells: where the reference stops (CURVE_WALL=1, initsetup.f90:71-78),
! a file nodes8.dat in the run directory gives every cell's edge nodes 5..8
! (NCELLGLOB records by global cell number, each the 4 x then the 4 y, edge k
! between vertices k and k+1); the set-up then runs with CURVE_WALL=0 and the
! library switches to the 8-node map of MAPDER_8_NODE.  No Tecplot output
! then (the reference's plot set-up has no 8-node branch either).
!
! usage: sd2d_mhd_hip [mixed] [refio]   (FP64 drop-in unless "mixed" is given)
program sd2d_mhd_hip
    use setup2d
    use chorus_hip
    use iso_c_binding, only: c_double, c_int
    implicit none
    include 'mpif.h'

    character(16) :: arg
    integer(c_int) :: prec
    logical :: refio, curved
    integer :: ia, icl
    real(c_double), allocatable :: rows8(:,:), xm8(:,:), ym8(:,:)
    real(c_double) :: residglob, divbglob, divbmax, eigv, eps0
    double precision :: t_start, t_finish
    integer :: fh

    call MPI_INIT(ierror)
    call MPI_COMM_SIZE(MPI_COMM_WORLD, size, ierror)
    call MPI_COMM_RANK(MPI_COMM_WORLD, rank, ierror)
    NPROC = size
    if (NPROC /= 1) stop 'sd2d_mhd_hip runs one rank: the GPU holds the whole mesh'
    prec = CHORUS_FP64
    refio = .false.
    do ia = 1, command_argument_count()
        call get_command_argument(ia, arg)
        if (trim(arg) == 'mixed') prec = CHORUS_MIXED
        if (trim(arg) == 'refio') refio = .true.
    end do

    t_start = MPI_WTIME()
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

    call chorus_gpu_setup(gam, prec)
    inquire(file='nodes8.dat', exist=curved)
    if (curved) then
        allocate(rows8(8, NCELLGLOB), xm8(NCELL, 4), ym8(NCELL, 4))
        open(newunit=fh, file='nodes8.dat', status='old')
        read(fh, *) rows8
        close(fh)
        do icl = 1, NCELL
            xm8(icl, :) = rows8(1:4, IC2ICG(icl))
            ym8(icl, :) = rows8(5:8, IC2ICG(icl))
        end do
        call chorus_gpu_set_nodes8(xm8, ym8)
    else
        call chorus_gpu_plot_setup(NRE)
    end if

    open(newunit=fh, file='resid.out', status='replace')
    write(fh, *) 'ITER    ctime     RESIDNORM   divB'
    do while (iter < MAXITER)
        iter = iter + 1
        call chorus_gpu_step(1)
        ctime = ctime + dt
        if (mod(iter, 10) == 1) then
            call chorus_gpu_diag(residglob, divbglob, divbmax, eigv, eps0)
            write(fh, *) iter, ctime, residglob, divbglob
        end if
        if (mod(iter, nwrite) == 1 .or. iter >= MAXITER) then
            if (.not. curved) call chorus_gpu_tecplot(ctime)
            if (refio) then
                call chorus_gpu_fetch_state()
                call WRITE_ALL_DATA_PARALLEL
            else
                call chorus_gpu_write_restart()
            end if
        end if
    end do
    close(fh)
    call chorus_gpu_free()
    t_finish = MPI_WTIME()
    print *, 'Elapsed time ', t_finish - t_start, ' seconds'
    call MPI_FINALIZE(ierror)
end program sd2d_mhd_hip
