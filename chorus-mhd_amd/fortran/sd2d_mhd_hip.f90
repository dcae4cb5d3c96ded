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
! usage: sd2d_mhd_hip [mixed] [refio]   (FP64 drop-in unless "mixed" is given)
program sd2d_mhd_hip
    use setup2d
    use chorus_hip
    use iso_c_binding, only: c_double, c_int
    implicit none
    include 'mpif.h'

    character(16) :: arg
    integer(c_int) :: prec
    logical :: refio
    integer :: ia
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
    call chorus_gpu_plot_setup(NRE)

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
            call chorus_gpu_tecplot(ctime)
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
