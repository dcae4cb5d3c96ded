! iso_c_binding interface of libchorus_mhd.so (include/chorus_mhd.h) for the
! reference's Fortran driver.
!
! The reference's hot path is a set of argument-less subroutines over
! `module setup2d` (compresid.f90, iterations.f90:92-217,
! setinitialcond.f90:82-92).  This module keeps that shape: after INIT_SETUP
! the driver calls chorus_gpu_setup once; each iteration of the time loop calls
! chorus_gpu_step (stepper + CT on the GPU) instead of RK_S5P3/FiveStageRk and
! correct_B_via_curl_Az; chorus_gpu_write_restart replaces the restart dump
! (WRITE_ALL_DATA_PARALLEL) and chorus_gpu_plot_setup / chorus_gpu_tecplot the
! Tecplot output chain, both written from the resident state;
! chorus_gpu_fetch_state copies Q back into setup2d's array for any other use.
! The state stays resident on the GPU in between.
module chorus_hip
    use iso_c_binding
    implicit none
    private
    public :: chorus_gpu_setup, chorus_gpu_step, chorus_gpu_compresid, chorus_gpu_correct_b, &
              chorus_gpu_fetch_state, chorus_gpu_diag, chorus_gpu_free, chorus_gpu_plot_setup, &
              chorus_gpu_tecplot, chorus_gpu_write_restart, chorus_gpu_set_nodes8

    integer(c_int), parameter, public :: CHORUS_FP64 = 0, CHORUS_MIXED = 1

    type, bind(c) :: chorus_mesh_t
        integer(c_int) :: ncell, nvert, n
        type(c_ptr) :: xv, yv, ivcell
        integer(c_int) :: nface
        type(c_ptr) :: if2c
        integer(c_int) :: ninter
        type(c_ptr) :: ibfinter
        integer(c_int) :: npair
        type(c_ptr) :: cyc2fpair
        integer(c_int) :: nfree
        type(c_ptr) :: ibfree
    end type

    type, bind(c) :: chorus_params_t
        integer(c_int) :: vismode, artiv, k_stage
        real(c_double) :: mu, eta0, gam, dt, dxcycl, dycycl
        integer(c_int) :: ncellglob, precision
    end type

    ! IBFINL, IBFOUT, IBFSYMP, IBFWAL (connect_bdry.f90:186-205) and the inlet state
    type, bind(c) :: chorus_bc_t
        integer(c_int) :: ninlet
        type(c_ptr) :: ibfinl
        integer(c_int) :: noutlet
        type(c_ptr) :: ibfout
        integer(c_int) :: nsymp
        type(c_ptr) :: ibfsymp
        integer(c_int) :: nwall
        type(c_ptr) :: ibfwal
        real(c_double) :: qin(9)
    end type

    interface
        integer(c_int) function c_chorus_init_bc(mesh, params, bc, ctx) bind(c, name='chorus_init_bc')
            import :: c_int, c_ptr, chorus_mesh_t, chorus_params_t, chorus_bc_t
            type(chorus_mesh_t), intent(in) :: mesh
            type(chorus_params_t), intent(in) :: params
            type(chorus_bc_t), intent(in) :: bc
            type(c_ptr), intent(out) :: ctx
        end function
        subroutine c_chorus_free(ctx) bind(c, name='chorus_free')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end subroutine
        type(c_ptr) function c_chorus_last_error(ctx) bind(c, name='chorus_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_chorus_set_state(ctx, q) bind(c, name='chorus_set_state')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: q(*)
        end function
        integer(c_int) function c_chorus_get_state(ctx, q) bind(c, name='chorus_get_state')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: q(*)
        end function
        integer(c_int) function c_chorus_compresid(ctx, resid) bind(c, name='chorus_compresid')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: resid(*)
        end function
        integer(c_int) function c_chorus_correct_b(ctx) bind(c, name='chorus_correct_b')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function c_chorus_step(ctx, nsteps) bind(c, name='chorus_step')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int), value :: nsteps
        end function
        integer(c_int) function c_chorus_diag(ctx, residglob, divbglob, divbmax, eigvmax, eps0) &
                bind(c, name='chorus_diag')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: residglob, divbglob, divbmax, eigvmax, eps0
        end function
        integer(c_int) function c_chorus_plot_setup(ctx, mesh, nre) bind(c, name='chorus_plot_setup')
            import :: c_int, c_ptr, chorus_mesh_t
            type(c_ptr), value :: ctx
            type(chorus_mesh_t), intent(in) :: mesh
            integer(c_int), value :: nre
        end function
        integer(c_int) function c_chorus_write_tecplot(ctx, path, ctime) bind(c, name='chorus_write_tecplot')
            import :: c_int, c_ptr, c_char, c_double
            type(c_ptr), value :: ctx
            character(kind=c_char), dimension(*), intent(in) :: path
            real(c_double), value :: ctime
        end function
        integer(c_int) function c_chorus_write_restart(ctx, path, iter, ctime, icg) bind(c, name='chorus_write_restart')
            import :: c_int, c_ptr, c_char, c_double
            type(c_ptr), value :: ctx
            character(kind=c_char), dimension(*), intent(in) :: path
            integer(c_int), value :: iter
            real(c_double), value :: ctime
            integer(c_int), intent(in) :: icg(*)
        end function
        integer(c_int) function c_chorus_set_nodes8(ctx, xm, ym) bind(c, name='chorus_set_nodes8')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: xm(*), ym(*)
        end function
        integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

    type(c_ptr), save :: ctx = c_null_ptr
    type(chorus_mesh_t), save :: h_mesh
    ! host copies handed to chorus_init (c_loc needs TARGET arrays)
    real(c_double), allocatable, target, save :: h_xv(:), h_yv(:)
    integer(c_int), allocatable, target, save :: h_iv(:,:), h_if2c(:,:), h_inter(:), h_cyc(:,:), h_free(:)
    integer(c_int), allocatable, target, save :: h_inl(:), h_out(:), h_symp(:), h_wal(:)

contains

    subroutine check(rc, what)
        integer(c_int), intent(in) :: rc
        character(*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (rc /= 0) then
            n = int(c_strlen(c_chorus_last_error(ctx)))
            call c_f_pointer(c_chorus_last_error(ctx), msg, [n])
            print *, 'chorus: ', what, ' failed (', rc, '): ', msg(1:n)
            stop
        end if
    end subroutine

    ! INIT_SETUP done: hand setup2d's mesh, connectivity, boundary lists,
    ! parameters and Q to the GPU.  qin: the state of INLE faces (the reference
    ! reads none; needed only when the mesh has INLE faces)
    subroutine chorus_gpu_setup(gam, precision, qin)
        use setup2d, only: NCELL, NVERT, N, XV, YV, IVCELL, NFACE, IF2C, NINTER, IBFINTER, NPAIR, &
                           CYC2FPAIR, NFREE, IBFREE, vismode, ARTIV, k_stage, mu, eta0, dt, DXCYCL, &
                           DYCYCL, NCELLGLOB, Q, NINLET, NOUTLET, NSYMP, NWALL, IBFINL, IBFOUT, &
                           IBFSYMP, IBFWAL
        real(c_double), intent(in) :: gam
        integer(c_int), intent(in) :: precision
        real(c_double), intent(in), optional :: qin(9)
        type(chorus_mesh_t) :: m
        type(chorus_params_t) :: p
        type(chorus_bc_t) :: b
        h_xv = XV(1:NVERT)
        h_yv = YV(1:NVERT)
        h_iv = IVCELL(1:NCELL, 1:4)
        h_if2c = IF2C(1:NFACE, 1:4)
        h_inter = IBFINTER(1:NINTER)
        allocate(h_cyc(max(NPAIR, 1), 2), h_free(max(NFREE, 1)))
        if (NPAIR > 0) h_cyc(1:NPAIR, :) = CYC2FPAIR(1:NPAIR, :)
        if (NFREE > 0) h_free(1:NFREE) = IBFREE(1:NFREE)
        m%ncell = NCELL; m%nvert = NVERT; m%n = N
        m%xv = c_loc(h_xv); m%yv = c_loc(h_yv); m%ivcell = c_loc(h_iv)
        m%nface = NFACE; m%if2c = c_loc(h_if2c)
        m%ninter = NINTER; m%ibfinter = c_loc(h_inter)
        m%npair = NPAIR; m%cyc2fpair = c_loc(h_cyc)
        m%nfree = NFREE; m%ibfree = c_loc(h_free)
        p%vismode = vismode; p%artiv = ARTIV; p%k_stage = k_stage
        p%mu = mu; p%eta0 = eta0; p%gam = gam; p%dt = dt
        p%dxcycl = DXCYCL; p%dycycl = DYCYCL; p%ncellglob = NCELLGLOB; p%precision = precision
        allocate(h_inl(max(NINLET, 1)), h_out(max(NOUTLET, 1)), h_symp(max(NSYMP, 1)), h_wal(max(NWALL, 1)))
        if (NINLET > 0) h_inl(1:NINLET) = IBFINL(1:NINLET)
        if (NOUTLET > 0) h_out(1:NOUTLET) = IBFOUT(1:NOUTLET)
        if (NSYMP > 0) h_symp(1:NSYMP) = IBFSYMP(1:NSYMP)
        if (NWALL > 0) h_wal(1:NWALL) = IBFWAL(1:NWALL)
        b%ninlet = NINLET; b%ibfinl = c_loc(h_inl)
        b%noutlet = NOUTLET; b%ibfout = c_loc(h_out)
        b%nsymp = NSYMP; b%ibfsymp = c_loc(h_symp)
        b%nwall = NWALL; b%ibfwal = c_loc(h_wal)
        b%qin = 0.0_c_double
        if (present(qin)) b%qin = qin
        call check(c_chorus_init_bc(m, p, b, ctx), 'chorus_init_bc')
        call check(c_chorus_set_state(ctx, Q), 'chorus_set_state')
        h_mesh = m
    end subroutine

    ! CURVE_WALL = 1 (initsetup.f90:71-78, where the reference stops): the edge
    ! nodes 5..8 of every cell, XM(NCELL,4) and YM(NCELL,4), column k on edge k
    ! (1: vertices 1-2, 2: 2-3, 3: 3-4, 4: 4-1), for the 8-node map of
    ! MAPDER_8_NODE (mapder.f90:117-253); after chorus_gpu_setup, before stepping
    subroutine chorus_gpu_set_nodes8(xm, ym)
        real(c_double), intent(in) :: xm(:,:), ym(:,:)
        real(c_double), allocatable :: x(:,:), y(:,:)
        x = xm
        y = ym
        call check(c_chorus_set_nodes8(ctx, x, y), 'chorus_set_nodes8')
    end subroutine

    ! tecplotter2dsetup (tecplotter2dsetup.f90:1-164) on the mesh of chorus_gpu_setup
    subroutine chorus_gpu_plot_setup(nre)
        integer, intent(in) :: nre
        call check(c_chorus_plot_setup(ctx, h_mesh, int(nre, c_int)), 'chorus_plot_setup')
    end subroutine

    ! tecplotter2d + tecpost_parallel (tecplotter.f90:1-418): ./TEC/TEC<time>.plt
    ! of the resident state, named as tecplotter.f90:190-191 does
    subroutine chorus_gpu_tecplot(ctime)
        double precision, intent(in) :: ctime
        character(6) :: decimal
        character(64) :: filename
        write(decimal, '(f6.3)') ctime - int(ctime)
        write(filename, '(a,i7.7,a4,a)') 'TEC', int(ctime), decimal(3:6), '.plt'
        call check(c_chorus_write_tecplot(ctx, './TEC/'//trim(filename)//c_null_char, real(ctime, c_double)), &
                   'chorus_write_tecplot')
    end subroutine

    ! WRITE_ALL_DATA_PARALLEL (write_data.f90:5-83) of the resident state:
    ! restart.<iter:8>.dat with setup2d's iter, ctime and global cell numbers
    subroutine chorus_gpu_write_restart()
        use setup2d, only: NCELL, IC2ICG, iter, ctime
        character(64) :: filename
        integer(c_int), allocatable :: icg(:)
        icg = IC2ICG(1:NCELL)
        write(filename, '(A,I8.8,A)') 'restart.', iter, '.dat'
        call check(c_chorus_write_restart(ctx, trim(filename)//c_null_char, int(iter, c_int), &
                   real(ctime, c_double), icg), 'chorus_write_restart')
    end subroutine

    ! loop body of iterations.f90:22-33 on the GPU: RK_S5P3 / FiveStageRk + correct_B_via_curl_Az
    subroutine chorus_gpu_step(nsteps)
        integer, intent(in) :: nsteps
        call check(c_chorus_step(ctx, int(nsteps, c_int)), 'chorus_step')
    end subroutine

    ! compresid (compresid.f90:1-67) of the resident state, into setup2d's resid
    subroutine chorus_gpu_compresid()
        use setup2d, only: resid
        call check(c_chorus_compresid(ctx, resid), 'chorus_compresid')
    end subroutine

    subroutine chorus_gpu_correct_b()
        call check(c_chorus_correct_b(ctx), 'chorus_correct_b')
    end subroutine

    ! copy the resident state into setup2d's Q
    subroutine chorus_gpu_fetch_state()
        use setup2d, only: Q
        call check(c_chorus_get_state(ctx, Q), 'chorus_get_state')
    end subroutine

    ! residual and div B norms of iterations.f90:66-67, plus max |div B|, eigvmax, epsilon0
    subroutine chorus_gpu_diag(residglob, divbglob, divbmax, eigvmax, eps0)
        real(c_double), intent(out) :: residglob, divbglob, divbmax, eigvmax, eps0
        call check(c_chorus_diag(ctx, residglob, divbglob, divbmax, eigvmax, eps0), 'chorus_diag')
    end subroutine

    subroutine chorus_gpu_free()
        call c_chorus_free(ctx)
        ctx = c_null_ptr
    end subroutine

end module chorus_hip
