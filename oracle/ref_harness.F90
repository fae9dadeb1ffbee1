! Oracle harness (test infrastructure only).
!
! A driver of our own that links against the reference's unmodified modules
! (MPI/src/{constants,SparseMatrix,vector_operations,numerics,timer}.F90) to extract
! full-precision values the reference's own main (MPI/src/main.F90) never prints:
!
!   hist    nx ny nz k0 k1 [bfile]   ||r_{K-1}|| for K=k0..k1 from cg(max_iter=K, tol=0),
!                                    each a fresh solve from x0=0 (cg returns normr =
!                                    sqrt(rtrans) of the last iteration's top, numerics.F90:122),
!                                    plus ||x-x*|| after the K=k1 solve
!   solve   nx ny nz maxit tol bfile xfile   cg to an absolute tolerance on b read from bfile
!   kernels nx ny nz xfile yfile alpha beta outfile   spmv/ddot/ddot1/waxpby/wwpby/wxbw/vcopy
!   gen     nx ny nz prefix          dump generate_matrix's CSR (1-based), b, x, xexact
!
! Binary files are raw little-endian float64 (int32 for indices), no record markers.
program ref_harness
    use constants
    use SparseMatrix
    use vector_operations
    use numerics
    use timer
    implicit none
    character(len=256) :: mode, arg, f1, f2, f3
    type(SparseMatrixT) :: A
    real(kind=dp), allocatable :: x(:), b(:), xexact(:), y(:), w(:), xs(:)
    integer :: nx, ny, nz, k0, k1, k, niters, n, maxit, nnz, i
    real(kind=dp) :: normr, tol, alpha, beta, d1, d2, emax, e2, diff
    logical :: have_b

    call get_command_argument(1, mode)
    call get_command_argument(2, arg); read (arg, *) nx
    call get_command_argument(3, arg); read (arg, *) ny
    call get_command_argument(4, arg); read (arg, *) nz
    call timer_init()
    call generate_matrix(nx, ny, nz, A, x, b, xexact)
    n = A%local_nrow
    nnz = A%rowindex(n+1) - 1

    select case (trim(mode))
    case ('hist')
        call get_command_argument(5, arg); read (arg, *) k0
        call get_command_argument(6, arg); read (arg, *) k1
        have_b = command_argument_count() >= 7
        if (have_b) then
            call get_command_argument(7, f1)
            call read_vec(f1, b, n)
            call get_command_argument(8, f2)
            allocate(xs(n))
            call read_vec(f2, xs, n)
        else
            allocate(xs(n))
            xs = xexact
        end if
        do k = k0, k1
            x = 0.0d0
            niters = 0
            normr = 0.0d0
            call cg(A, b, x, k, 0.0d0, niters, normr)
            write (*, '(a,i6,1x,i6,1x,es25.17)') 'H NORMR ', k, niters, normr
        end do
        call xerr(x, xs, n, emax, e2)
        write (*, '(a,i6,1x,es25.17,1x,es25.17)') 'H XERR ', k1, emax, e2
    case ('solve')
        call get_command_argument(5, arg); read (arg, *) maxit
        call get_command_argument(6, arg); read (arg, *) tol
        call get_command_argument(7, f1)
        call get_command_argument(8, f2)
        call read_vec(f1, b, n)
        x = 0.0d0
        niters = 0
        normr = 0.0d0
        call cg(A, b, x, maxit, tol, niters, normr)
        call write_vec(f2, x, n)
        write (*, '(a,i6,1x,es25.17)') 'H SOLVE ', niters, normr
    case ('kernels')
        call get_command_argument(5, f1)
        call get_command_argument(6, f2)
        call get_command_argument(7, arg); read (arg, *) alpha
        call get_command_argument(8, arg); read (arg, *) beta
        call get_command_argument(9, f3)
        allocate(y(n), w(n))
        call read_vec(f1, x, n)
        call read_vec(f2, y, n)
        open (unit=21, file=trim(f3), access='stream', form='unformatted', status='replace')
        call spmv(A, x, w)
        write (21) w
        call waxpby(n, alpha, x, beta, y, w)
        write (21) w
        w = y
        call wwpby(n, beta, x, w)
        write (21) w
        w = y
        call wxbw(n, x, beta, w)
        write (21) w
        call vcopy(n, x, w)
        write (21) w
        close (21)
        d1 = 0.0d0
        d2 = 0.0d0
        call ddot(n, x, y, d1)
        call ddot1(n, x, d2)
        write (*, '(a,es25.17,1x,es25.17)') 'H DOTS ', d1, d2
    case ('gen')
        call get_command_argument(5, f1)
        open (unit=21, file=trim(f1)//'.rowindex', access='stream', form='unformatted', status='replace')
        write (21) A%rowindex(1:n+1)
        close (21)
        open (unit=21, file=trim(f1)//'.colindex', access='stream', form='unformatted', status='replace')
        write (21) A%colindex(1:nnz)
        close (21)
        open (unit=21, file=trim(f1)//'.values', access='stream', form='unformatted', status='replace')
        write (21) A%values(1:nnz)
        close (21)
        call write_vec(trim(f1)//'.b', b, n)
        call write_vec(trim(f1)//'.x', x, n)
        call write_vec(trim(f1)//'.xexact', xexact, n)
        write (*, '(a,i12,1x,i12,1x,i12,1x,i12)') 'H GEN ', n, nnz, A%total_nrow, A%total_nnz
    case default
        print *, 'unknown mode'
        stop 1
    end select

contains

    subroutine read_vec(fname, v, len)
        character(len=*), intent(in) :: fname
        integer, intent(in) :: len
        real(kind=dp), intent(inout) :: v(len)
        open (unit=20, file=trim(fname), access='stream', form='unformatted', status='old')
        read (20) v
        close (20)
    end subroutine read_vec

    subroutine write_vec(fname, v, len)
        character(len=*), intent(in) :: fname
        integer, intent(in) :: len
        real(kind=dp), intent(in) :: v(len)
        open (unit=20, file=trim(fname), access='stream', form='unformatted', status='replace')
        write (20) v
        close (20)
    end subroutine write_vec

    ! max-norm as compute_residual (numerics.F90:164-198) defines it, and a sequential 2-norm
    subroutine xerr(v, ref, len, emax, e2)
        integer, intent(in) :: len
        real(kind=dp), intent(in) :: v(len), ref(len)
        real(kind=dp), intent(out) :: emax, e2
        integer :: j
        real(kind=dp) :: d
        emax = 0.0d0
        e2 = 0.0d0
        do j = 1, len
            d = abs(v(j) - ref(j))
            if (d > emax) emax = d
            e2 = e2 + d*d
        end do
        e2 = sqrt(e2)
    end subroutine xerr
end program ref_harness
