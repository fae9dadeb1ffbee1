! SparseMatrix module: the reference's matrix type and procedures (MPI/src/SparseMatrix.F90),
! with the CSR kept in device memory by libhpccg_hip.
!
!   SparseMatrixT    same components as SparseMatrix.F90:44-64 (single rank: no halo lists),
!                    plus dev_tag, which tells whether this matrix is the one resident on the GPU
!   generate_matrix  SparseMatrix.F90:68-185; built on the GPU (same bytes), copied to the
!                    host arrays the caller gets back, and left resident for spmv / cg
!   spmv             SparseMatrix.F90:213-232; y = A*x on the GPU (uploads A on first use)
!   dump_matrix      SparseMatrix.F90:187-211 (debug print)
module SparseMatrix
    use constants
    use iso_c_binding
    use hpccg_capi
    implicit none
    integer, public :: ierror
    integer, public :: single = 1  ! one rank: the whole grid is this process's subdomain

    type, public :: SparseMatrixT
        integer :: start_row
        integer :: stop_row
        integer :: total_nrow
        integer(kind=4) :: total_nnz
        integer :: local_nrow
        integer :: local_ncol
        integer :: local_nnz
        integer, allocatable :: colindex(:)
        integer, allocatable :: rowindex(:)
        real(kind=dp), allocatable :: values(:)
        integer :: dev_tag = 0
    end type SparseMatrixT

    public :: generate_matrix, dump_matrix, spmv, sparsematrix_make_resident

contains

    subroutine generate_matrix(nx, ny, nz, A, x, b, xexact)
        integer, intent(in) :: nx, ny, nz
        type(SparseMatrixT), intent(out) :: A
        real(kind=dp), allocatable, intent(out) :: x(:)
        real(kind=dp), allocatable, intent(out) :: b(:)
        real(kind=dp), allocatable, intent(out) :: xexact(:)
        integer(kind=8) :: n8, nnz_alloc
        integer(c_int64_t) :: nnz

        n8 = int(nx, 8) * int(ny, 8) * int(nz, 8)
        nnz_alloc = 27_8 * n8
        if (n8 < 1 .or. nnz_alloc > int(huge(1), 8)) then
            print *, 'generate_matrix: grid too large for default-integer CSR (27*N must fit)'
            error stop 1
        end if
        A%local_nrow = int(n8)
        A%local_nnz = int(nnz_alloc)       ! slots allocated, as SparseMatrix.F90:108-109
        A%total_nrow = A%local_nrow
        A%total_nnz = 27 * A%total_nrow    ! nominal, as SparseMatrix.F90:111
        A%start_row = 1
        A%stop_row = A%local_nrow
        A%local_ncol = A%local_nrow
        allocate(A%values(A%local_nnz))
        allocate(A%rowindex(A%local_nrow + 1))
        allocate(A%colindex(A%local_nnz))
        allocate(x(A%local_nrow))
        allocate(b(A%local_nrow))
        allocate(xexact(A%local_nrow))
        call hpccg_check(c_hpccg_generate_matrix(hpccg_context(), int(nx, c_int32_t), int(ny, c_int32_t), &
            int(nz, c_int32_t), A%rowindex, A%colindex, A%values, x, b, xexact, nnz), 'generate_matrix')
        ! slots past the true nnz stay unused (the reference leaves them unset)
        A%values(nnz + 1:) = 0.0_dp
        A%colindex(nnz + 1:) = 0
        A%dev_tag = hpccg_new_tag()
        call hpccg_set_resident(A%dev_tag)
    end subroutine generate_matrix

    ! Make A the matrix resident on the GPU (uploads its CSR if another one is resident).
    subroutine sparsematrix_make_resident(A)
        type(SparseMatrixT), intent(inout) :: A
        integer :: nnz
        if (A%dev_tag /= 0 .and. A%dev_tag == hpccg_resident_tag()) return
        nnz = A%rowindex(A%local_nrow + 1) - 1
        call hpccg_check(c_hpccg_matrix_upload(hpccg_context(), int(A%local_nrow, c_int32_t), &
            int(A%local_ncol, c_int32_t), A%rowindex, A%colindex, A%values), 'matrix_upload')
        if (A%dev_tag == 0) A%dev_tag = hpccg_new_tag()
        call hpccg_set_resident(A%dev_tag)
    end subroutine sparsematrix_make_resident

    subroutine dump_matrix(nx, ny, nz, A, x, b)
        integer, intent(in) :: nx, ny, nz
        type(SparseMatrixT), intent(in) :: A
        real(kind=dp), allocatable, intent(in) :: x(:)
        real(kind=dp), allocatable, intent(in) :: b(:)
        integer :: i, k
        print *, 'Matrix has ', A%total_nrow, ' total rows and ', A%local_ncol, ' columns.'
        print *, 'Matrix has ', A%local_nrow, ' local rows.'
        do i = 1, A%local_nrow
            print *, 'Row id', i, 'Start', A%rowindex(i), 'Stop', A%rowindex(i + 1) - 1, &
                'Nonzeroes', A%rowindex(i + 1) - A%rowindex(i)
            do k = A%rowindex(i), A%rowindex(i + 1) - 1
                print '(i0,a,f4.1)', A%colindex(k), ' : ', A%values(k)
            end do
        end do
    end subroutine dump_matrix

    subroutine spmv(A, x, y)
        type(SparseMatrixT), intent(in) :: A
        real(kind=dp), allocatable, intent(in) :: x(:)
        real(kind=dp), allocatable, intent(inout) :: y(:)
        call make_resident_const(A)
        call hpccg_check(c_hpccg_spmv(hpccg_context(), x, y), 'spmv')
    end subroutine spmv

    ! spmv receives A as intent(in) (SparseMatrix.F90:215); a matrix that was never on the
    ! device is uploaded without recording a tag, so it is uploaded again on its next use.
    subroutine make_resident_const(A)
        type(SparseMatrixT), intent(in) :: A
        integer :: nnz
        if (A%dev_tag /= 0 .and. A%dev_tag == hpccg_resident_tag()) return
        nnz = A%rowindex(A%local_nrow + 1) - 1
        call hpccg_check(c_hpccg_matrix_upload(hpccg_context(), int(A%local_nrow, c_int32_t), &
            int(A%local_ncol, c_int32_t), A%rowindex, A%colindex, A%values), 'matrix_upload')
        call hpccg_set_resident(A%dev_tag)
    end subroutine make_resident_const
end module SparseMatrix
