! numerics module: cg and compute_residual with the reference's interfaces
! (MPI/src/numerics.F90:41-53, :164-175).  cg runs the whole solve on the GPU (libhpccg_hip:
! Prepare, then the iteration loop as replayed HIP graphs) and prints the reference's lines
! afterwards from the per-iteration residual history: "Initial Residual", an ITER line at
! every print_freq-th iteration and at k = max_iter-1, "Converged Solution" when the lagging
! tolerance test ended the loop, and the "Total" timer (numerics.F90:102-150).
#include "timer.h"

module numerics
    use constants
contains
    subroutine cg(A, b, x, max_iter, tolerance, niters, normr)
        use iso_c_binding
        use SparseMatrix
        use hpccg_capi
        use timer
        implicit none
        type(SparseMatrixT), intent(inout) :: A
        real(kind=dp), allocatable, intent(inout) :: b(:)
        real(kind=dp), allocatable, intent(inout) :: x(:)
        integer, intent(in) :: max_iter
        real(kind=dp), intent(in) :: tolerance
        integer, intent(out) :: niters
        real(kind=dp), intent(out) :: normr
        real(kind=dp), allocatable :: hist(:)
        integer(c_int32_t) :: nit
        integer(c_int) :: status
        integer :: k, print_freq
        type(interval) :: timertotal

        print_freq = max_iter / 10              ! numerics.F90:68-74
        if (print_freq > 50) print_freq = 50
        if (print_freq < 1) print_freq = 1

        call sparsematrix_make_resident(A)
        allocate(hist(0:max(max_iter, 0)))
        TIMER_START(0, timertotal, "Total")
        status = c_hpccg_cg(hpccg_context(), b, x, int(max_iter, c_int32_t), tolerance, nit, normr, hist)
        TIMER_STOP(0, timertotal, "Total")
        call hpccg_check(status, 'cg')
        if (status > 0) print *, 'cg: p.Ap was zero or not finite (breakdown)'
        niters = nit

        print *, "Initial Residual = ", hist(0)
        do k = 1, niters
            if (mod(k, print_freq) == 0 .or. k + 1 == max_iter) then
                print "(a,i5,tr2,a,e20.10)", 'ITER', k, 'Residual', hist(k - 1)
            end if
        end do
        if (niters < max_iter) print *, "Converged Solution"
        TIMER_PRINT(0, timertotal, "Total")
        deallocate(hist)
    end subroutine cg

    ! max_i |v1(i) - v2(i)| on the GPU (numerics.F90:164-198; single rank, no MAX allreduce);
    ! every call starts from 0 (the reference's local_residual is implicitly SAVEd, so only its
    ! first call is meaningful)
    subroutine compute_residual(n, v1, v2, residual)
        use iso_c_binding
        use hpccg_capi
        implicit none
        integer, intent(in) :: n
        real(kind=dp), allocatable, intent(inout) :: v1(:)
        real(kind=dp), allocatable, intent(inout) :: v2(:)
        real(kind=dp), intent(out) :: residual
        real(c_double) :: r
        call hpccg_check(c_hpccg_compute_residual(hpccg_context(), int(n, c_int32_t), v1, v2, r), 'compute_residual')
        residual = r
    end subroutine compute_residual
end module numerics
