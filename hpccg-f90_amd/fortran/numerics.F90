! numerics module: cg and compute_residual with the reference's interfaces
! (MPI/src/numerics.F90:41-53, :164-175).  cg runs the whole solve on the GPU (libhpccg_hip)
! and prints the reference's lines afterwards, in the reference's order, from the
! per-iteration residual history and the device's timer regions (numerics.F90:91-156):
!   TIMINGLEV > 1: "TIMER: Prepare" before "Initial Residual", and "TIMER: Iteration" after
!                  every iteration (iterations are then launched one at a time, each timed
!                  with HIP events);
!   always:        "Initial Residual", an ITER line at every print_freq-th iteration and at
!                  k = max_iter-1, "Converged Solution" when the lagging tolerance test ended
!                  the loop;
!   TIMINGLEV > 0: "TIMER: Total", the iteration loop only (numerics.F90:106,149);
!   TIMINGLEV > 2: "MFLOPS SPARSEMV" = niters * 2 * total_nnz / (sum of the SpMV times) / 1e6
!                  with the nominal total_nnz = 27 N (numerics.F90:152-156).
! The region times are device times measured by the library; TIMER_PRINT reports them through
! the timer module like any other interval.  roctx ranges named like the regions bracket the
! same device work (visible under rocprofv3 --marker-trace).
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
        real(kind=dp), allocatable :: hist(:), t_iter(:), t_spmv(:)
        real(c_double) :: t_total, t_prepare
        integer(c_int32_t) :: nit, level
        integer(c_int) :: status
        integer :: k, print_freq
        type(interval) :: timerdata, timertotal
        ! the reference initialises it in its declaration (implicit SAVE, numerics.F90:66): the
        ! SpMV time accumulates over every cg call of the process
        real(kind=dp), save :: total_time_spmv = 0.0d0
        real(kind=dp) :: fnops_sparsemv

        print_freq = max_iter / 10              ! numerics.F90:68-74
        if (print_freq > 50) print_freq = 50
        if (print_freq < 1) print_freq = 1

        call sparsematrix_make_resident(A)
        allocate(hist(0:max(max_iter, 0)), t_iter(max(max_iter, 1)), t_spmv(max(max_iter, 1)))
        t_iter = 0.0d0
        t_spmv = 0.0d0
        level = max(1, TIMINGLEV)
        status = c_hpccg_cg_timed(hpccg_context(), b, x, int(max_iter, c_int32_t), tolerance, level, nit, normr, &
                                  hist, t_total, t_prepare, t_iter, t_spmv)
        call hpccg_check(status, 'cg')
        if (status > 0) print *, 'cg: p.Ap was zero or not finite (breakdown)'
        niters = nit

        call device_interval(timerdata, t_prepare)
        TIMER_PRINT(1, timerdata, "Prepare")
        print *, "Initial Residual = ", hist(0)
        do k = 1, niters
            if (mod(k, print_freq) == 0 .or. k + 1 == max_iter) then
                print "(a,i5,tr2,a,e20.10)", 'ITER', k, 'Residual', hist(k - 1)
            end if
            ! timer_print returns a default REAL, added to the double accumulator (numerics.F90:135)
            call device_interval(timerdata, t_spmv(k))
            total_time_spmv = total_time_spmv + timer_print(timerdata)
            call device_interval(timerdata, t_iter(k))
            TIMER_PRINT(1, timerdata, "Iteration")
        end do
        if (niters < max_iter) print *, "Converged Solution"
        call device_interval(timertotal, t_total)
        TIMER_PRINT(0, timertotal, "Total")
        if (TIMINGLEV > 2) then
            fnops_sparsemv = niters * 2.0d0 * A%total_nnz
            print *, "MFLOPS SPARSEMV", fnops_sparsemv / total_time_spmv / 1.0d6
        end if
        deallocate(hist, t_iter, t_spmv)
    end subroutine cg

    ! an interval spanning `seconds` of device time
    subroutine device_interval(timing, seconds)
        use timer
        implicit none
        type(interval), intent(out) :: timing
        real(kind=dp), intent(in) :: seconds
        timing%startt = 0.0d0
        timing%stopt = seconds
    end subroutine device_interval

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
