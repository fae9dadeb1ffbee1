! ISO_C_BINDING layer over libhpccg_hip (include/hpccg_hip.h).
!
! Owns the process-wide context handle (one device, one HIP stream), created on first use:
!   HPCCG_DEVICE    device ordinal (default 0)
!   HPCCG_DOT_MODE  "tree" (default: deterministic tree dots, fast) or "sequential" (the
!                   reference's single left-to-right chain: cg reproduces its iterates bit for
!                   bit, numerics.F90:41-162)
! A nonzero (negative) status from the library stops the program with its message, the way
! the reference stops on fatal errors (main.F90:113-114).
module hpccg_capi
    use iso_c_binding
    implicit none
    private

    type(c_ptr), save :: ctx_handle = c_null_ptr
    integer, save :: resident_tag = 0   ! tag of the SparseMatrixT whose CSR is on the device
    integer, save :: next_tag = 0

    public :: hpccg_context, hpccg_check, hpccg_new_tag, hpccg_resident_tag, hpccg_set_resident
    public :: c_hpccg_matrix_upload, c_hpccg_generate_matrix, c_hpccg_spmv, c_hpccg_ddot, c_hpccg_ddot1
    public :: c_hpccg_waxpby, c_hpccg_wwpby, c_hpccg_wxbw, c_hpccg_vcopy, c_hpccg_cg
    public :: c_hpccg_compute_residual, c_hpccg_cg_timed

    interface
        integer(c_int) function c_hpccg_create(ctx, device) bind(C, name='hpccg_create')
            import :: c_ptr, c_int
            type(c_ptr) :: ctx
            integer(c_int), value :: device
        end function
        integer(c_int) function c_hpccg_set_dot_mode(ctx, mode) bind(C, name='hpccg_set_dot_mode')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: mode
        end function
        type(c_ptr) function c_hpccg_last_error() bind(C, name='hpccg_last_error')
            import :: c_ptr
        end function
        integer(c_int) function c_hpccg_matrix_upload(ctx, nrow, ncol, rowindex, colindex, values) &
                bind(C, name='hpccg_matrix_upload')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nrow, ncol
            integer(c_int32_t), intent(in) :: rowindex(*), colindex(*)
            real(c_double), intent(in) :: values(*)
        end function
        integer(c_int) function c_hpccg_generate_matrix(ctx, nx, ny, nz, rowindex, colindex, values, &
                x, b, xexact, nnz) bind(C, name='hpccg_generate_matrix')
            import :: c_ptr, c_int, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nx, ny, nz
            integer(c_int32_t) :: rowindex(*), colindex(*)
            real(c_double) :: values(*), x(*), b(*), xexact(*)
            integer(c_int64_t) :: nnz
        end function
        integer(c_int) function c_hpccg_spmv(ctx, x, y) bind(C, name='hpccg_spmv')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: x(*)
            real(c_double) :: y(*)
        end function
        integer(c_int) function c_hpccg_ddot(ctx, n, x, y, r) bind(C, name='hpccg_ddot')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: x(*), y(*)
            real(c_double) :: r
        end function
        integer(c_int) function c_hpccg_ddot1(ctx, n, x, r) bind(C, name='hpccg_ddot1')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: x(*)
            real(c_double) :: r
        end function
        integer(c_int) function c_hpccg_compute_residual(ctx, n, v1, v2, r) bind(C, name='hpccg_compute_residual')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: v1(*), v2(*)
            real(c_double) :: r
        end function
        integer(c_int) function c_hpccg_waxpby(ctx, n, alpha, x, beta, y, w) bind(C, name='hpccg_waxpby')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            real(c_double), value :: alpha, beta
            real(c_double), intent(in) :: x(*), y(*)
            real(c_double) :: w(*)
        end function
        integer(c_int) function c_hpccg_wwpby(ctx, n, beta, y, w) bind(C, name='hpccg_wwpby')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            real(c_double), value :: beta
            real(c_double), intent(in) :: y(*)
            real(c_double) :: w(*)
        end function
        integer(c_int) function c_hpccg_wxbw(ctx, n, x, beta, w) bind(C, name='hpccg_wxbw')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            real(c_double), value :: beta
            real(c_double), intent(in) :: x(*)
            real(c_double) :: w(*)
        end function
        integer(c_int) function c_hpccg_vcopy(ctx, n, x, y) bind(C, name='hpccg_vcopy')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n
            real(c_double), intent(in) :: x(*)
            real(c_double) :: y(*)
        end function
        integer(c_int) function c_hpccg_cg(ctx, b, x, max_iter, tolerance, niters, normr, hist) &
                bind(C, name='hpccg_cg')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: b(*)
            real(c_double) :: x(*)
            integer(c_int32_t), value :: max_iter
            real(c_double), value :: tolerance
            integer(c_int32_t) :: niters
            real(c_double) :: normr
            real(c_double) :: hist(*)
        end function
        integer(c_int) function c_hpccg_cg_timed(ctx, b, x, max_iter, tolerance, level, niters, normr, hist, &
                t_total, t_prepare, t_iter, t_spmv) bind(C, name='hpccg_cg_timed')
            import :: c_ptr, c_int, c_int32_t, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: b(*)
            real(c_double) :: x(*)
            integer(c_int32_t), value :: max_iter
            real(c_double), value :: tolerance
            integer(c_int32_t), value :: level
            integer(c_int32_t) :: niters
            real(c_double) :: normr
            real(c_double) :: hist(*)
            real(c_double) :: t_total, t_prepare
            real(c_double) :: t_iter(*), t_spmv(*)
        end function
    end interface

contains

    function hpccg_context() result(ctx)
        type(c_ptr) :: ctx
        character(len=32) :: val
        integer :: dev, ln, st, mode
        if (.not. c_associated(ctx_handle)) then
            dev = 0
            call get_environment_variable('HPCCG_DEVICE', val, ln, st)
            if (st == 0 .and. ln > 0) read (val, *) dev
            call hpccg_check(c_hpccg_create(ctx_handle, int(dev, c_int)), 'hpccg_create')
            mode = 1
            call get_environment_variable('HPCCG_DOT_MODE', val, ln, st)
            if (st == 0 .and. ln > 0) then
                if (val(1:1) == 's' .or. val(1:1) == 'S') mode = 0
            end if
            call hpccg_check(c_hpccg_set_dot_mode(ctx_handle, int(mode, c_int)), 'hpccg_set_dot_mode')
        end if
        ctx = ctx_handle
    end function hpccg_context

    subroutine hpccg_check(status, where)
        integer(c_int), intent(in) :: status
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: i
        if (status >= 0) return
        text = ''
        call c_f_pointer(c_hpccg_last_error(), msg, [512])
        do i = 1, 512
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
        end do
        write (*, '(a,a,a,i0,a,a)') 'libhpccg_hip: ', where, ' failed (', status, '): ', trim(text)
        error stop 1
    end subroutine hpccg_check

    integer function hpccg_new_tag()
        next_tag = next_tag + 1
        hpccg_new_tag = next_tag
    end function hpccg_new_tag

    integer function hpccg_resident_tag()
        hpccg_resident_tag = resident_tag
    end function hpccg_resident_tag

    subroutine hpccg_set_resident(tag)
        integer, intent(in) :: tag
        resident_tag = tag
    end subroutine hpccg_set_resident
end module hpccg_capi
