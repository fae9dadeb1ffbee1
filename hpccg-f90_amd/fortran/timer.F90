! Region timer with the interface of MPI/src/timer.F90:53-142 (type interval, timer_init,
! timer_start, timer_stop, timer_output, timer_print, timer_finalize).  Wall clock from
! system_clock at its finest int64 rate; results are default REAL like the reference's.
module timer
    implicit none
    private
    public :: interval, timer_init, timer_start, timer_stop, timer_output, timer_print, timer_finalize

    type interval
        double precision startt
        double precision stopt
    end type interval

contains

    double precision function wtime()
        integer(kind=8) :: cnt, rate
        call system_clock(cnt, rate)
        wtime = dble(cnt) / dble(rate)
    end function wtime

    subroutine timer_init()
    end subroutine timer_init

    subroutine timer_start(timing, region_name)
        type(interval), intent(inout) :: timing
        character(len=*), intent(in) :: region_name
        timing%startt = wtime()
    end subroutine timer_start

    subroutine timer_stop(timing, region_name)
        type(interval), intent(inout) :: timing
        character(len=*), intent(in) :: region_name
        timing%stopt = wtime()
    end subroutine timer_stop

    subroutine timer_output(timing, region_name)
        type(interval), intent(in) :: timing
        character(len=*), intent(in) :: region_name
        real :: timing_result
        timing_result = real(timing%stopt - timing%startt)
        print *, "TIMER: ", region_name, timing_result, "sec"
    end subroutine timer_output

    function timer_print(timing) result(timing_result)
        type(interval), intent(in) :: timing
        real :: timing_result
        timing_result = real(timing%stopt - timing%startt)
    end function timer_print

    subroutine timer_finalize()
    end subroutine timer_finalize
end module timer
