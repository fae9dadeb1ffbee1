/* Leveled timing macros with the interface of MPI/includes/timer.h:36-58: a region at level
 * `lev` is timed when TIMINGLEV > lev.  NOTIMING removes every region. */
#ifndef HPCCG_TIMER_H
#define HPCCG_TIMER_H
#ifndef TIMINGLEV
#define TIMINGLEV 1
#endif
#ifdef NOTIMING
#define TIMER_START(lev, timing, region_name)
#define TIMER_STOP(lev, timing, region_name)
#define TIMER_PRINT(lev, timing, region_name)
#else
#define TIMER_START(lev, timing, region_name) if (TIMINGLEV > lev) then; call timer_start(timing, region_name); endif
#define TIMER_STOP(lev, timing, region_name) if (TIMINGLEV > lev) then; call timer_stop(timing, region_name); endif
#define TIMER_PRINT(lev, timing, region_name) if (TIMINGLEV > lev) then; call timer_output(timing, region_name); endif
#endif
#endif
