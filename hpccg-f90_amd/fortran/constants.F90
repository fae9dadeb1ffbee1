! Kind parameters of the HPCCG-F90 interface (MPI/src/constants.F90:37-42).  The capacity
! limits are kept so that code written against the reference still compiles; they size the
! reference's multi-rank halo lists, which this single-GPU build does not have.
module constants
    implicit none
    integer, parameter :: sp = kind(0.0e0)
    integer, parameter :: dp = kind(0.0d0)
    integer, parameter :: max_external = 100000
    integer, parameter :: max_num_messages = 500
    integer, parameter :: max_num_neighbors = max_num_messages
end module constants
