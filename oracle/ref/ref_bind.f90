! C bindings of the reference's MPI/NetCDF-free leaf modules (stafft,
! scherzinger), compiled from the reference sources by oracle/ref/Makefile.
! TEST INFRASTRUCTURE ONLY: used by tests/ to pin the oracle's transforms and
! eigen-solver to the reference's own arithmetic.
module ref_bind
    use iso_c_binding
    use stafft, only : initfft, forfft, revfft, dct, dst
    use scherzinger, only : scherzinger_diagonalise, scherzinger_eigenvalues
    implicit none

contains

    subroutine ref_forfft(n, x) bind(C, name="ref_forfft")
        integer(c_int), value         :: n
        real(c_double), intent(inout) :: x(n)
        integer                       :: factors(5)
        double precision              :: trig(2 * n)
        call initfft(n, factors, trig)
        call forfft(1, n, x, trig, factors)
    end subroutine ref_forfft

    subroutine ref_revfft(n, x) bind(C, name="ref_revfft")
        integer(c_int), value         :: n
        real(c_double), intent(inout) :: x(n)
        integer                       :: factors(5)
        double precision              :: trig(2 * n)
        call initfft(n, factors, trig)
        call revfft(1, n, x, trig, factors)
    end subroutine ref_revfft

    ! x(0:n)
    subroutine ref_dct(n, x) bind(C, name="ref_dct")
        integer(c_int), value         :: n
        real(c_double), intent(inout) :: x(0:n)
        integer                       :: factors(5)
        double precision              :: trig(2 * n)
        call initfft(n, factors, trig)
        call dct(1, n, x, trig, factors)
    end subroutine ref_dct

    ! x(1:n)
    subroutine ref_dst(n, x) bind(C, name="ref_dst")
        integer(c_int), value         :: n
        real(c_double), intent(inout) :: x(n)
        integer                       :: factors(5)
        double precision              :: trig(2 * n)
        call initfft(n, factors, trig)
        call dst(1, n, x, trig, factors)
    end subroutine ref_dst

    ! A(3,3) symmetric (overwritten), D(3) descending, V(3,3) eigenvectors in columns
    subroutine ref_scherzinger(A, D, V) bind(C, name="ref_scherzinger")
        real(c_double), intent(inout) :: A(3, 3)
        real(c_double), intent(out)   :: D(3), V(3, 3)
        call scherzinger_diagonalise(A, D, V)
    end subroutine ref_scherzinger

end module ref_bind
