! ref_pieces.f90 -- bind(C) entry points around the reference's own
! self-contained Fortran units, so that tests can run them from Python.
!
! TEST INFRASTRUCTURE ONLY.  This file holds no numerics of its own: it is
! linked against the reference's unmodified sources, compiled where they lie
! (see oracle/Makefile):
!   Vortex/src/others/TDMA.f90          solve_tridiagonal            (:1-36)
!   Vortex/lib/solver/matrix_solver.f90 TWIN_DEC / TWIN_BKS          (:6-78)
!   Vortex/lib/solver/ccd_solvers.f90   ccd_root init / solve "ccd"  (:382-427)
!   Vortex/src/interface_solver/level_set_rk3_solver.f90
!                                       wenojs_flux_split,
!                                       crweno_flux_split            (:288-414)
!   Vortex/src/interface_solver/level_set_rk3_redis.f90
!                                       phyn                         (:456-471)
! The last two files also hold units that `use all`; only the argument-only
! units above are compiled from them (whole, unmodified).  The rest of the
! reference cannot be built in this image (flang rejects the REAL subscript
! at lib/data/mutligrid_root.f90:32-34 and LAPACK dgesv, used at
! lib/data/tree_multigrid.f90:159,227, is absent), so the hot-path units
! that work on module `all` are not reachable from here.

module ref_pieces
use iso_c_binding
use ccd_solvers
implicit none

interface
    subroutine solve_tridiagonal(A, B, C, S, X, M, N)
        integer :: M, N
        real(8), dimension(m:n) :: A, B, C, S, X
    end subroutine
    subroutine crweno_flux_split(f, g, fp, gm, is, ie, ghc)
        integer :: is, ie, ghc
        real(8), dimension(is-ghc:ie+ghc) :: f, g, fp, gm
    end subroutine
    function phyn(a, b, c, d)
        real(8) :: a, b, c, d, phyn
    end function
end interface

contains

subroutine ref_solve_tridiagonal(A, B, C, S, X, n) bind(C, name="ref_solve_tridiagonal")
    integer(c_int), value :: n
    real(c_double), dimension(n) :: A, B, C, S, X
    call solve_tridiagonal(A, B, C, S, X, 1, n)
end subroutine

! CCD first/second derivative of one full line f(1:n) with spacing dx,
! exactly as surface_norms calls it (derivatives.f90:17-18).
subroutine ref_ccd_line(n, dx, f, fx, fxx) bind(C, name="ref_ccd_line")
    integer(c_int), value :: n
    real(c_double), value :: dx
    real(c_double), dimension(n) :: f, fx, fxx
    type(ccd_root) :: r
    real(8) :: dt
    dt = 0.0d0
    call r%init(1, n, dx, dt)
    call r%solve("ccd", f, fx, fxx)
end subroutine

! One CRWENO5 line as level_set_rk3_source_weno calls it
! (level_set_rk3_solver.f90:149-151): split fluxes f (negative) and g
! (positive) on (1-3 : n+3); fp and gm receive the face fluxes.
subroutine ref_crweno_flux_split(f, g, fp, gm, n) bind(C, name="ref_crweno_flux_split")
    integer(c_int), value :: n
    real(c_double), dimension(1-3:n+3) :: f, g, fp, gm
    call crweno_flux_split(f, g, fp, gm, 1, n, 3)
end subroutine

! The WENO-HJ correction term (level_set_rk3_redis.f90:456-471).
function ref_phyn(a, b, c, d) bind(C, name="ref_phyn")
    real(c_double), value :: a, b, c, d
    real(c_double) :: ref_phyn
    ref_phyn = phyn(a, b, c, d)
end function

! Upwind CCD first derivative (root_solve "uccd", ccd_solvers.f90:428-438).
subroutine ref_uccd_line(n, dx, f, u, fx) bind(C, name="ref_uccd_line")
    integer(c_int), value :: n
    real(c_double), value :: dx
    real(c_double), dimension(n) :: f, u, fx
    real(8), dimension(n) :: fxx
    type(ccd_root) :: r
    real(8) :: dt
    dt = 0.0d0
    call r%init(1, n, dx, dt)
    call r%solve("uccd", f, fx, fxx, u)
end subroutine

! The factorised CCD coefficient tables after TWIN_DEC (root_init :390-392).
! out holds A, B, AA, BB one after another, each (3, n) column-major.
subroutine ref_ccd_tables(n, dx, out) bind(C, name="ref_ccd_tables")
    integer(c_int), value :: n
    real(c_double), value :: dx
    real(c_double), dimension(3, n, 4) :: out
    type(ccd_root) :: r
    real(8) :: dt
    dt = 0.0d0
    call r%init(1, n, dx, dt)
    out(:, :, 1) = r%ccd%AU
    out(:, :, 2) = r%ccd%BU
    out(:, :, 3) = r%ccd%AAU
    out(:, :, 4) = r%ccd%BBU
end subroutine

end module ref_pieces
