! Writes the initial phi and VOF boxes of one case, from the driver's
! lsmp_cases module, for the CPU tests (tests/test_driver_cases_host.py):
!   cases_check <disk|vortex2d|vortex> <n> <output file>
! on the unit box with n cells per direction, ghc = 3, 7 z planes in 2D.
program cases_check
use lsmp_cases
implicit none
character(len=64) :: casename, arg
character(len=512) :: path
integer :: n, nx, ny, nz, ghc, kb0, kb1, u
real(8) :: dx, dy, dz
real(8), allocatable :: phi(:,:,:), vof(:,:,:), xc(:), yc(:), zc(:)

call get_command_argument(1, casename)
call get_command_argument(2, arg)
read(arg, *) n
call get_command_argument(3, path)
ghc = 3
nx = n
ny = n
dx = 1.0d0 / nx
dy = 1.0d0 / ny
if (trim(casename) == "vortex") then
    nz = n
    dz = 1.0d0 / nz
    kb0 = 1 - ghc; kb1 = nz + ghc
else
    nz = 1
    dz = 1.0d0
    kb0 = 1 - ghc; kb1 = 1 + ghc
endif
allocate(xc(0:nx+1), yc(0:ny+1), zc(0:nz+1))
call case_coords(xc, nx, 0.0d0, 1.0d0, dx)
call case_coords(yc, ny, 0.0d0, 1.0d0, dy)
call case_coords(zc, nz, 0.0d0, 1.0d0, dz)
allocate(phi(1-ghc:nx+ghc, 1-ghc:ny+ghc, kb0:kb1), vof(1-ghc:nx+ghc, 1-ghc:ny+ghc, kb0:kb1))
phi = 0.0d0
vof = 0.0d0
call case_initial_phi(casename, nx, ny, nz, ghc, kb0, kb1, dx, dy, xc, yc, zc, phi, vof)
if (trim(casename) == "disk") then
    call case_bc_box(vof, nx, ny, nz, ghc, kb0, kb1)
else
    call case_initial_vof(casename, nx, ny, nz, ghc, kb0, kb1, dx, dy, dz, xc, yc, zc, vof)
endif
open(newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
write(u) phi, vof
close(u)
end program cases_check
