! lsmp_cases.f90 -- initial conditions of the reference's three transport
! cases on host boxes (1-ghc:n+ghc per direction, kb0:kb1 in z): the grid's
! cell centres (tree.f90:254-274), phi and the sub-sampled VOF of
! problem_init.f90 for Disk (Zalesak disk), Vortex2D (circle) and Vortex
! (sphere), and the zero-gradient ghost fill (branches.f90:283-310).  Used by
! lsmp_driver; a CPU test checks them against the oracle bit for bit.
module lsmp_cases
implicit none
private
public :: case_coords, case_initial_phi, case_initial_vof, case_bc_box

contains

! cell centres x(i) = start + (i - 0.5) dx with (i - 0.5) in default real,
! x(0) = start, x(n+1) = end (tree.f90:254-274)
subroutine case_coords(x, n, s, e, d)
    integer, intent(in) :: n
    real(8), intent(out) :: x(0:n+1)
    real(8), intent(in) :: s, e, d
    integer :: i
    do i = 1, n
        x(i) = s + (i - 0.5) * d
    enddo
    x(0) = s
    x(n+1) = e
end subroutine

! problem_init.f90:32-79 for the three cases, then the zero-gradient ghosts.
! The disk's phi is 2 vof - 1 of its 30^3 sub-samples; vof, when present,
! receives that VOF (interior cells).
subroutine case_initial_phi(casename, nx, ny, nz, ghc, kb0, kb1, dx, dy, xc, yc, zc, phi, vof)
    character(len=*), intent(in) :: casename
    integer, intent(in) :: nx, ny, nz, ghc, kb0, kb1
    real(8), intent(in) :: dx, dy, xc(0:nx+1), yc(0:ny+1), zc(0:nz+1)
    real(8), intent(inout) :: phi(1-ghc:nx+ghc, 1-ghc:ny+ghc, kb0:kb1)
    real(8), intent(inout), optional :: vof(1-ghc:nx+ghc, 1-ghc:ny+ghc, kb0:kb1)
    integer :: i, j, k, ii, jj, kk, sub
    real(8) :: x, y, v
    sub = 30
    do k = 1, nz
    do j = 1, ny
    do i = 1, nx
        select case (trim(casename))
        case ("disk")
            v = 0.0d0
            do ii = 1, sub
            do jj = 1, sub
            do kk = 1, sub
                x = 0.5d0*( xc(i)+xc(i-1) ) + real(ii,8)*dx/real(sub,8)
                y = 0.5d0*( yc(j)+yc(j-1) ) + real(jj,8)*dy/real(sub,8)
                if (sqrt((x-0.5)**2+(y-0.75)**2) <= 0.15) then
                    v = v + 1.0d0/real(sub,8)**3.0d0
                end if
                if (abs(x-0.5) <= 0.025 .and. y < 0.8 .and. (x-0.5)**2+(y-0.75)**2 <= 0.15**2) then
                    v = v - 1.0d0/real(sub,8)**3.0d0
                end if
            end do
            end do
            end do
            phi(i,j,k) = v * 2.0 - 1.0
            if (present(vof)) vof(i,j,k) = v
        case ("vortex2d")
            phi(i,j,k) = -sqrt((xc(i)-0.5)**2+(yc(j)-0.75)**2) + 0.15
        case default
            phi(i,j,k) = - sqrt((xc(i)-0.35)**2+(yc(j)-0.35)**2+(zc(k)-0.35)**2) + 0.15
        end select
    enddo
    enddo
    enddo
    call case_bc_box(phi, nx, ny, nz, ghc, kb0, kb1)
end subroutine

! The VOF of the circle (Vortex2D) or sphere (Vortex) from 30^3 sub-samples
! per cell (problem_init.f90:32-59; the 2D case loops over the z samples too),
! with its ghosts.  Every inside sample adds the same 1/30^3, so a cell whose
! centre lies more than 1.5 dx from the surface -- all its samples on one
! side -- holds the full sum or zero, and only the others are sampled.
subroutine case_initial_vof(casename, nx, ny, nz, ghc, kb0, kb1, dx, dy, dz, xc, yc, zc, vof)
    character(len=*), intent(in) :: casename
    integer, intent(in) :: nx, ny, nz, ghc, kb0, kb1
    real(8), intent(in) :: dx, dy, dz, xc(0:nx+1), yc(0:ny+1), zc(0:nz+1)
    real(8), intent(inout) :: vof(1-ghc:nx+ghc, 1-ghc:ny+ghc, kb0:kb1)
    integer :: i, j, k, ii, jj, kk, sub
    real(8) :: x, y, z, v, full, d
    logical :: circle
    circle = trim(casename) == "vortex2d"
    sub = 30
    full = 0.0d0
    do ii = 1, sub**3
        full = full + 1.0d0/real(sub,8)**3.0d0
    enddo
    do k = 1, nz
    do j = 1, ny
    do i = 1, nx
        if (circle) then
            d = sqrt((xc(i)-0.5)**2+(yc(j)-0.75)**2)
        else
            d = sqrt((xc(i)-0.35)**2+(yc(j)-0.35)**2+(zc(k)-0.35)**2)
        endif
        if (abs(d - 0.15d0) >= 1.5d0*dx) then
            if (d < 0.15d0) then
                vof(i,j,k) = full
            else
                vof(i,j,k) = 0.0d0
            endif
            cycle
        endif
        v = 0.0d0
        do ii = 1, sub
        do jj = 1, sub
        do kk = 1, sub
            x = 0.5d0*( xc(i)+xc(i-1) ) + real(ii,8)*dx/real(sub,8)
            y = 0.5d0*( yc(j)+yc(j-1) ) + real(jj,8)*dy/real(sub,8)
            z = 0.5d0*( zc(k)+zc(k-1) ) + real(kk,8)*dz/real(sub,8)
            if (circle) then
                if (sqrt((x-0.5)**2+(y-0.75)**2) <= 0.15) v = v + 1.0d0/real(sub,8)**3.0d0
            else
                if (sqrt((x-0.35)**2+(y-0.35)**2+(z-0.35)**2) <= 0.15) v = v + 1.0d0/real(sub,8)**3.0d0
            endif
        enddo
        enddo
        enddo
        vof(i,j,k) = v
    enddo
    enddo
    enddo
    call case_bc_box(vof, nx, ny, nz, ghc, kb0, kb1)
end subroutine

! zero-gradient ghost fill, x then y then z (branches.f90:283-310)
subroutine case_bc_box(dat, nx, ny, nz, ghc, kb0, kb1)
    integer, intent(in) :: nx, ny, nz, ghc, kb0, kb1
    real(8), intent(inout) :: dat(1-ghc:nx+ghc, 1-ghc:ny+ghc, kb0:kb1)
    integer :: i, j, k
    do k = kb0, kb1
    do j = 1-ghc, ny+ghc
    do i = 1, ghc
        dat(1-i,j,k) = dat(1,j,k)
        dat(nx+i,j,k) = dat(nx,j,k)
    enddo
    enddo
    enddo
    do k = kb0, kb1
    do i = 1-ghc, nx+ghc
    do j = 1, ghc
        dat(i,1-j,k) = dat(i,1,k)
        dat(i,ny+j,k) = dat(i,ny,k)
    enddo
    enddo
    enddo
    do j = 1-ghc, ny+ghc
    do i = 1-ghc, nx+ghc
    do k = 1, ghc
        dat(i,j,1-k) = dat(i,j,1)
        dat(i,j,nz+k) = dat(i,j,nz)
    enddo
    enddo
    enddo
end subroutine

end module lsmp_cases
