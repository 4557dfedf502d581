! lsmp_driver.f90 -- Fortran driver of the gfx950 level-set transport library.
!
! Plays the role of the reference's program main (Vortex/src/main.f90:10-53)
! with problem_init (src/setup/problem_init.f90), the deck reader
! (lib/data/tree.f90:90-184, 230-285, 352) and the per-step output
! (src/others/output.f90:1-36, 95-143, 169-231), for the three transport
! cases of the reference: Disk (2D Zalesak disk, rigid rotation), Vortex2D
! (2D single vortex) and Vortex (3D LeVeque deformation).  The hot path --
! advection, reinitialisation, mass correction, mass/volume sums and the L1
! metric -- runs on the GPU through lsmp_mod's ISO_C_BINDING interface.
!
! usage:  lsmp_driver <disk|vortex2d|vortex> [deck=./input/default.txt] [max_steps]
!                     [fp64|fp32] [crweno|uccd] [energy]
! (max_steps 0 = no limit; the next two choose the bulk precision and the
! level_set_rk3_source variant, level_set_rk3_solver.f90:96-102; "energy"
! adds energy_and_momentum and its Energy, Momentum and Position rows every
! step, output.f90:33-36, 250-397 -- the reference's Navier-Stokes
! diagnostics, nine CCD solves of the velocity per step, off by default)
!
! The deck has the reference's format (odd lines are comments).  Like the
! reference's shipped cases it needs threads grid 1 1 1 and ghc = 3; methods 1
! (LS), 2 (MPLS) and 3 (CLSVOF, its merged branch) are supported.
program lsmp_driver
use iso_c_binding
use lsmp
use lsmp_cases
implicit none

character(len=256) :: casename, deck, arg
character(len=30) :: name
integer :: method, gx, gy, gz, nthreads, nplots, level, ug, ghc, how_to_paras, inverse
integer :: t2red, iter, max_steps, dim, field, iostat, ubc(2), vbc(2), wbc(2), per(3)
integer(c_int) :: precision, scheme
integer(8) :: clk0, clk1, clkrate
real(8) :: xstart, xend, ystart, yend, zstart, zend, t2s, t2p, dt, rdt, ls_wid
real(8) :: p_tol, p_w1, p_w2, p_b, t_tol, t_w, rho_1, mu_1, rho_2, mu_2, sigma, g, ggx, ggy, ggz
real(8) :: cl, cu, ct, re, fr, we, rho_12, mu_12
real(8) :: energy_unit, ek, ep, es, ev, ek0, ep0, es0, em(11)
logical :: energy
integer :: btn_sf, btn_g
integer :: nx, ny, nz, kb0, kb1
real(8) :: dx, dy, dz, time, mass, vol, imass, ivol, red_error, l1
real(8) :: loss_mass_avg, loss_mass_max, loss_vol_avg, loss_vol_max
real(8), allocatable :: phi(:,:,:), vofb(:,:,:), xc(:), yc(:), zc(:)
real(8) :: massv, volv, imassv, ivolv, l1v
real(8) :: loss_mass_avgv, loss_mass_maxv, loss_vol_avgv, loss_vol_maxv
type(lsmp_params) :: prm
type(lsmp_stats) :: st
type(c_ptr) :: h
integer(c_int) :: npseudo, nmass, nrec0, nrec1
integer :: fmass, fvol, fen, fmom, fpos

! ------------------------------------------------------------ arguments
call get_command_argument(1, casename)
if (len_trim(casename) == 0) then
    write(*,*) "usage: lsmp_driver <disk|vortex2d|vortex> [deck] [max_steps] [fp64|fp32] [crweno|uccd]"
    stop 1
endif
deck = "./input/default.txt"
call get_command_argument(2, arg)
if (len_trim(arg) > 0) deck = arg
max_steps = huge(1)
call get_command_argument(3, arg)
if (len_trim(arg) > 0) read(arg, *) max_steps
if (max_steps <= 0) max_steps = huge(1)
precision = LSMP_FP64
call get_command_argument(4, arg)
if (trim(arg) == "fp32") precision = LSMP_FP32
scheme = LSMP_SCHEME_CRWENO
call get_command_argument(5, arg)
if (trim(arg) == "uccd") scheme = LSMP_SCHEME_UCCD
call get_command_argument(6, arg)
energy = trim(arg) == "energy"

select case (trim(casename))
case ("disk")
    dim = 2; field = LSMP_VEL_ROTATION
case ("vortex2d")
    dim = 2; field = LSMP_VEL_VORTEX2D
case ("vortex")
    dim = 3; field = LSMP_VEL_LEVEQUE3D
case default
    write(*,*) "unknown case ", trim(casename), " (disk | vortex2d | vortex)"
    stop 1
end select

! ------------------------------------------------------------ deck (tree.f90:96-157)
open(unit=526, file=trim(deck), status='old')
read(526,*); read(526,*) method
read(526,*); read(526,*) name
read(526,*); read(526,*) gx, gy, gz
read(526,*); read(526,*) nthreads
read(526,*); read(526,*) nplots
read(526,*); read(526,*) level
read(526,*); read(526,*) ug
read(526,*); read(526,*) ghc
read(526,*); read(526,*) xstart, xend
read(526,*); read(526,*) ystart, yend
read(526,*); read(526,*) zstart, zend
read(526,*); read(526,*) t2s, t2p, t2red
read(526,*); read(526,*) dt, rdt
read(526,*); read(526,*) p_tol, p_w1, p_w2, p_b
read(526,*); read(526,*) t_tol, t_w
read(526,*); read(526,*) ls_wid
read(526,*); read(526,*) how_to_paras
read(526,*); read(526,*) rho_1, mu_1
read(526,*); read(526,*) rho_2, mu_2
read(526,*); read(526,*) sigma, btn_sf
read(526,*); read(526,*) g, btn_g
read(526,*); read(526,*) ggx, ggy, ggz
read(526,*); read(526,*) cl, cu, ct
read(526,*); read(526,*) re, fr, we
read(526,*); read(526,*) rho_12, mu_12
read(526,*); read(526,*) inverse
read(526,*); read(526,*) ubc(1), ubc(2)
read(526,*); read(526,*) vbc(1), vbc(2)
read(526,*); read(526,*) wbc(1), wbc(2)
read(526,*); read(526,*) per(1), per(2), per(3)
close(526)

if (method < 1 .or. method > 3) then
    write(*,*) "Error >> method ", method, " is not provided by this driver (1 LS, 2 MPLS, 3 CLSVOF)"
    stop 1
endif
if (gx /= 1 .or. gy /= 1 .or. gz /= 1) then
    ! the reference truncates line solves at subdomain edges (branches.f90:97-136);
    ! the device path solves whole lines
    write(*,*) "Error >> threads grid must be 1 1 1"
    stop 1
endif
if (any(per /= 0)) then
    write(*,*) "Error >> periodic directions are not supported"
    stop 1
endif
if (how_to_paras < 1 .or. how_to_paras > 5) then
    write(*,*) "Error >> Wrong parameter selector "          ! tree.f90:304-306
    stop 1
endif
if (how_to_paras > 1) rho_12 = rho_2 / rho_1                 ! tree.f90:310-312
! reference scales and the dimensionless groups (tree.f90:289-318): cl, cu, ct
! are the deck's L, U, T
select case (how_to_paras)
case (1)
    cl = ((re * mu_1 / rho_1)**2.0d0 * g / fr)**(1.0d0/3.0d0)
    cu = dsqrt(fr * g * cl)
    ct = cl / cu
case (2)
    cu = dsqrt(cl * g)
    ct = cl / cu
case (3)
    ct = cl / cu
case (4)
    cu = cl / ct
case (5)
    cl = cu * ct
end select
if (how_to_paras > 1) then
    mu_12 = mu_2 / mu_1
    re = rho_1 * cu * cl / mu_1
    we = rho_1 * cu**2.0d0 * cl / sigma
    fr = cu**2.0d0 / (g * cl)
endif
energy_unit = rho_1 * g * cl**4.0

! ------------------------------------------------------------ grid (tree.f90:230-280)
nx = int(ug * (xend - xstart))
ny = int(ug * (yend - ystart))
if (dim == 3) then
    nz = int(ug * (zend - zstart))
    dz = (zend - zstart) / nz
    kb0 = 1 - ghc; kb1 = nz + ghc
else
    nz = 1
    dz = 1.0d0
    kb0 = 1 - ghc; kb1 = 1 + ghc
endif
dx = (xend - xstart) / nx
dy = (yend - ystart) / ny
dt = dt * dx
rdt = rdt * dx
ls_wid = ls_wid * dx
if (t2red < 0) t2red = int(t2s / dt) + 1                      ! tree.f90:352

allocate(xc(0:nx+1), yc(0:ny+1), zc(0:nz+1))
call case_coords(xc, nx, xstart, xend, dx)
call case_coords(yc, ny, ystart, yend, dy)
call case_coords(zc, nz, zstart, zend, dz)

write(*,*) " --- Problem information --- "
write(*,'(A30)') name
write(*,'(A20,F15.8)') "dx:", dx
write(*,'(A20,F15.8)') "dt:", dt
write(*,'(A20,I5,A3,I5,A3,I5)') "Grids:", nx, "x", ny, "x", nz

prm%nx = nx; prm%ny = ny; prm%nz = nz; prm%ghc = ghc; prm%dim = dim
prm%dx = dx; prm%dy = dy; prm%dz = dz
prm%dt = dt; prm%rdt = rdt; prm%ls_wid = ls_wid
prm%rho_12 = rho_12; prm%inverse = inverse; prm%precision = precision
prm%xstart = xstart; prm%xend = xend; prm%ystart = ystart; prm%yend = yend
prm%zstart = zstart; prm%zend = zend
call lsmp_check(c_lsmp_init(prm, h), "lsmp_init")
call lsmp_check(c_lsmp_set_advection_scheme(h, scheme), "lsmp_set_advection_scheme")

! ------------------------------------------------------------ initial condition (problem_init.f90:32-103)
allocate(phi(1-ghc:nx+ghc, 1-ghc:ny+ghc, kb0:kb1))
if (method == 3) allocate(vofb(1-ghc:nx+ghc, 1-ghc:ny+ghc, kb0:kb1))
if (method == 3) then
    call case_initial_phi(casename, nx, ny, nz, ghc, kb0, kb1, dx, dy, xc, yc, zc, phi, vofb)
else
    call case_initial_phi(casename, nx, ny, nz, ghc, kb0, kb1, dx, dy, xc, yc, zc, phi)
endif
call lsmp_check(c_lsmp_upload_phi(h, phi), "lsmp_upload_phi")
if (trim(casename) == "disk") then
    call lsmp_check(c_lsmp_reinit(h, 0_c_int, -1.0d0, red_error, npseudo), "lsmp_reinit(0)")  ! :103
    write(*,'("LS Init:",I8,ES20.4)') npseudo, red_error
endif
call lsmp_check(c_lsmp_store_phi0(h), "lsmp_store_phi0")     ! phi%tmp2 (roots.f90:150-166)
call lsmp_check(c_lsmp_mass_vol(h, imass, ivol), "lsmp_mass_vol")   ! problem_init.f90:140-142
massv = 0.0d0; volv = 0.0d0; imassv = 0.0d0; ivolv = 0.0d0; l1v = 0.0d0
if (trim(casename) /= "disk") red_error = 0.0d0              ! printed before any btn = 1 reinitialisation
if (method == 3) then
    if (trim(casename) == "disk") then
        call case_bc_box(vofb, nx, ny, nz, ghc, kb0, kb1)   ! the disk's VOF was sampled with its phi
    else
        call case_initial_vof(casename, nx, ny, nz, ghc, kb0, kb1, dx, dy, dz, xc, yc, zc, vofb)
    endif
    call lsmp_check(c_lsmp_upload_vof(h, vofb), "lsmp_upload_vof")
    call lsmp_check(c_lsmp_store_vof0(h), "lsmp_store_vof0")           ! vof%tmp2
    call lsmp_check(c_lsmp_vof_mass_vol(h, imassv, ivolv), "lsmp_vof_mass_vol")
endif

! ------------------------------------------------------------ outputs (tree.f90:199-217)
call execute_command_line("mkdir -p out", exitstat=iostat)
fmass = 15; fvol = 16; fen = 17; fmom = 18; fpos = 19
open(unit=fmass, file="./out/"//trim(name)//"_MassLoss.plt")
write(fmass,*) 'variables = "T" "LS" "VOF" "B1" "B2" '
open(unit=fvol, file="./out/"//trim(name)//"_VolumeLoss.plt")
write(fvol,*) 'variables = "T" "LS" "VOF" "B1" "B2" '
open(unit=fen, file="./out/"//trim(name)//"_Energy.plt")
write(fen,*) 'variables = "T" "Ek" "Ep" "Es" "Ev" "E" "E0" '
open(unit=fmom, file="./out/"//trim(name)//"_Momentum.plt")
write(fmom,*) 'variables = "T" "Px" "Py" "Pz" '
open(unit=fpos, file="./out/"//trim(name)//"_Position.plt")
write(fpos,*) 'variables = "T" "x" "y" "z" '

! ------------------------------------------------------------ time loop (main.f90:17-51)
time = 0.0d0
iter = 0
loss_mass_avg = 0.0d0; loss_mass_max = 0.0d0
loss_vol_avg = 0.0d0; loss_vol_max = 0.0d0
loss_mass_avgv = 0.0d0; loss_mass_maxv = 0.0d0
loss_vol_avgv = 0.0d0; loss_vol_maxv = 0.0d0
ev = -1.0d0                                                  ! problem_init.f90:138
ek0 = 0.0d0; ep0 = 0.0d0; es0 = 0.0d0
call lsmp_check(c_lsmp_synchronize(h), "lsmp_synchronize")
call system_clock(clk0, clkrate)
do
    if (time + dt > t2s) dt = t2s - time + 1.0d-10
    time = time + dt
    iter = iter + 1
    call lsmp_check(c_lsmp_set_velocity_analytic(h, int(field, c_int), time, t2s), "velocity")

    if (method == 3) then
        ! clsvof merged branch (interface_solver.f90:234-244) after p%switch
        call lsmp_check(c_lsmp_set_face_velocity_analytic(h, int(field, c_int), time, t2s), "face velocity")
        call lsmp_check(c_lsmp_store_vof_old(h), "lsmp_store_vof_old")
        call lsmp_check(c_lsmp_vof_wlic(h, dt), "lsmp_vof_wlic")
        call lsmp_check(c_lsmp_rk3_advect(h, dt), "lsmp_rk3_advect")
        if (mod(iter, t2red) == 0) then
            call lsmp_check(c_lsmp_clsvof_recon(h, nrec0, nrec1), "lsmp_clsvof_recon")
        endif
    else
        ! interface_solver -> classical_ls / mpls merged branch (interface_solver.f90:15-23, 123-129)
        call lsmp_check(c_lsmp_rk3_advect(h, dt), "lsmp_rk3_advect")
        if (mod(iter, t2red) == 0) then
            call lsmp_check(c_lsmp_reinit(h, 1_c_int, -1.0d0, red_error, npseudo), "lsmp_reinit")
        endif
        if (method == 2) then
            call lsmp_check(c_lsmp_mass_fix(h, imass, mass, nmass), "lsmp_mass_fix")
        endif
    endif
    call lsmp_check(c_lsmp_mass_vol(h, mass, vol), "lsmp_mass_vol")      ! main.f90:35
    if (method == 3) call lsmp_check(c_lsmp_vof_mass_vol(h, massv, volv), "lsmp_vof_mass_vol")

    ! output (output.f90:7-31); the marker and NS columns of the reference
    ! are not transported on this path and are written as zero, and so is
    ! the VOF column unless the VOF is transported (CLSVOF)
    if (iter == 1) ivol = vol
    if (iter == 1) ivolv = volv
    if (method == 3) then
        write(fmass,'(7ES15.4)') time, 100.0d0*(mass-imass)/imass, 100.0d0*(massv-imassv)/imassv, 0.0d0, 0.0d0
        write(fvol,'(7ES15.4)') time, 100.0d0*(vol-ivol)/ivol, 100.0d0*(volv-ivolv)/ivolv, 0.0d0, 0.0d0
        loss_mass_avgv = loss_mass_avgv + abs(imassv-massv)/imassv
        loss_vol_avgv = loss_vol_avgv + abs(ivolv-volv)/ivolv
        loss_mass_maxv = max(loss_mass_maxv, abs(imassv-massv)/imassv)
        loss_vol_maxv = max(loss_vol_maxv, abs(ivolv-volv)/ivolv)
    else
        write(fmass,'(7ES15.4)') time, 100.0d0*(mass-imass)/imass, 0.0d0, 0.0d0, 0.0d0
        write(fvol,'(7ES15.4)') time, 100.0d0*(vol-ivol)/ivol, 0.0d0, 0.0d0, 0.0d0
    endif
    loss_mass_avg = loss_mass_avg + abs(imass-mass)/imass
    loss_vol_avg = loss_vol_avg + abs(ivol-vol)/ivol
    loss_mass_max = max(loss_mass_max, abs(imass-mass)/imass)
    loss_vol_max = max(loss_vol_max, abs(ivol-vol)/ivol)

    if (energy) then
        ! energy_and_momentum (output.f90:250-397) and its rows (:33-36)
        call lsmp_check(c_lsmp_energy_momentum(h, mu_12, merge(1_c_int, 0_c_int, method == 3), em), &
                        "lsmp_energy_momentum")
        es = em(1) * fr / we * energy_unit * btn_sf
        ek = em(9) * fr * energy_unit
        ep = em(10) * energy_unit
        if (ev < 0.0d0) then
            es0 = es; ek0 = ek; ep0 = ep
            ev = 0.0d0
        else
            ev = ev + em(11) * dt * fr / re * energy_unit
        endif
        write(fen, '(7ES15.4)') time, ek, ep, es, ev, ek+es+ep+ev, ek0+ep0+es0
        write(fmom, '(7ES15.4)') time, em(3), em(4), em(5)
        write(fpos, '(7ES15.4)') time, em(6), em(7), em(8)
    endif

    if (mod(iter, 5) == 0) then
        write(*,*) "========================================="
        write(*,'(A,",",I8,",",F15.6)') trim(name), iter, time
        write(*,*) ''
        write(*,'("LS,  Loss of mass  (%) :",ES15.4)') 100.0d0*(mass-imass)/imass
        write(*,'("LS,  Loss of volume(%) :",ES15.4)') 100.0d0*(ivol-vol)/ivol
        write(*,'("LS,  redistance error  :",ES15.4)') red_error
        write(*,*) ''
        call lsmp_check(c_lsmp_get_stats(h, st), "lsmp_get_stats")
        call print_device_time(st)
        write(*,*) "========================================="
    endif

    if (time + 1.0d-10 > t2s) exit
    if (iter >= max_steps) exit
enddo
call lsmp_check(c_lsmp_synchronize(h), "lsmp_synchronize")
call system_clock(clk1)

! ------------------------------------------------------------ calculate_interface_loss (output.f90:169-231)
call lsmp_check(c_lsmp_interface_l1(h, l1), "lsmp_interface_l1")
if (method == 3) call lsmp_check(c_lsmp_vof_interface_l1(h, l1v), "lsmp_vof_interface_l1")
open(unit=456, file="./Vortex_Results.txt", action='write', position='append')
if (method == 1) then
    write(456, *) "Method: LS"
else if (method == 2) then
    write(456, *) "Method: MPLS"
else
    write(456, *) "Method: CLSVOF"
endif
write(456,*) "Density Ratio:", rho_12
write(456,*) "Nx :", nx
if (method == 3) then
    write(456,'(A10, 2ES15.4)') "VOF Mass:", loss_mass_avgv / iter, loss_mass_maxv
    write(456,'(A10, 2ES15.4)') "VOF Vol:", loss_vol_avgv / iter, loss_vol_maxv
    write(456,'(A10, 2ES15.4)') "VOF Int:", l1v
else
    write(456,'(A10, 2ES15.4)') "LS Mass:", loss_mass_avg / iter, loss_mass_max
    write(456,'(A10, 2ES15.4)') "LS Vol:", loss_vol_avg / iter, loss_vol_max
    write(456,'(A10, 2ES15.4)') "LS Int:", l1
endif
close(456)
write(*,'(A,I8,A,ES24.16,A,ES24.16)') " finished: steps", iter, "  L1", l1, "  max |dM|/M", loss_mass_max
if (method == 3) write(*,'(A,ES24.16,A,ES24.16)') " VOF: L1", l1v, "  max |dM|/M", loss_mass_maxv
write(*,'(A,F12.4,A,ES12.4)') " wall ms/step", 1.0d3 * real(clk1 - clk0, 8) / real(clkrate, 8) / iter, &
    "  cell-updates/s", real(nx, 8) * ny * nz * iter / (real(clk1 - clk0, 8) / real(clkrate, 8))

close(fmass); close(fvol); close(fen); close(fmom); close(fpos)
call c_lsmp_finalize(h)

contains

! print_CPU_info (output.f90:129-143) with device time per phase
subroutine print_device_time(s)
    type(lsmp_stats), intent(in) :: s
    real(8) :: total
    total = s%t_advect_ms + s%t_reinit_ms + s%t_massfix_ms
    if (total <= 0.0d0) return
    write(*,'("Total device time(s) :",F15.6)') total * 1.0d-3
    write(*,'(3A18)') "Inter. Adv.", "Inter. Recon.", "Mass corr."
    write(*,'(F17.2,"%",F17.2,"%",F17.2,"%")') 100.0d0*s%t_advect_ms/total, &
        & 100.0d0*s%t_reinit_ms/total, 100.0d0*s%t_massfix_ms/total
end subroutine

end program lsmp_driver
