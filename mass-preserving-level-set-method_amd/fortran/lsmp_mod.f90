! lsmp_mod.f90 -- ISO_C_BINDING interface of the gfx950 level-set transport
! library (include/lsmp.h) for Fortran drivers.
!
! The reference reaches its hot path through argument-less internal
! subroutines (interface_solver.f90:1-13, level_set_rk3_solver.f90:1-94,
! level_set_rk3_redis.f90:1-79, mass_preserving_level_set.f90:1-74,
! ns_solver/others.f90:172-247).  A driver replaces those calls with the
! wrappers at the end of this module; each stops the program with the
! library's message on a non-zero return code, as the reference stops on its
! own error conditions.
module lsmp
use iso_c_binding
implicit none
private

integer(c_int), parameter, public :: LSMP_FP64 = 0, LSMP_FP32 = 1
integer(c_int), parameter, public :: LSMP_VEL_ROTATION = 0, LSMP_VEL_VORTEX2D = 1, LSMP_VEL_LEVEQUE3D = 2
integer(c_int), parameter, public :: LSMP_SCHEME_CRWENO = 0, LSMP_SCHEME_UCCD = 1
integer(c_int), parameter, public :: LSMP_FALLBACK_RESTART = 0, LSMP_FALLBACK_CONTINUE = 1, &
    LSMP_FALLBACK_CONTINUE_RELATIVE = 2

type, bind(C), public :: lsmp_params
    integer(c_int) :: nx, ny, nz, ghc, dim
    real(c_double) :: dx, dy, dz
    real(c_double) :: dt, rdt, ls_wid
    real(c_double) :: rho_12
    integer(c_int) :: inverse, precision
    real(c_double) :: xstart, xend, ystart, yend, zstart, zend
end type lsmp_params

type, bind(C), public :: lsmp_stats
    integer(c_int) :: last_redis_iters, last_mass_iters
    real(c_double) :: last_red_error
    real(c_double) :: t_advect_ms, t_reinit_ms, t_massfix_ms
    integer(c_long_long) :: n_advect, n_reinit, n_massfix
end type lsmp_stats

interface
    integer(c_int) function c_lsmp_init(p, h) bind(C, name="lsmp_init")
        import :: c_int, c_ptr, lsmp_params
        type(lsmp_params), intent(in) :: p
        type(c_ptr), intent(out) :: h
    end function
    subroutine c_lsmp_finalize(h) bind(C, name="lsmp_finalize")
        import :: c_ptr
        type(c_ptr), value :: h
    end subroutine
    type(c_ptr) function c_lsmp_error_string(code) bind(C, name="lsmp_error_string")
        import :: c_int, c_ptr
        integer(c_int), value :: code
    end function
    integer(c_int) function c_lsmp_upload_phi(h, box) bind(C, name="lsmp_upload_phi")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), intent(in) :: box(*)
    end function
    integer(c_int) function c_lsmp_download_phi(h, box) bind(C, name="lsmp_download_phi")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), intent(out) :: box(*)
    end function
    integer(c_int) function c_lsmp_store_phi0(h) bind(C, name="lsmp_store_phi0")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
    end function
    integer(c_int) function c_lsmp_set_velocity(h, u, v, w) bind(C, name="lsmp_set_velocity")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), intent(in) :: u(*), v(*), w(*)
    end function
    integer(c_int) function c_lsmp_set_velocity_analytic(h, field, t, t2s) bind(C, name="lsmp_set_velocity_analytic")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int), value :: field
        real(c_double), value :: t, t2s
    end function
    integer(c_int) function c_lsmp_rk3_advect(h, dt) bind(C, name="lsmp_rk3_advect")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), value :: dt
    end function
    integer(c_int) function c_lsmp_reinit(h, btn, timestop, red_error, n_pseudo) bind(C, name="lsmp_reinit")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int), value :: btn
        real(c_double), value :: timestop
        real(c_double), intent(out) :: red_error
        integer(c_int), intent(out) :: n_pseudo
    end function
    integer(c_int) function c_lsmp_mass_vol(h, mass, vol) bind(C, name="lsmp_mass_vol")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), intent(out) :: mass, vol
    end function
    ! no counterpart in the reference: the sums of lsmp_mass_vol on a device
    ! field, allocating nothing (for callers that cannot allocate)
    integer(c_int) function c_lsmp_mass_vol_noalloc(phi, nx, ny, nz, dim, dx, dy, dz, ls_wid, rho_12, inverse, &
                                                    out, stream) bind(C, name="lsmp_mass_vol_noalloc")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: phi, out, stream
        integer(c_int), value :: nx, ny, nz, dim, inverse
        real(c_double), value :: dx, dy, dz, ls_wid, rho_12
    end function
    integer(c_int) function c_lsmp_mass_fix(h, imass, mass, n_iter) bind(C, name="lsmp_mass_fix")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), value :: imass
        real(c_double), intent(out) :: mass
        integer(c_int), intent(out) :: n_iter
    end function
    integer(c_int) function c_lsmp_interface_l1(h, l1) bind(C, name="lsmp_interface_l1")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), intent(out) :: l1
    end function
    integer(c_int) function c_lsmp_get_stats(h, s) bind(C, name="lsmp_get_stats")
        import :: c_int, c_ptr, lsmp_stats
        type(c_ptr), value :: h
        type(lsmp_stats), intent(out) :: s
    end function
    integer(c_int) function c_lsmp_synchronize(h) bind(C, name="lsmp_synchronize")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
    end function
    ! level_set_rk3_source selector (level_set_rk3_solver.f90:96-102)
    integer(c_int) function c_lsmp_set_advection_scheme(h, scheme) bind(C, name="lsmp_set_advection_scheme")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
        integer(c_int), value :: scheme
    end function
    ! mass correction on every line (0), on the band lines where that pays (1, default) or always (2)
    integer(c_int) function c_lsmp_set_band_lines(h, mode) bind(C, name="lsmp_set_band_lines")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
        integer(c_int), value :: mode
    end function
    ! lines the last mass correction solved along x, y, z (-1: every line)
    integer(c_int) function c_lsmp_band_line_counts(h, counts) bind(C, name="lsmp_band_line_counts")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
        integer(c_int), intent(out) :: counts(3)
    end function
    ! narrow band of half-width width_cells (3D; 0 = off): the chosen forms only
    integer(c_int) function c_lsmp_set_narrow_band(h, width_cells) bind(C, name="lsmp_set_narrow_band")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), value :: width_cells
    end function
    ! two-marker MPLS branch (interface_solver.f90:34-121)
    integer(c_int) function c_lsmp_set_marker(h, mid, box) bind(C, name="lsmp_set_marker")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int), value :: mid
        real(c_double), intent(in) :: box(*)
    end function
    integer(c_int) function c_lsmp_get_marker(h, mid, box) bind(C, name="lsmp_get_marker")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int), value :: mid
        real(c_double), intent(out) :: box(*)
    end function
    integer(c_int) function c_lsmp_marker_load(h, mid) bind(C, name="lsmp_marker_load")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
        integer(c_int), value :: mid
    end function
    integer(c_int) function c_lsmp_marker_store(h, mid) bind(C, name="lsmp_marker_store")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
        integer(c_int), value :: mid
    end function
    integer(c_int) function c_lsmp_combine_markers(h, reconstruct) bind(C, name="lsmp_combine_markers")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
        integer(c_int), intent(out) :: reconstruct
    end function
    ! device-time accounting per kernel (HIP events on the solver's stream)
    integer(c_int) function c_lsmp_set_profiling(h, on) bind(C, name="lsmp_set_profiling")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
        integer(c_int), value :: on
    end function
    integer(c_int) function c_lsmp_kernel_time(h, name, launches, ms) bind(C, name="lsmp_kernel_time")
        import :: c_int, c_ptr, c_char, c_long_long, c_double
        type(c_ptr), value :: h
        character(kind=c_char), dimension(*), intent(in) :: name   ! NUL-terminated
        integer(c_long_long), intent(out) :: launches
        real(c_double), intent(out) :: ms
    end function
    ! connected components of phi > phi_core grown through phi_min < phi <= phi_core, and their masses
    integer(c_int) function c_lsmp_component_masses(h, phi_core, phi_min, nquery, query_cells, query_comp, &
                                                    maxcomp, comp_mass, ncomp) bind(C, name="lsmp_component_masses")
        import :: c_int, c_ptr, c_double, c_long_long
        type(c_ptr), value :: h
        real(c_double), value :: phi_core, phi_min
        integer(c_int), value :: nquery, maxcomp
        integer(c_long_long), intent(in) :: query_cells(*)
        integer(c_int), intent(out) :: query_comp(*)
        real(c_double), intent(out) :: comp_mass(*)
        integer(c_int), intent(out) :: ncomp
    end function
    ! learned reinitialisation: the network (BF16 MFMA, FP32 accumulation) and its use
    integer(c_int) function c_lsmp_net_create(capacity, seed, net) bind(C, name="lsmp_net_create")
        import :: c_int, c_ptr
        integer(c_int), value :: capacity, seed
        type(c_ptr), intent(out) :: net
    end function
    subroutine c_lsmp_net_destroy(net) bind(C, name="lsmp_net_destroy")
        import :: c_ptr
        type(c_ptr), value :: net
    end subroutine
    integer(c_int) function c_lsmp_net_param_count() bind(C, name="lsmp_net_param_count")
        import :: c_int
    end function
    integer(c_int) function c_lsmp_net_set_params(net, params, n) bind(C, name="lsmp_net_set_params")
        import :: c_int, c_ptr, c_float
        type(c_ptr), value :: net
        real(c_float), intent(in) :: params(*)
        integer(c_int), value :: n
    end function
    integer(c_int) function c_lsmp_net_get_params(net, params, n) bind(C, name="lsmp_net_get_params")
        import :: c_int, c_ptr, c_float
        type(c_ptr), value :: net
        real(c_float), intent(out) :: params(*)
        integer(c_int), value :: n
    end function
    integer(c_int) function c_lsmp_net_dataset_size(net, n) bind(C, name="lsmp_net_dataset_size")
        import :: c_int, c_ptr
        type(c_ptr), value :: net
        integer(c_int), intent(out) :: n
    end function
    integer(c_int) function c_lsmp_net_collect(net, h, band_cells, stride, n_added) bind(C, name="lsmp_net_collect")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: net, h
        real(c_double), value :: band_cells
        integer(c_int), value :: stride
        integer(c_int), intent(out) :: n_added
    end function
    integer(c_int) function c_lsmp_net_add_samples(net, x, y, n) bind(C, name="lsmp_net_add_samples")
        import :: c_int, c_ptr, c_float
        type(c_ptr), value :: net
        real(c_float), intent(in) :: x(*), y(*)
        integer(c_int), value :: n
    end function
    integer(c_int) function c_lsmp_net_clear_samples(net) bind(C, name="lsmp_net_clear_samples")
        import :: c_int, c_ptr
        type(c_ptr), value :: net
    end function
    integer(c_int) function c_lsmp_net_apply(net, h, band_cells, n_band) bind(C, name="lsmp_net_apply")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: net, h
        real(c_double), value :: band_cells
        integer(c_int), intent(out) :: n_band
    end function
    integer(c_int) function c_lsmp_net_train(net, steps, batch, lr, seed, weight_c, loss) bind(C, name="lsmp_net_train")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: net
        integer(c_int), value :: steps, batch, seed
        real(c_double), value :: lr, weight_c
        real(c_double), intent(out) :: loss
    end function
    integer(c_int) function c_lsmp_net_eval_grad(net, x, y, n, loss, grad) bind(C, name="lsmp_net_eval_grad")
        import :: c_int, c_ptr, c_float, c_double
        type(c_ptr), value :: net
        real(c_float), intent(in) :: x(*), y(*)
        integer(c_int), value :: n
        real(c_double), intent(out) :: loss
        real(c_float), intent(out) :: grad(*)
    end function
    ! in place of c_lsmp_reinit(gpu, 1, ...): network, sweeps plain pseudo-steps, residual check, fallback
    integer(c_int) function c_lsmp_reinit_learned(h, net, band_cells, sweeps, residual_tol, residual_band, &
                                                  fallback_mode, residual, fallback, n_pseudo) &
                                                  bind(C, name="lsmp_reinit_learned")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h, net
        real(c_double), value :: band_cells, residual_tol, residual_band
        integer(c_int), value :: sweeps, fallback_mode
        real(c_double), intent(out) :: residual
        integer(c_int), intent(out) :: fallback, n_pseudo
    end function
    integer(c_int) function c_lsmp_reinit_residual(h, band_cells, residual) bind(C, name="lsmp_reinit_residual")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), value :: band_cells
        real(c_double), intent(out) :: residual
    end function
    ! CLSVOF (clsvof, interface_solver.f90:133-244)
    integer(c_int) function c_lsmp_upload_vof(h, vof) bind(C, name="lsmp_upload_vof")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), intent(in) :: vof(*)
    end function
    integer(c_int) function c_lsmp_download_vof(h, vof) bind(C, name="lsmp_download_vof")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), intent(out) :: vof(*)
    end function
    integer(c_int) function c_lsmp_store_vof_old(h) bind(C, name="lsmp_store_vof_old")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
    end function
    integer(c_int) function c_lsmp_store_vof0(h) bind(C, name="lsmp_store_vof0")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
    end function
    integer(c_int) function c_lsmp_vof_mass_vol(h, mass, vol) bind(C, name="lsmp_vof_mass_vol")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), intent(out) :: mass, vol
    end function
    integer(c_int) function c_lsmp_vof_interface_l1(h, l1) bind(C, name="lsmp_vof_interface_l1")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), intent(out) :: l1
    end function
    ! energy_and_momentum's integrals (output.f90:250-397), out(11)
    integer(c_int) function c_lsmp_energy_momentum(h, mu_12, vof_marker, out) bind(C, name="lsmp_energy_momentum")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), value :: mu_12
        integer(c_int), value :: vof_marker
        real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function c_lsmp_set_face_velocity(h, u, v, w) bind(C, name="lsmp_set_face_velocity")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), intent(in) :: u(*), v(*), w(*)
    end function
    integer(c_int) function c_lsmp_set_face_velocity_analytic(h, field, t, t2s) &
                                                              bind(C, name="lsmp_set_face_velocity_analytic")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        integer(c_int), value :: field
        real(c_double), value :: t, t2s
    end function
    integer(c_int) function c_lsmp_vof_wlic(h, dt) bind(C, name="lsmp_vof_wlic")
        import :: c_int, c_ptr, c_double
        type(c_ptr), value :: h
        real(c_double), value :: dt
    end function
    integer(c_int) function c_lsmp_clsvof_recon(h, n_pseudo0, n_pseudo1) bind(C, name="lsmp_clsvof_recon")
        import :: c_int, c_ptr
        type(c_ptr), value :: h
        integer(c_int), intent(out) :: n_pseudo0, n_pseudo1
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
        import :: c_size_t, c_ptr
        type(c_ptr), value :: s
    end function
end interface

public :: c_lsmp_init, c_lsmp_finalize, c_lsmp_upload_phi, c_lsmp_download_phi, c_lsmp_store_phi0
public :: c_lsmp_set_velocity, c_lsmp_set_velocity_analytic, c_lsmp_rk3_advect, c_lsmp_reinit
public :: c_lsmp_mass_vol, c_lsmp_mass_vol_noalloc, c_lsmp_mass_fix, c_lsmp_interface_l1, c_lsmp_get_stats, c_lsmp_synchronize
public :: c_lsmp_set_advection_scheme, c_lsmp_set_band_lines, c_lsmp_band_line_counts, c_lsmp_set_narrow_band
public :: c_lsmp_set_marker, c_lsmp_get_marker, c_lsmp_marker_load, c_lsmp_marker_store, c_lsmp_combine_markers
public :: c_lsmp_set_profiling, c_lsmp_kernel_time, c_lsmp_component_masses
public :: c_lsmp_net_create, c_lsmp_net_destroy, c_lsmp_net_param_count, c_lsmp_net_set_params
public :: c_lsmp_net_get_params, c_lsmp_net_dataset_size, c_lsmp_net_collect, c_lsmp_net_add_samples
public :: c_lsmp_net_clear_samples, c_lsmp_net_apply, c_lsmp_net_train, c_lsmp_net_eval_grad
public :: c_lsmp_reinit_learned, c_lsmp_reinit_residual
public :: c_lsmp_upload_vof, c_lsmp_download_vof, c_lsmp_store_vof_old, c_lsmp_set_face_velocity
public :: c_lsmp_set_face_velocity_analytic, c_lsmp_vof_wlic, c_lsmp_clsvof_recon
public :: c_lsmp_store_vof0, c_lsmp_vof_mass_vol, c_lsmp_vof_interface_l1, c_lsmp_energy_momentum
public :: lsmp_check, lsmp_error_message

contains

function lsmp_error_message(code) result(msg)
    integer(c_int), intent(in) :: code
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = c_lsmp_error_string(code)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
        msg(i:i) = chars(i)
    end do
end function

! Stop with the library's message on failure (the reference's `stop`).
subroutine lsmp_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    if (rc /= 0) then
        write(*,'(A,A,A,I0,A,A)') " lsmp error in ", where, ": code ", rc, ", ", lsmp_error_message(rc)
        stop 1
    endif
end subroutine

end module lsmp
