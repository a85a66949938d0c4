!> iso_c_binding interface to the HIP library (include/nuvort.h).
!>
!> The reference's hot-path procedures (src/run.f90:24-65 call sequence:
!> ghostp, construct_tree, pmotion, pinsert, pclose / pclose_tree, precon,
!> premove, enforce_boundary, empty_tree, NAN_finder) each have one nv_*
!> counterpart here.  Every nv_* returns an integer status (NV_OK = 0); the
!> caller maps a non-zero status to its fatal_error (src/cdata.f90:245-255),
!> which nv_check below does.
!>
!> Arrays cross the boundary as structure-of-arrays: never pass type(qvort)
!> itself, whose memory layout is processor-defined.  nv_state holds c_loc
!> pointers to caller arrays x(3,n), u(3,n), ... and infront(n), ...; the
!> library copies them inside nv_set_state / nv_get_state and keeps no pointer.
module nuvort_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: NV_OK = 0
  integer(c_int), parameter :: NV_VEL_LIA = 0, NV_VEL_BS = 1, NV_VEL_TREE = 2
  integer(c_int), parameter :: NV_BC_OPEN = 0, NV_BC_PERIODIC = 1, NV_BC_SOLID = 2
  integer(c_int), parameter :: NV_NF_ZERO = 0, NV_NF_XFLOW = 1, NV_NF_NOISEY_XFLOW = 2, NV_NF_SHEAR_XFLOW = 3
  integer(c_int), parameter :: NV_NF_YCHANNEL = 4, NV_NF_CHANNEL = 5
  integer(c_int), parameter :: NV_PREC_FP64 = 0, NV_PREC_FP32C = 1
  integer(c_int), parameter :: NV_CLOSE_AUTO = 0, NV_CLOSE_BRUTE = 1, NV_CLOSE_TREE = 2

  type, bind(C) :: nv_params
    real(c_double) :: delta, dt, quant_circ, corea
    real(c_double) :: box_size(3)
    integer(c_int32_t) :: boundary(3)
    integer(c_int32_t) :: velocity
    real(c_double) :: tree_theta
    integer(c_int32_t) :: tree_extra_correction
    integer(c_int32_t) :: normal_velocity
    real(c_double) :: norm_vel_xflow
    real(c_double) :: alpha(2)
    real(c_double) :: super_velocity(3)
    real(c_double) :: normal_fluid_cutoff, t_zero_normal_fluid
    real(c_double) :: norm_shear_omega
    integer(c_int32_t) :: recon_shots
    integer(c_int32_t) :: overide_timestep_check
    integer(c_int32_t) :: precision
    real(c_double) :: near_radius
    integer(c_int32_t) :: close_search
    integer(c_int32_t) :: exact_order
    integer(c_int32_t) :: capacity
    integer(c_int32_t) :: targets_per_lane
    integer(c_int32_t) :: kahan_block
    integer(c_int32_t) :: step_launch
    integer(c_int32_t) :: source_tile
  end type

  type, bind(C) :: nv_state
    integer(c_int32_t) :: pcount, capacity, itime, recon_count, remove_count, wall_recon_count
    real(c_double) :: t
    type(c_ptr) :: x, u, u1, u2, u_sup, u_mf, ghosti, ghostb
    type(c_ptr) :: infront, behind, closest
    type(c_ptr) :: closestd
    type(c_ptr) :: pinnedi, pinnedb
    type(c_ptr) :: wpinned
    type(c_ptr) :: curv, stretch
  end type

  !> calculate_diagnostics / the last pinsert (src/diagnostic.f90, src/line.f90:38-39,95)
  type, bind(C) :: nv_diag
    integer(c_int32_t) :: active, recon_count, remove_count, itime
    real(c_double) :: t, maxu, maxdu, mean_u_sup, mean_u_mf, kappa_bar, kappa_min, kappa_max
    real(c_double) :: total_length, avg_sep
    integer(c_int64_t) :: tree_eval
    integer(c_int32_t) :: wall_recon_count, pad
  end type

  interface
    integer(c_int) function nv_create(params, ctx) bind(C, name='nv_create')
      import :: nv_params, c_ptr, c_int
      type(nv_params), intent(in) :: params
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function nv_destroy(ctx) bind(C, name='nv_destroy')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function nv_last_error(ctx) bind(C, name='nv_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_set_state(ctx, st) bind(C, name='nv_set_state')
      import :: c_ptr, c_int, nv_state
      type(c_ptr), value :: ctx
      type(nv_state), intent(in) :: st
    end function
    integer(c_int) function nv_get_state(ctx, st) bind(C, name='nv_get_state')
      import :: c_ptr, c_int, nv_state
      type(c_ptr), value :: ctx
      type(nv_state), intent(inout) :: st
    end function
    integer(c_int) function nv_set_time(ctx, t, itime) bind(C, name='nv_set_time')
      import :: c_ptr, c_int, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: t
      integer(c_int32_t), value :: itime
    end function
    !> the noisey_xflow table that load_noise reads (src/normal_fluid.f90:156-162)
    integer(c_int) function nv_set_noise(ctx, n, noise) bind(C, name='nv_set_noise')
      import :: c_ptr, c_int, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
      real(c_double), dimension(*), intent(in) :: noise
    end function
    integer(c_int) function nv_ghostp(ctx) bind(C, name='nv_ghostp')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_construct_tree(ctx) bind(C, name='nv_construct_tree')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_pmotion(ctx) bind(C, name='nv_pmotion')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_pinsert(ctx) bind(C, name='nv_pinsert')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_pclose(ctx) bind(C, name='nv_pclose')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_pclose_tree(ctx) bind(C, name='nv_pclose_tree')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_precon(ctx) bind(C, name='nv_precon')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_wall_recon(ctx) bind(C, name='nv_wall_recon')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_premove(ctx) bind(C, name='nv_premove')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_enforce_boundary(ctx) bind(C, name='nv_enforce_boundary')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_empty_tree(ctx) bind(C, name='nv_empty_tree')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_nan_check(ctx) bind(C, name='nv_nan_check')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function nv_step(ctx, nsteps) bind(C, name='nv_step')
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function
    !> pclose on device arrays the caller holds, allocating nothing (x component-major,
    !> leading dimension ld; stream a hipStream_t or c_null_ptr); returns before the kernels end
    integer(c_int) function nv_pclose_noalloc(n, ld, x, infront, behind, closest, closestd, stream) &
        bind(C, name='nv_pclose_noalloc')
      import :: c_int, c_int32_t, c_ptr
      integer(c_int32_t), value :: n, ld
      type(c_ptr), value :: x, infront, behind, closest, closestd, stream
    end function
    integer(c_int) function nv_diagnostics(ctx, d) bind(C, name='nv_diagnostics')
      import :: c_ptr, c_int, nv_diag
      type(c_ptr), value :: ctx
      type(nv_diag), intent(out) :: d
    end function
    !> path arguments: c_null_char-terminated file names
    integer(c_int) function nv_data_dump(ctx, path) bind(C, name='nv_data_dump')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
    end function
    integer(c_int) function nv_data_restore(ctx, path) bind(C, name='nv_data_restore')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
    end function
    integer(c_int) function nv_printf(ctx, path) bind(C, name='nv_printf')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
    end function
    integer(c_int) function nv_print_info(ctx, d, shots, path) bind(C, name='nv_print_info')
      import :: c_ptr, c_int, c_char, c_int32_t, nv_diag
      type(c_ptr), value :: ctx
      type(nv_diag), intent(in) :: d
      integer(c_int32_t), value :: shots
      character(kind=c_char), dimension(*), intent(in) :: path
    end function
    integer(c_int) function nv_tree_print(ctx, path) bind(C, name='nv_tree_print')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: path
    end function
    !> cache_path: a c_null_char-terminated file name; c_null_char alone tunes without a cache
    integer(c_int) function nv_autotune(ctx, cache_path) bind(C, name='nv_autotune')
      import :: c_ptr, c_int, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: cache_path
    end function
  end interface

contains

  !> the message of the last error, as a Fortran string
  function nv_error_text(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: k
    msg = ' '
    call c_f_pointer(nv_last_error(ctx), cs, [256])
    do k = 1, 256
      if (cs(k) == c_null_char) exit
      msg(k:k) = cs(k)
    end do
  end function

  !> stop the run the way the reference's fatal_error does on a bad status
  subroutine nv_check(ctx, status, location)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: location
    if (status /= NV_OK) then
      write (*,*) '-------------------------FATAL ERROR-------------------------'
      write (*,*) trim(location), ': ', trim(nv_error_text(ctx)), ' (status ', status, ')'
      write (*,*) '-------------------------------------------------------------'
      stop 1
    end if
  end subroutine
end module
