!> A Fortran host that runs the reference's main loop (src/run.f90:24-65)
!> with every hot-path procedure replaced by its HIP C-ABI counterpart.
!>
!> usage: run_hip <state_in> <state_out> <nsteps>
!>   run.in is read from the working directory with the reference's parser
!>   semantics (src/cdata.f90:131-241: label up to the first blank,
!>   list-directed value, unknown labels skipped, the first read error ends
!>   parsing).  The state files use the layout documented in
!>   oracle/ref_harness.f90 (pcount, itime, recon_count, t, then every field
!>   of the point array as contiguous arrays).
program run_hip
  use iso_c_binding
  use nuvort_hip
  implicit none
  character(len=512) :: fin, fout, arg
  integer :: nrun, istep
  type(nv_params), target :: prm
  type(c_ptr) :: ctx
  integer(c_int) :: rc
  ! run.in values the loop itself needs
  integer :: recon_shots, shots
  logical :: tree_print
  real(c_double) :: dt, tree_theta
  character(len=40) :: print_file
  ! the point array as structure of arrays
  integer :: n, itime, recon_count
  real(c_double) :: t
  real(c_double), allocatable, target :: x(:,:), u(:,:), u1(:,:), u2(:,:), usup(:,:), umf(:,:), gi(:,:), gb(:,:)
  real(c_double), allocatable, target :: closestd(:)
  integer(c_int32_t), allocatable, target :: infront(:), behind(:), closest(:), pinnedi(:), pinnedb(:), wpinned(:,:)
  type(nv_state) :: st
  character(len=30) :: noise_file
  real(c_double), allocatable :: noise(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg) ; read(arg,*) nrun

  call read_params(prm, recon_shots, dt, tree_theta, noise_file, shots, tree_print)
  rc = nv_create(prm, ctx)
  if (rc /= NV_OK) then
    write (*,*) 'nv_create rejected the run.in parameters, status ', rc
    stop 1
  end if
  if (prm%normal_velocity == NV_NF_NOISEY_XFLOW) then
    ! load_noise (src/normal_fluid.f90:156-162): 100000 list-directed values
    allocate(noise(100000))
    open(unit=37, file=trim(noise_file), status='old')
    read(37,*) noise
    close(37)
    call nv_check(ctx, nv_set_noise(ctx, int(size(noise), c_int32_t), noise), 'normal_fluid.mod:load_noise')
  end if
  call load_state(trim(fin))
  call bind_state(n)
  call nv_check(ctx, nv_set_state(ctx, st), 'nv_set_state')

  do istep = 1, nrun
    itime = itime + 1
    call nv_check(ctx, nv_set_time(ctx, t, int(itime, c_int32_t)), 'set_time')
    call nv_check(ctx, nv_ghostp(ctx), 'boundary.mod:ghostp')
    if (tree_theta > 0) then
      call nv_check(ctx, nv_construct_tree(ctx), 'tree.mod:construct_tree')
      ! the leaf list build_tree appends every shots steps (src/tree.f90:96-102)
      if (tree_print .and. mod(itime, shots) == 0) then
        write(unit=print_file, fmt="(a,i3.3,a)") "./data/tree_mesh", itime/shots, ".log"
        call nv_check(ctx, nv_tree_print(ctx, trim(print_file)//c_null_char), 'tree.mod:build_tree (tree_print)')
      end if
    end if
    call nv_check(ctx, nv_pmotion(ctx), 'timestep.mod:pmotion')
    call nv_check(ctx, nv_pinsert(ctx), 'line.mod:pinsert')
    if (mod(itime, recon_shots) == 0) then
      if (tree_theta > 0) then
        call nv_check(ctx, nv_pclose_tree(ctx), 'tree.mod:pclose_tree')
      else
        call nv_check(ctx, nv_pclose(ctx), 'line.mod:pclose')
      end if
      call nv_check(ctx, nv_precon(ctx), 'reconnection.mod:precon')
      call nv_check(ctx, nv_wall_recon(ctx), 'reconnection.mod:wall_recon')
      call nv_check(ctx, nv_premove(ctx), 'line.mod:premove')
    end if
    call nv_check(ctx, nv_enforce_boundary(ctx), 'boundary.mod:enforce_boundary')
    t = t + dt
    if (tree_theta > 0) call nv_check(ctx, nv_empty_tree(ctx), 'tree.mod:empty_tree')
    call nv_check(ctx, nv_nan_check(ctx), 'general.mod:NAN_finder')
  end do

  ! the point array may have grown (pinsert): ask, resize, fetch
  rc = nv_get_state(ctx, st)
  if (rc == 3) then
    call resize_state(int(st%pcount))
    call bind_state(size(infront))
    rc = nv_get_state(ctx, st)
  end if
  call nv_check(ctx, rc, 'nv_get_state')
  n = st%pcount
  recon_count = st%recon_count
  call save_state(trim(fout))
  rc = nv_destroy(ctx)

contains

  !> the reference's run.in semantics for the keys this library reads
  subroutine read_params(p, rshots, dt_out, theta_out, nfile, shots_out, tprint)
    type(nv_params), intent(out) :: p
    integer, intent(out) :: rshots, shots_out
    logical, intent(out) :: tprint
    real(c_double), intent(out) :: dt_out, theta_out
    character(len=30), intent(out) :: nfile
    character(len=100) :: buffer, label
    character(len=40) :: bx, by, bz
    character(len=30) :: vel, nf, prec
    integer :: pos, ios, fh
    logical :: tec, otc
    real(c_double) :: delta, dtv, kappa, corea, box(3), theta, xflow, alpha(2), sup(3), cutoff, tzero, nearr
    real(c_double) :: omega
    ! defaults of src/cdata.f90:76-127
    delta = 0 ; dtv = 0 ; kappa = 9.97E-4_c_double ; corea = 8.244023E-9_c_double ; box = 0
    bx = 'periodic' ; by = 'periodic' ; bz = 'periodic' ; vel = ' ' ; nf = 'zero'
    theta = 0 ; tec = .true. ; xflow = 0.5_c_double ; alpha = 0 ; sup = 0
    cutoff = 1E8_c_double ; tzero = 1E8_c_double ; rshots = 1 ; otc = .false.
    shots_out = 1 ; tprint = .false.
    prec = 'fp64' ; nearr = 8 ; omega = 0 ; nfile = 'set_me!'
    ios = 0
    open(newunit=fh, file='run.in', status='old')
    do while (ios == 0)
      read(fh, '(A)', iostat=ios) buffer
      if (ios /= 0) exit
      pos = scan(buffer, ' ')
      label = buffer(1:pos)
      buffer = buffer(pos+1:)
      select case (label)
      case ('recon_shots') ; read(buffer, *, iostat=ios) rshots
      case ('shots') ; read(buffer, *, iostat=ios) shots_out
      case ('tree_print') ; read(buffer, *, iostat=ios) tprint
      case ('dt') ; read(buffer, *, iostat=ios) dtv
      case ('delta') ; read(buffer, *, iostat=ios) delta
      case ('quant_circ') ; read(buffer, *, iostat=ios) kappa
      case ('corea') ; read(buffer, *, iostat=ios) corea
      case ('box_size') ; read(buffer, *, iostat=ios) box
      case ('velocity') ; read(buffer, *, iostat=ios) vel
      case ('boundary_x') ; read(buffer, *, iostat=ios) bx
      case ('boundary_y') ; read(buffer, *, iostat=ios) by
      case ('boundary_z') ; read(buffer, *, iostat=ios) bz
      case ('tree_theta') ; read(buffer, *, iostat=ios) theta
      case ('tree_extra_correction') ; read(buffer, *, iostat=ios) tec
      case ('normal_velocity') ; read(buffer, *, iostat=ios) nf
      case ('normal_fluid_cutoff') ; read(buffer, *, iostat=ios) cutoff
      case ('t_zero_normal_fluid') ; read(buffer, *, iostat=ios) tzero
      case ('norm_vel_xflow') ; read(buffer, *, iostat=ios) xflow
      case ('norm_shear_omega') ; read(buffer, *, iostat=ios) omega
      case ('noise_file') ; read(buffer, *, iostat=ios) nfile
      case ('alpha') ; read(buffer, *, iostat=ios) alpha
      case ('super_velocity') ; read(buffer, *, iostat=ios) sup
      case ('overide_timestep_check') ; read(buffer, *, iostat=ios) otc
      case ('hip_precision') ; read(buffer, *, iostat=ios) prec
      case ('hip_near_radius') ; read(buffer, *, iostat=ios) nearr
      case default
      end select
    end do
    close(fh)
    p%delta = delta ; p%dt = dtv ; p%quant_circ = kappa ; p%corea = corea ; p%box_size = box
    p%boundary(1) = bc_code(bx) ; p%boundary(2) = bc_code(by) ; p%boundary(3) = bc_code(bz)
    select case (vel)
    case ('LIA') ; p%velocity = NV_VEL_LIA
    case ('BS') ; p%velocity = NV_VEL_BS
    case ('Tree') ; p%velocity = NV_VEL_TREE
    case default ; p%velocity = -1
    end select
    p%tree_theta = theta
    p%tree_extra_correction = merge(1, 0, tec)
    select case (nf)
    case ('zero') ; p%normal_velocity = NV_NF_ZERO
    case ('xflow') ; p%normal_velocity = NV_NF_XFLOW
    case ('noisey_xflow') ; p%normal_velocity = NV_NF_NOISEY_XFLOW
    case ('shear_xflow') ; p%normal_velocity = NV_NF_SHEAR_XFLOW
    case ('ychannel') ; p%normal_velocity = NV_NF_YCHANNEL
    case ('channel') ; p%normal_velocity = NV_NF_CHANNEL
    case default ; p%normal_velocity = 9
    end select
    p%norm_shear_omega = omega
    p%norm_vel_xflow = xflow ; p%alpha = alpha ; p%super_velocity = sup
    p%normal_fluid_cutoff = cutoff ; p%t_zero_normal_fluid = tzero
    p%recon_shots = rshots
    p%overide_timestep_check = merge(1, 0, otc)
    p%precision = merge(NV_PREC_FP32C, NV_PREC_FP64, prec == 'fp32c')
    p%near_radius = nearr
    p%close_search = NV_CLOSE_AUTO
    p%exact_order = 0 ; p%capacity = 0 ; p%targets_per_lane = 0 ; p%kahan_block = 0
    p%step_launch = 0 ; p%source_tile = 0
    dt_out = dtv ; theta_out = theta
  end subroutine

  !> run.in's boundary word as NV_BC_* (9 for a word init_setup rejects)
  integer(c_int) function bc_code(word)
    character(len=*), intent(in) :: word
    select case (trim(word))
    case ('periodic') ; bc_code = NV_BC_PERIODIC
    case ('open') ; bc_code = NV_BC_OPEN
    case ('solid') ; bc_code = NV_BC_SOLID
    case default ; bc_code = 9
    end select
  end function

  subroutine resize_state(m)
    integer, intent(in) :: m
    if (allocated(x)) deallocate(x, u, u1, u2, usup, umf, gi, gb, closestd, infront, behind, closest, &
                                 pinnedi, pinnedb, wpinned)
    allocate(x(3,m), u(3,m), u1(3,m), u2(3,m), usup(3,m), umf(3,m), gi(3,m), gb(3,m), closestd(m))
    allocate(infront(m), behind(m), closest(m), pinnedi(m), pinnedb(m), wpinned(3,m))
  end subroutine

  subroutine bind_state(cap)
    integer, intent(in) :: cap
    st%pcount = n ; st%capacity = cap ; st%itime = itime ; st%recon_count = recon_count
    st%remove_count = 0 ; st%wall_recon_count = 0 ; st%t = t
    st%x = c_loc(x) ; st%u = c_loc(u) ; st%u1 = c_loc(u1) ; st%u2 = c_loc(u2)
    st%u_sup = c_loc(usup) ; st%u_mf = c_loc(umf) ; st%ghosti = c_loc(gi) ; st%ghostb = c_loc(gb)
    st%infront = c_loc(infront) ; st%behind = c_loc(behind) ; st%closest = c_loc(closest)
    st%closestd = c_loc(closestd) ; st%pinnedi = c_loc(pinnedi) ; st%pinnedb = c_loc(pinnedb)
    st%wpinned = c_loc(wpinned)
  end subroutine

  subroutine read_vec(u_, a)
    integer, intent(in) :: u_
    real(c_double), intent(inout) :: a(:,:)
    integer :: k
    do k = 1, 3
      read(u_) a(k,:)
    end do
  end subroutine

  subroutine load_state(fname)
    character(len=*), intent(in) :: fname
    integer :: fh, k
    open(newunit=fh, file=fname, access='stream', form='unformatted', status='old')
    read(fh) n, itime, recon_count, t
    call resize_state(n)
    call read_vec(fh, x) ; call read_vec(fh, u) ; call read_vec(fh, u1) ; call read_vec(fh, u2)
    call read_vec(fh, usup) ; call read_vec(fh, umf) ; call read_vec(fh, gi) ; call read_vec(fh, gb)
    read(fh) infront ; read(fh) behind ; read(fh) closest ; read(fh) closestd
    read(fh) pinnedi ; read(fh) pinnedb
    do k = 1, 3
      read(fh) wpinned(k,:)
    end do
    close(fh)
  end subroutine

  subroutine save_state(fname)
    character(len=*), intent(in) :: fname
    integer :: fh, k
    open(newunit=fh, file=fname, access='stream', form='unformatted', status='replace')
    write(fh) int(n, c_int32_t), int(itime, c_int32_t), int(recon_count, c_int32_t), t
    do k = 1, 3 ; write(fh) x(k,1:n) ; end do
    do k = 1, 3 ; write(fh) u(k,1:n) ; end do
    do k = 1, 3 ; write(fh) u1(k,1:n) ; end do
    do k = 1, 3 ; write(fh) u2(k,1:n) ; end do
    do k = 1, 3 ; write(fh) usup(k,1:n) ; end do
    do k = 1, 3 ; write(fh) umf(k,1:n) ; end do
    do k = 1, 3 ; write(fh) gi(k,1:n) ; end do
    do k = 1, 3 ; write(fh) gb(k,1:n) ; end do
    write(fh) infront(1:n) ; write(fh) behind(1:n) ; write(fh) closest(1:n) ; write(fh) closestd(1:n)
    write(fh) pinnedi(1:n) ; write(fh) pinnedb(1:n)
    do k = 1, 3 ; write(fh) wpinned(k,1:n) ; end do
    close(fh)
  end subroutine
end program
