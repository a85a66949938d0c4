!> Test infrastructure only (the oracle's reference build): a driver of this
!> project's own that links the UNMODIFIED NuVort hot-path modules compiled
!> straight from the reference tree (cdata, derivatives, general, boundary,
!> normal_fluid, tree, timestep, reconnection, line) and replays the per-step
!> call sequence of src/run.f90:24-65 on a state read from a file.
!>
!> It exists so that golden fixtures and calibration timings come from the
!> reference's own FP64 arithmetic.  Nothing in the product links it.
!>
!> usage: nuvort_ref <state_in> <out_prefix> <nsteps> <dump_phases 0|1> [output 0|1]
!>   run.in is read from the working directory by the reference's own parser
!>   (src/cdata.f90:131-241); ./data must exist (precon appends to
!>   ./data/recon_location.log, src/reconnection.f90:47-49).
!>   A state_in named *.dat is read as the reference's restart file, with the
!>   statements of data_restore (src/initial.f90:151-166).
!>   With output = 1 every step ends with the reference's output phase
!>   (src/run.f90:50-56): calculate_diagnostics, then, when mod(itime,shots)
!>   is 0, the statements of data_dump (./data/var.dat, src/output.f90:85-94),
!>   of printf (./data/var%04d.log, :47-70) and print_info's ts.log line
!>   (./data/ts.log, :23-37); output.f90 itself is not compiled (its banner
!>   calls GNU extensions).
!>
!> State file (little-endian stream, also written by oracle/ref_io.py):
!>   int32 pcount, int32 itime, int32 recon_count, real8 t, then per field
!>   arrays of length pcount: x,y,z, u(3), u1(3), u2(3), u_sup(3), u_mf(3),
!>   ghosti(3), ghostb(3) as real8; infront, behind, closest as int32;
!>   closestd real8; pinnedi, pinnedb int32 (0/1); wpinned(3) int32.
program ref_harness
  use cdata
  use boundary
  use timestep
  use tree
  use line
  use reconnection
  use normal_fluid
  use diagnostic
  implicit none
  character(len=512) :: state_in, out_prefix, arg
  integer :: nrun, dump_phases, istep, itime0, do_output
  real(8) :: t0, t1, t_vel, t_tree, t_close, ta, tb

  call get_command_argument(1, state_in)
  call get_command_argument(2, out_prefix)
  call get_command_argument(3, arg) ; read(arg,*) nrun
  call get_command_argument(4, arg) ; read(arg,*) dump_phases
  do_output = 0
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg) ; read(arg,*) do_output
  end if

  call read_run_file()
  ! boundary bookkeeping exactly as src/initial.f90:37-78 does it
  if (boundary_x=='periodic') n_periodic=n_periodic+1
  if (boundary_y=='periodic') n_periodic=n_periodic+1
  if (boundary_z=='periodic') n_periodic=n_periodic+1
  if (boundary_x=='solid') n_mirror=n_mirror+1
  if (boundary_y=='solid') n_mirror=n_mirror+1
  if (boundary_z=='solid') n_mirror=n_mirror+1
  call create_periodic_loop_array
  call create_mirror_loop_array
  ! the normal-fluid models with set-up (the noise table of noisey_xflow is
  ! read here from run.in's noise_file), as src/initial.f90 calls it
  if (normal_velocity/='zero' .and. normal_velocity/='xflow') call setup_normal_fluid

  if (index(state_in, '.dat') == len_trim(state_in) - 3) then
    call restore_var(trim(state_in))
  else
    call load_state(trim(state_in))
  end if
  itime0 = itime
  t_vel = 0.d0 ; t_tree = 0.d0 ; t_close = 0.d0
  t0 = omp_get_wtime()
  do istep = 1, nrun
    itime = itime0 + istep
    call ghostp
    if (dump_phases==1 .and. istep==1) call dump_state(trim(out_prefix)//'_ghost.bin')
    if (tree_theta>0) then
      ta = omp_get_wtime()
      call construct_tree
      tb = omp_get_wtime() ; t_tree = t_tree + (tb-ta)
      if (dump_phases==1 .and. istep==1) call dump_tree(trim(out_prefix)//'_tree.bin')
    end if
    ta = omp_get_wtime()
    call pmotion
    tb = omp_get_wtime() ; t_vel = t_vel + (tb-ta)
    if (dump_phases==1 .and. istep==1) call dump_state(trim(out_prefix)//'_motion.bin')
    call pinsert
    if (dump_phases==1 .and. istep==1) call dump_state(trim(out_prefix)//'_insert.bin')
    if (mod(itime, recon_shots)==0) then
      ta = omp_get_wtime()
      if (tree_theta>0) then
        call pclose_tree
      else
        call pclose
      end if
      tb = omp_get_wtime() ; t_close = t_close + (tb-ta)
      if (dump_phases==1 .and. istep==1) call dump_state(trim(out_prefix)//'_close.bin')
      call precon
      if (dump_phases==1 .and. istep==1) call dump_state(trim(out_prefix)//'_recon.bin')
      call wall_recon
      if (dump_phases==1 .and. istep==1) call dump_state(trim(out_prefix)//'_wall.bin')
      call premove
      if (dump_phases==1 .and. istep==1) call dump_state(trim(out_prefix)//'_remove.bin')
    end if
    call enforce_boundary
    t = t + dt
    if (do_output == 1) then
      call calculate_diagnostics
      if (mod(itime, shots) == 0) then
        call dump_var
        call write_var_log(itime / shots)
        call write_ts_line
      end if
    end if
    if (tree_theta>0) then
      call empty_tree(vtree)
      deallocate(vtree%parray) ; deallocate(vtree)
      tree_eval = 0 ; nullify(vtree)
    end if
    if (NAN_test) call NAN_finder
  end do
  t1 = omp_get_wtime()
  call dump_state(trim(out_prefix)//'_final.bin')
  open(unit=71, file=trim(out_prefix)//'_timing.txt', status='replace')
  write(71,'(a,i0)') 'threads ', omp_get_max_threads()
  write(71,'(a,i0)') 'steps ', nrun
  write(71,'(a,es24.16)') 'wall_s ', t1-t0
  write(71,'(a,es24.16)') 'velocity_s ', t_vel
  write(71,'(a,es24.16)') 'tree_build_s ', t_tree
  write(71,'(a,es24.16)') 'close_search_s ', t_close
  write(71,'(a,i0)') 'recon_count ', recon_count
  write(71,'(a,i0)') 'remove_count ', remove_count
  write(71,'(a,i0)') 'wall_recon_count ', wall_recon_count
  write(71,'(a,i0)') 'pcount ', pcount
  close(71)

contains

  !> src/initial.f90:151-166 data_restore's statements
  subroutine restore_var(fname)
    character(len=*), intent(in) :: fname
    integer :: dummy_itime
    open(unit=63, file=fname, FORM='unformatted')
      read(63) pcount
      read(63) recon_count
      read(63) dummy_itime
      read(63) t
      allocate(f(pcount))
      read(63) f
    close(63)
    itime = dummy_itime
  end subroutine

  !> src/output.f90:85-94 data_dump's statements
  subroutine dump_var
    open(unit=53, file="./data/var.dat", FORM='unformatted', status='replace')
      write(53) pcount
      write(53) recon_count
      write(53) itime
      write(53) t
      write(53) f
    close(53)
  end subroutine

  !> src/output.f90:47-70 printf's statements
  subroutine write_var_log(filenumber)
    integer, intent(in) :: filenumber
    character(len=40) :: print_file
    write(unit=print_file, fmt="(a,i4.4,a)") "./data/var", filenumber, ".log"
    open(unit=98, file=print_file, status='replace', form='unformatted', access='stream')
      write(98) t
      write(98) pcount
      write(98) f(:)%x(1)
      write(98) f(:)%x(2)
      write(98) f(:)%x(3)
      write(98) f(:)%infront
      write(98) f(:)%u(1)
      write(98) f(:)%u(2)
      write(98) f(:)%u(3)
      write(98) f(:)%u_mf(1)
      write(98) f(:)%u_mf(2)
      write(98) f(:)%u_mf(3)
      write(98) f(:)%curv
      write(98) f(:)%stretch
    close(98)
  end subroutine

  !> src/output.f90:35-37 print_info's ts.log line (the header is cosmetic)
  subroutine write_ts_line
    open(unit=78, file='data/ts.log', position='append')
    write(78,'(i6.4,f13.7,i10.1,i13.1,i13.1,f7.4,f13.6,f13.5,f13.5,f10.2,i13.1,i13.1)') &
      itime/shots, t, count(mask=f(:)%infront>0), recon_count, wall_recon_count, avg_sep/delta, &
      total_length, maxu, maxdu, kappa_bar, remove_count, tree_eval
    close(78)
    open(unit=79, file='data/diag.log', position='append')
    write(79,'(10es26.17e3)') maxu, maxdu, mean_u_sup, mean_u_mf, kappa_bar, kappa_min, kappa_max, &
      total_length, avg_sep, real(tree_eval)
    close(79)
  end subroutine

  subroutine load_state(fname)
    character(len=*), intent(in) :: fname
    integer :: n, k
    real(8), allocatable :: r(:)
    integer, allocatable :: iv(:)
    open(unit=61, file=fname, access='stream', form='unformatted', status='old')
    read(61) n
    read(61) itime
    read(61) recon_count
    read(61) t
    pcount = n
    allocate(f(n), r(n), iv(n))
    do k = 1, 3 ; read(61) r ; f(:)%x(k) = r ; end do
    do k = 1, 3 ; read(61) r ; f(:)%u(k) = r ; end do
    do k = 1, 3 ; read(61) r ; f(:)%u1(k) = r ; end do
    do k = 1, 3 ; read(61) r ; f(:)%u2(k) = r ; end do
    do k = 1, 3 ; read(61) r ; f(:)%u_sup(k) = r ; end do
    do k = 1, 3 ; read(61) r ; f(:)%u_mf(k) = r ; end do
    do k = 1, 3 ; read(61) r ; f(:)%ghosti(k) = r ; end do
    do k = 1, 3 ; read(61) r ; f(:)%ghostb(k) = r ; end do
    read(61) iv ; f(:)%infront = iv
    read(61) iv ; f(:)%behind = iv
    read(61) iv ; f(:)%closest = iv
    read(61) r ; f(:)%closestd = r
    read(61) iv ; f(:)%pinnedi = (iv /= 0)
    read(61) iv ; f(:)%pinnedb = (iv /= 0)
    do k = 1, 3 ; read(61) iv ; f(:)%wpinned(k) = iv ; end do
    f(:)%curv = 0.d0 ; f(:)%stretch = 0.d0
    close(61)
    deallocate(r, iv)
  end subroutine

  subroutine dump_state(fname)
    character(len=*), intent(in) :: fname
    integer :: k, i
    open(unit=62, file=fname, access='stream', form='unformatted', status='replace')
    write(62) pcount
    write(62) itime
    write(62) recon_count
    write(62) t
    do k = 1, 3 ; write(62) (f(i)%x(k), i=1,pcount) ; end do
    do k = 1, 3 ; write(62) (f(i)%u(k), i=1,pcount) ; end do
    do k = 1, 3 ; write(62) (f(i)%u1(k), i=1,pcount) ; end do
    do k = 1, 3 ; write(62) (f(i)%u2(k), i=1,pcount) ; end do
    do k = 1, 3 ; write(62) (f(i)%u_sup(k), i=1,pcount) ; end do
    do k = 1, 3 ; write(62) (f(i)%u_mf(k), i=1,pcount) ; end do
    do k = 1, 3 ; write(62) (f(i)%ghosti(k), i=1,pcount) ; end do
    do k = 1, 3 ; write(62) (f(i)%ghostb(k), i=1,pcount) ; end do
    write(62) (f(i)%infront, i=1,pcount)
    write(62) (f(i)%behind, i=1,pcount)
    write(62) (f(i)%closest, i=1,pcount)
    write(62) (f(i)%closestd, i=1,pcount)
    write(62) (merge(1,0,f(i)%pinnedi), i=1,pcount)
    write(62) (merge(1,0,f(i)%pinnedb), i=1,pcount)
    do k = 1, 3 ; write(62) (f(i)%wpinned(k), i=1,pcount) ; end do
    close(62)
  end subroutine

  !> every node of vtree in depth-first pre-order with the reference's child
  !> order (src/tree.f90:311-314): one record of 10 real8 + 2 int32 each.
  subroutine dump_tree(fname)
    character(len=*), intent(in) :: fname
    open(unit=63, file=fname, access='stream', form='unformatted', status='replace')
    call dump_node(vtree)
    close(63)
  end subroutine

  recursive subroutine dump_node(nd)
    type(node), pointer :: nd
    integer :: leaf_infront
    leaf_infront = 0
    if (nd%pcount >= 1) leaf_infront = nd%parray(1)%infront
    write(63) nd%posx, nd%posy, nd%posz, nd%width, nd%centx, nd%centy, nd%centz, &
              nd%circ(1), nd%circ(2), nd%circ(3), nd%pcount, leaf_infront
    if (nd%pcount > 1) then
      call dump_node(nd%fbl) ; call dump_node(nd%fbr)
      call dump_node(nd%ftl) ; call dump_node(nd%ftr)
      call dump_node(nd%bbl) ; call dump_node(nd%bbr)
      call dump_node(nd%btl) ; call dump_node(nd%btr)
    end if
  end subroutine
end program
