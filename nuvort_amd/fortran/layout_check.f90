!> Prints the byte size and field offsets of the bind(C) types of
!> nuvort_hip, so a test can compare them with the C structs of nuvort.h.
program layout_check
  use iso_c_binding
  use nuvort_hip
  implicit none
  type(nv_params), target :: p
  type(nv_state), target :: s
  write(*,'(a,i0)') 'nv_params ', c_sizeof(p)
  write(*,'(a,i0)') 'nv_state ', c_sizeof(s)
  write(*,'(a,i0)') 'nv_params.precision ', loc(p%precision) - loc(p)
  write(*,'(a,i0)') 'nv_params.kahan_block ', loc(p%kahan_block) - loc(p)
  write(*,'(a,i0)') 'nv_params.step_launch ', loc(p%step_launch) - loc(p)
  write(*,'(a,i0)') 'nv_params.source_tile ', loc(p%source_tile) - loc(p)
  write(*,'(a,i0)') 'nv_state.t ', loc(s%t) - loc(s)
  write(*,'(a,i0)') 'nv_state.wpinned ', loc(s%wpinned) - loc(s)
end program
