! desc_layout.f90 -- the bind(C) layout of the descriptor types of
! wql3d_hip_binding: sizes and field offsets, printed one per line as
! "name value" for comparison with the C compiler's sizeof / offsetof.
program desc_layout
  use iso_c_binding
  use wql3d_hip_binding
  implicit none
  type(wql_block_desc), target :: d
  type(wql_source), target :: s
  type(wql_iface_desc), target :: i
  integer(c_intptr_t) :: b
  b = transfer(c_loc(d), b)
  write(*, '(A,I0)') 'sizeof_block_desc ', c_sizeof(d)
  write(*, '(A,I0)') 'bc ', transfer(c_loc(d%bc), b) - b
  write(*, '(A,I0)') 'arrays_on_device ', transfer(c_loc(d%arrays_on_device), b) - b
  write(*, '(A,I0)') 'hq ', transfer(c_loc(d%hq), b) - b
  write(*, '(A,I0)') 'tau ', transfer(c_loc(d%tau), b) - b
  write(*, '(A,I0)') 'weight ', transfer(c_loc(d%weight), b) - b
  write(*, '(A,I0)') 'body ', transfer(c_loc(d%body), b) - b
  write(*, '(A,I0)') 'metricx ', transfer(c_loc(d%metricx), b) - b
  write(*, '(A,I0)') 'F ', transfer(c_loc(d%F), b) - b
  write(*, '(A,I0)') 'eta ', transfer(c_loc(d%eta), b) - b
  write(*, '(A,I0)') 'Deta ', transfer(c_loc(d%Deta), b) - b
  write(*, '(A,I0)') 'Qs_inv ', transfer(c_loc(d%Qs_inv), b) - b
  write(*, '(A,I0)') 'pml_box ', transfer(c_loc(d%pml_box), b) - b
  write(*, '(A,I0)') 'pml_Q ', transfer(c_loc(d%pml_Q), b) - b
  write(*, '(A,I0)') 'pml_DQ ', transfer(c_loc(d%pml_DQ), b) - b
  write(*, '(A,I0)') 'face_box ', transfer(c_loc(d%face_box), b) - b
  write(*, '(A,I0)') 'n_l ', transfer(c_loc(d%n_l), b) - b
  write(*, '(A,I0)') 'n_n ', transfer(c_loc(d%n_n), b) - b
  write(*, '(A,I0)') 'sat_form ', transfer(c_loc(d%sat_form), b) - b
  b = transfer(c_loc(s), b)
  write(*, '(A,I0)') 'sizeof_source ', c_sizeof(s)
  write(*, '(A,I0)') 'src_duration ', transfer(c_loc(s%duration), b) - b
  write(*, '(A,I0)') 'src_location ', transfer(c_loc(s%location), b) - b
  write(*, '(A,I0)') 'src_near_phys ', transfer(c_loc(s%near_phys), b) - b
  b = transfer(c_loc(i), b)
  write(*, '(A,I0)') 'sizeof_iface_desc ', c_sizeof(i)
  write(*, '(A,I0)') 'iface_y ', transfer(c_loc(i%y), b) - b
  write(*, '(A,I0)') 'iface_Dc ', transfer(c_loc(i%Dc), b) - b
  write(*, '(A,I0)') 'iface_W0 ', transfer(c_loc(i%W0), b) - b
  write(*, '(A,I0)') 'iface_alpha ', transfer(c_loc(i%alpha), b) - b
end program desc_layout
