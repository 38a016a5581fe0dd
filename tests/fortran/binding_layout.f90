! Layout of the bind(C) types of coral_amd_binding: "name size" and
! "name.field offset" lines, compared by the tests with the C structs (the
! ctypes mirror of coral_amd.h).
program binding_layout
  use coral_amd_binding
  implicit none
  type(coral_amd_piece), target :: p
  type(coral_amd_nlterm), target :: t
  type(coral_amd_system), target :: s
  type(coral_amd_fullterm), target :: ft
  type(coral_amd_fullvar), target :: fv
  type(coral_amd_quadvar), target :: q
  type(coral_amd_problem), target :: pr
  print '(a, i0)', 'coral_amd_piece ', c_sizeof(p)
  print '(a, i0)', 'coral_amd_nlterm ', c_sizeof(t)
  print '(a, i0)', 'coral_amd_system ', c_sizeof(s)
  print '(a, i0)', 'coral_amd_fullterm ', c_sizeof(ft)
  print '(a, i0)', 'coral_amd_fullvar ', c_sizeof(fv)
  print '(a, i0)', 'coral_amd_quadvar ', c_sizeof(q)
  print '(a, i0)', 'coral_amd_problem ', c_sizeof(pr)
  call off('coral_amd_piece.eqn', c_loc(p), c_loc(p%eqn))
  call off('coral_amd_nlterm.qvar', c_loc(t), c_loc(t%qvar))
  call off('coral_amd_system.bc_codes', c_loc(s), c_loc(s%bc_codes))
  call off('coral_amd_system.mass', c_loc(s), c_loc(s%mass))
  call off('coral_amd_system.stif', c_loc(s), c_loc(s%stif))
  call off('coral_amd_system.nl', c_loc(s), c_loc(s%nl))
  call off('coral_amd_fullterm.position', c_loc(ft), c_loc(ft%position))
  call off('coral_amd_fullvar.terms', c_loc(fv), c_loc(fv%terms))
  call off('coral_amd_problem.Lx', c_loc(pr), c_loc(pr%Lx))
  call off('coral_amd_problem.nstages', c_loc(pr), c_loc(pr%nstages))
  call off('coral_amd_problem.a_arr', c_loc(pr), c_loc(pr%a_arr))
  call off('coral_amd_problem.kxky', c_loc(pr), c_loc(pr%kxky))
  call off('coral_amd_problem.zero', c_loc(pr), c_loc(pr%zero))
  call off('coral_amd_problem.full', c_loc(pr), c_loc(pr%full))
  call off('coral_amd_problem.quad', c_loc(pr), c_loc(pr%quad))
contains
  subroutine off(name, base, field)
    character(*), intent(in) :: name
    type(c_ptr), intent(in) :: base, field
    print '(a, a, i0)', name, ' ', transfer(field, 0_c_intptr_t) - transfer(base, 0_c_intptr_t)
  end subroutine
end program binding_layout
