! The 5x5 known-answer test of the reference's algebra driver
! (src/plane_layer/driver_algebra.f90:24-47), run from Fortran through the
! C-ABI: band storage as zgbtrf takes it (kl = 1, ku = 2, ldab = 2kl+ku+1),
! factor and solve on the GPU, solution -1, 2, -3, 4, -5.  Exit status 0 when
! it matches to 1e-13, 1 otherwise.
program banded_kat
  use coral_amd_binding
  implicit none
  integer, parameter :: n = 5, kl = 1, ku = 2, ldab = 2 * kl + ku + 1
  complex(c_double_complex), target :: ab(ldab, n), x(n)
  integer(c_int32_t), target :: info(1)
  real(c_double) :: a(n, n), expect(n)
  integer :: i, j
  a = 0.d0
  a(1, 1:3) = [0.d0, 1.d0, 2.d0]
  a(2, 1:4) = [9.d0, 4.d0, 6.d0, 8.d0]
  a(3, 2:5) = [1.d0, -3.d0, 5.d0, 7.d0]
  a(4, 3:5) = [-2.d0, 0.d0, 4.d0]
  a(5, 4:5) = [1.d0, 3.d0]
  ab = (0.d0, 0.d0)
  do j = 1, n
     do i = max(1, j - ku), min(n, j + kl)
        ab(kl + ku + 1 + i - j, j) = cmplx(a(i, j), 0.d0, c_double_complex)
     end do
  end do
  x = cmplx([-4.d0, 13.d0, -4.d0, -14.d0, -11.d0], 0.d0, c_double_complex)
  expect = [-1.d0, 2.d0, -3.d0, 4.d0, -5.d0]
  call coral_amd_check(coral_amd_banded_solve(0_c_int32_t, int(n, c_int32_t), int(kl, c_int32_t), &
                                              int(ku, c_int32_t), 1_c_int32_t, c_loc(ab), c_loc(x), c_loc(info)))
  print '(a, i0)', 'info ', info(1)
  print '(a, 5f12.8)', 'x ', real(x)
  if (info(1) /= 0 .or. maxval(abs(x - expect)) > 1.d-13) then
     print '(a)', 'banded_kat: FAILED'
     stop 1
  end if
  print '(a)', 'banded_kat: ok'
end program banded_kat
