! A Fortran host driving whole time steps through the C-ABI, the way the
! reference's main loop would (plane_layer_main.f90:147,157): create the
! context from a parsed recipe, factorize for dt, upload the state, step,
! download.  The parsed recipe (what PL_equations.f90 produces) arrives in a
! small binary file, since the reference's parser is not part of this build:
!
!   problem file (stream, native byte order):
!     int32 NX, NY, NZ; real64 Lx, Ly, gap, center; int32 s;
!     real64 a_arr(s*s), a_hat((s+1)*(s+1))                (row-major)
!     int32 n_kxky, then per system: int32 nvars, bc(nvars), orders(nvars),
!       int32 n_mass, pieces; int32 n_stif, pieces; int32 n_nl, NL terms
!     int32 n_zero, the same per mean system
!     int32 n_full, then per full variable: int32 nterms, terms
!     int32 n_quad, quadratic variables
!   (a piece / NL term: real64 + 6 int32; a full term: real64 + 7 int32; a
!   quadratic variable: 3 int32)
!   state file: real64 dt; int32 nsteps; int32 n_kxky, then per system int64
!   count and count complex128 values (the layout coral_amd_upload_state
!   takes); int32 n_zero, then per system int64 count and count real64 values
!
! Usage: step_driver <problem file> <state file> <output file>.  The output
! file has the state file's layout after nsteps steps (dt and nsteps as given),
! followed by real64 CFL time step of the last stage.
program step_driver
  use coral_amd_binding
  implicit none

  type :: sysbuf
     integer(c_int32_t), allocatable :: bc(:), eo(:)
     type(coral_amd_piece), allocatable :: mass(:), stif(:)
     type(coral_amd_nlterm), allocatable :: nl(:)
  end type
  type :: fullbuf
     type(coral_amd_fullterm), allocatable :: terms(:)
  end type
  type :: kstate
     complex(c_double_complex), allocatable :: v(:)
  end type
  type :: zstate
     real(c_double), allocatable :: v(:)
  end type

  type(coral_amd_problem) :: prob
  type(sysbuf), allocatable, target :: kb(:), zb(:)
  type(coral_amd_system), allocatable, target :: ks(:), zs(:)
  type(fullbuf), allocatable, target :: fb(:)
  type(coral_amd_fullvar), allocatable, target :: fv(:)
  type(coral_amd_quadvar), allocatable, target :: qv(:)
  real(c_double), allocatable, target :: a_arr(:), a_hat(:)
  type(kstate), allocatable, target :: kx(:)
  type(zstate), allocatable, target :: zx(:)
  character(len=1024) :: fprob, fstate, fout
  integer(c_int32_t) :: NX, NY, NZ, s, n, i, nsteps
  integer(c_int64_t) :: cnt
  real(c_double) :: Lx, Ly, gap, center, dt, cfl
  type(c_ptr) :: ctx
  integer :: u

  call get_command_argument(1, fprob)
  call get_command_argument(2, fstate)
  call get_command_argument(3, fout)

  open (newunit=u, file=trim(fprob), access='stream', form='unformatted', status='old')
  read (u) NX, NY, NZ, Lx, Ly, gap, center, s
  allocate (a_arr(s * s), a_hat((s + 1) * (s + 1)))
  read (u) a_arr, a_hat
  read (u) n
  call read_systems(u, n, kb, ks)
  prob%n_kxky = n
  read (u) n
  call read_systems(u, n, zb, zs)
  prob%n_zero = n
  read (u) n
  allocate (fb(max(n, 1)), fv(max(n, 1)))
  do i = 1, n
     read (u) fv(i)%nterms
     allocate (fb(i)%terms(max(fv(i)%nterms, 1)))
     read (u) fb(i)%terms(1:fv(i)%nterms)
     fv(i)%terms = c_loc(fb(i)%terms)
  end do
  prob%n_full = n
  read (u) n
  allocate (qv(max(n, 1)))
  read (u) qv(1:n)
  prob%n_quad = n
  close (u)

  prob%NX = NX; prob%NY = NY; prob%NZ = NZ
  prob%Lx = Lx; prob%Ly = Ly; prob%gap = gap; prob%center = center
  prob%nstages = s
  prob%a_arr = c_loc(a_arr); prob%a_hat = c_loc(a_hat)
  prob%kxky = c_loc(ks); prob%zero = c_loc(zs)
  prob%full = c_loc(fv); prob%quad = c_loc(qv)

  open (newunit=u, file=trim(fstate), access='stream', form='unformatted', status='old')
  read (u) dt, nsteps
  read (u) n
  allocate (kx(max(n, 1)))
  do i = 1, n
     read (u) cnt
     allocate (kx(i)%v(cnt))
     read (u) kx(i)%v
  end do
  read (u) n
  allocate (zx(max(n, 1)))
  do i = 1, n
     read (u) cnt
     allocate (zx(i)%v(cnt))
     read (u) zx(i)%v
  end do
  close (u)

  call coral_amd_check(coral_amd_create(prob, 0_c_int, ctx))
  call coral_amd_check(coral_amd_set_dt(ctx, dt))
  do i = 1, prob%n_kxky
     call coral_amd_check(coral_amd_upload_state(ctx, 0_c_int32_t, i - 1, c_loc(kx(i)%v)))
  end do
  do i = 1, prob%n_zero
     call coral_amd_check(coral_amd_upload_state(ctx, 1_c_int32_t, i - 1, c_loc(zx(i)%v)))
  end do
  call coral_amd_check(coral_amd_step(ctx, nsteps))
  do i = 1, prob%n_kxky
     call coral_amd_check(coral_amd_download_state(ctx, 0_c_int32_t, i - 1, c_loc(kx(i)%v)))
  end do
  do i = 1, prob%n_zero
     call coral_amd_check(coral_amd_download_state(ctx, 1_c_int32_t, i - 1, c_loc(zx(i)%v)))
  end do
  call coral_amd_check(coral_amd_get_cfl_dt(ctx, cfl))
  call coral_amd_check(coral_amd_destroy(ctx))

  open (newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
  write (u) dt, nsteps
  write (u) prob%n_kxky
  do i = 1, prob%n_kxky
     write (u) int(size(kx(i)%v), c_int64_t)
     write (u) kx(i)%v
  end do
  write (u) prob%n_zero
  do i = 1, prob%n_zero
     write (u) int(size(zx(i)%v), c_int64_t)
     write (u) zx(i)%v
  end do
  write (u) cfl
  close (u)
  print '(a, i0, a)', 'step_driver: ', nsteps, ' steps done'

contains

  subroutine read_systems(u, n, b, sy)
    integer, intent(in) :: u
    integer(c_int32_t), intent(in) :: n
    type(sysbuf), allocatable, target, intent(inout) :: b(:)
    type(coral_amd_system), allocatable, target, intent(inout) :: sy(:)
    integer :: k
    allocate (b(max(n, 1)), sy(max(n, 1)))
    do k = 1, n
       read (u) sy(k)%nvars
       allocate (b(k)%bc(sy(k)%nvars), b(k)%eo(sy(k)%nvars))
       read (u) b(k)%bc, b(k)%eo
       read (u) sy(k)%n_mass
       allocate (b(k)%mass(max(sy(k)%n_mass, 1)))
       read (u) b(k)%mass(1:sy(k)%n_mass)
       read (u) sy(k)%n_stif
       allocate (b(k)%stif(max(sy(k)%n_stif, 1)))
       read (u) b(k)%stif(1:sy(k)%n_stif)
       read (u) sy(k)%n_nl
       allocate (b(k)%nl(max(sy(k)%n_nl, 1)))
       read (u) b(k)%nl(1:sy(k)%n_nl)
       sy(k)%bc_codes = c_loc(b(k)%bc)
       sy(k)%eqn_orders = c_loc(b(k)%eo)
       sy(k)%mass = c_loc(b(k)%mass)
       sy(k)%stif = c_loc(b(k)%stif)
       sy(k)%nl = c_loc(b(k)%nl)
    end do
  end subroutine

end program step_driver
