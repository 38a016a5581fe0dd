! Fortran (iso_c_binding) interface to the C-ABI declared in coral_amd.h.
!
! The derived types mirror the C structs field by field (bind(C) gives them the
! C layout); the recipe types they carry are those of PL_equations.f90:22-79.
! Every function returns 0 on success or a negative code; coral_amd_check turns
! a failure into `error stop`, the fail-stop convention of the reference
! (PL_algebra_z_3D.f90:210-212, PL_IMEX_timestepping.f90:870-875).
module coral_amd_binding
  use, intrinsic :: iso_c_binding
  implicit none

  type, bind(C) :: coral_amd_piece            ! linear_operator_recipe_T, PL_equations.f90:59-68
     real(c_double)     :: dsca
     integer(c_int32_t) :: kx, ky, iz, mz, var, eqn
  end type
  type, bind(C) :: coral_amd_nlterm           ! atom_NL_term_T, PL_equations.f90:22-30
     real(c_double)     :: dsca
     integer(c_int32_t) :: kx, ky, iz, mz, qvar, eqn
  end type
  type, bind(C) :: coral_amd_system
     integer(c_int32_t) :: nvars
     type(c_ptr)        :: bc_codes, eqn_orders
     integer(c_int32_t) :: n_mass
     type(c_ptr)        :: mass
     integer(c_int32_t) :: n_stif
     type(c_ptr)        :: stif
     integer(c_int32_t) :: n_nl
     type(c_ptr)        :: nl
  end type
  type, bind(C) :: coral_amd_fullterm         ! a_linear_contribution_T, PL_equations.f90:71-79
     real(c_double)     :: dsca
     integer(c_int32_t) :: kx, ky, dz, mz, kind, system, position
  end type
  type, bind(C) :: coral_amd_fullvar
     integer(c_int32_t) :: nterms
     type(c_ptr)        :: terms
  end type
  type, bind(C) :: coral_amd_quadvar
     integer(c_int32_t) :: ivar1, ivar2, remove_z_integral
  end type
  type, bind(C) :: coral_amd_problem
     integer(c_int32_t) :: NX, NY, NZ
     real(c_double)     :: Lx, Ly, gap, center
     integer(c_int32_t) :: nstages
     type(c_ptr)        :: a_arr, a_hat
     integer(c_int32_t) :: n_kxky
     type(c_ptr)        :: kxky
     integer(c_int32_t) :: n_zero
     type(c_ptr)        :: zero
     integer(c_int32_t) :: n_full
     type(c_ptr)        :: full
     integer(c_int32_t) :: n_quad
     type(c_ptr)        :: quad
  end type

  interface
     integer(c_int) function coral_amd_create(prob, device, ctx) bind(C)
       import
       type(coral_amd_problem), intent(in) :: prob
       integer(c_int), value :: device
       type(c_ptr) :: ctx
     end function
     integer(c_int) function coral_amd_destroy(ctx) bind(C)
       import
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function coral_amd_set_dt(ctx, dt) bind(C)
       import
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
     end function
     integer(c_int) function coral_amd_upload_state(ctx, kind, isys, data) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: kind, isys
       type(c_ptr), value :: data
     end function
     integer(c_int) function coral_amd_download_state(ctx, kind, isys, data) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: kind, isys
       type(c_ptr), value :: data
     end function
     integer(c_int) function coral_amd_step(ctx, nsteps) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nsteps
     end function
     integer(c_int) function coral_amd_synchronize(ctx) bind(C)
       import
       type(c_ptr), value :: ctx
     end function
     ! steps replayed as HIP graphs (same results, less launch overhead)
     integer(c_int) function coral_amd_use_graphs(ctx, enable) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: enable
     end function
     ! boundary-value extraction (1 BotSurf, 2 TopSurf) and vertical
     ! decomposition (1 barotropic, 2 baroclinic) of full variable ifull (0-based)
     integer(c_int) function coral_amd_set_full_variable_filter(ctx, ifull, extract, zdecomp) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ifull, extract, zdecomp
     end function
     ! fixed physical values of full variable ifull (a penalisation mask),
     ! NXAA x NYAA x NZAA in the layout of coral_amd_get_physical; c_null_ptr
     ! returns it to its spectral definition
     integer(c_int) function coral_amd_set_fixed_physical(ctx, ifull, phys) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ifull
       type(c_ptr), value :: phys
     end function
     ! horizontal Smagorinsky viscosity in full variable ifull (the last; -1
     ! off) from the squares of full variables grads(1:ngrads)
     integer(c_int) function coral_amd_set_smagorinsky(ctx, ifull, grads, ngrads, prefactor) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ifull
       integer(c_int32_t), intent(in) :: grads(*)
       integer(c_int32_t), value :: ngrads
       real(c_double), value :: prefactor
     end function
     ! mean-system sources: a term dsca Iz^iz sourceNN of equation eqn of
     ! mean system isys (0-based); source NN from its NZAA Gauss-point values
     ! (top wall first; c_null_ptr removes it)
     integer(c_int) function coral_amd_add_source_term(ctx, isys, eqn, iz, dsca, source_index) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: isys, eqn, iz
       real(c_double), value :: dsca
       integer(c_int32_t), value :: source_index
     end function
     integer(c_int) function coral_amd_clear_source_terms(ctx) bind(C)
       import
       type(c_ptr), value :: ctx
     end function
     integer(c_int) function coral_amd_set_source(ctx, source_index, physical) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: source_index
       type(c_ptr), value :: physical
     end function
     ! precision 0: FP64; 1: FP32 transforms and products, FP64 solves
     ! (refine = 1: one FP64 residual-refinement sweep per stage)
     integer(c_int) function coral_amd_set_precision(ctx, precision, refine) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: precision, refine
     end function
     ! Hermitian projection of the kx = 0 plane after every step
     integer(c_int) function coral_amd_set_kx0_projection(ctx, enable) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: enable
     end function
     integer(c_int) function coral_amd_get_cfl_dt(ctx, dt) bind(C)
       import
       type(c_ptr), value :: ctx
       real(c_double) :: dt
     end function
     integer(c_int) function coral_amd_get_physical(ctx, ifull, out) bind(C)
       import
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ifull
       type(c_ptr), value :: out
     end function
     ! spectra of the quadratic variables of the current state, complex
     ! (NZ, NYAA, NX/2, nquad) in Fortran order
     integer(c_int) function coral_amd_products(ctx, out) bind(C)
       import
       type(c_ptr), value :: ctx, out
     end function
     ! batched zgbtrf + zgbtrs on the device (PL_algebra_z_3D.f90:189-276):
     ! ab holds nsys complex band matrices in LAPACK storage (ldab = 2kl+ku+1),
     ! x the right-hand sides, overwritten with the solutions
     integer(c_int) function coral_amd_banded_solve(device, n, kl, ku, nsys, ab, x, info) bind(C)
       import
       integer(c_int32_t), value :: device, n, kl, ku, nsys
       type(c_ptr), value :: ab, x, info
     end function
     integer(c_int) function coral_amd_last_error(buf, len) bind(C)
       import
       character(kind=c_char) :: buf(*)
       integer(c_int32_t), value :: len
     end function
  end interface

contains

  subroutine coral_amd_check(ierr)
    integer(c_int), intent(in) :: ierr
    character(kind=c_char) :: msg(512)
    character(len=512) :: text
    integer(c_int) :: i2
    integer :: n
    if (ierr == 0) return
    i2 = coral_amd_last_error(msg, 512_c_int32_t)
    text = ''
    n = 0
    do while (n < 512)
       if (msg(n + 1) == c_null_char) exit
       n = n + 1
       text(n:n) = msg(n)
    end do
    print '(a, a)', 'coral_amd: ', text(1:n)
    error stop 'coral_amd failure'
  end subroutine

end module coral_amd_binding
