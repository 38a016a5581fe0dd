/*
 * coral_amd -- MI355X-native plane-layer IMEX time-stepper, C ABI.
 *
 * The ABI replaces the seams a Coral host crosses on its hot path:
 *   - the step seam  full_problem_data_structure_T%march_forward(dt) and
 *     %factorize_operators(dt)  (src/plane_layer/PL_IMEX_timestepping.f90:191,196,
 *     called from src/plane_layer/plane_layer_main.f90:147,157);
 *   - the transform seam PS_fields_T%spec_to_phys / %phys_to_spec
 *     (src/slabs.fftw3mpi/transforms.f90:10-25, :155-301);
 *   - the algebra seam zOperator_3d_1coupled_T%factorize / %backsolve
 *     (src/plane_layer/PL_algebra_z_3D.f90:12-36, :189-276).
 * The problem is handed over as the parsed recipe (full_problem_recipe_T,
 * src/plane_layer/PL_equations.f90:59-219) plus the geometry of
 * coral.parameters.in (src/plane_layer/PL_geometry.f90:269-305); the library
 * assembles every operator itself.
 *
 * Conventions
 *   - every function returns 0 on success and a negative code on failure; the
 *     message is available from coral_amd_last_error (thread-local);
 *   - all host buffers are caller owned; device memory and the HIP stream are
 *     owned by the context; calls enqueue on the context's stream and return
 *     once host-visible results (downloads) are complete;
 *   - indices of systems, variables and equations are 0-based;
 *   - spectral state of a kxky system uses Coral's layout: complex double
 *     (re, im interleaved) array of Fortran shape (total, NYAA, NX/2), i.e. C
 *     shape [NX/2][NYAA][total]; mean (zero) systems are real double [total].
 */
#ifndef CORAL_AMD_H
#define CORAL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct coral_amd_ctx coral_amd_ctx;

/* One linear piece dsca * (dx)^kx (dy)^ky Iz^iz mz^mz var, in equation eqn
 * (linear_operator_recipe_T, PL_equations.f90:59-68). */
typedef struct {
    double dsca;
    int32_t kx, ky, iz, mz, var, eqn;
} coral_amd_piece;

/* One nonlinear term dsca * (dx)^kx (dy)^ky Iz^iz mz^mz Q[qvar] in equation eqn
 * (atom_NL_term_T, PL_equations.f90:22-30). */
typedef struct {
    double dsca;
    int32_t kx, ky, iz, mz, qvar, eqn;
} coral_amd_nlterm;

/* A coupled set of variables (kxky or mean).  Mean systems must not carry
 * horizontal derivatives. */
typedef struct {
    int32_t nvars;
    const int32_t *bc_codes;   /* [nvars], Galerkin codes of chebyshev_galerkin_2.f90 */
    const int32_t *eqn_orders; /* [nvars] */
    int32_t n_mass;
    const coral_amd_piece *mass;
    int32_t n_stif;
    const coral_amd_piece *stif;
    int32_t n_nl;
    const coral_amd_nlterm *nl;
} coral_amd_system;

/* One term of a full variable (a_linear_contribution_T, PL_equations.f90:71-79).
 * kind: 0 = kxky variable, 1 = mean variable. */
typedef struct {
    double dsca;
    int32_t kx, ky, dz, mz, kind, system, position;
} coral_amd_fullterm;

typedef struct {
    int32_t nterms;
    const coral_amd_fullterm *terms;
} coral_amd_fullvar;

/* Quadratic variable Q = F[ivar1] * F[ivar2]; -1 stands for ONE. */
typedef struct {
    int32_t ivar1, ivar2, remove_z_integral;
} coral_amd_quadvar;

typedef struct {
    int32_t NX, NY, NZ;            /* dealiased mode counts (coral.parameters.in lines 1-3) */
    double Lx, Ly, gap, center;    /* lines 5-8 */
    int32_t nstages;               /* implicit stages s of the ARS scheme */
    const double *a_arr;           /* s x s, row-major (IMEX_schemes.f90 A_arr) */
    const double *a_hat;           /* (s+1) x (s+1), row-major (A_hat) */
    int32_t n_kxky;
    const coral_amd_system *kxky;
    int32_t n_zero;
    const coral_amd_system *zero;
    int32_t n_full;
    const coral_amd_fullvar *full;
    int32_t n_quad;
    const coral_amd_quadvar *quad;
} coral_amd_problem;

/* Build the solver on HIP device `device`: assembles every operator from the
 * recipe, allocates all device memory.  Replaces the setup of
 * full_problem_data_structure_T%init (PL_IMEX_timestepping.f90:510-569). */
int coral_amd_create(const coral_amd_problem *prob, int device, coral_amd_ctx **out);
int coral_amd_destroy(coral_amd_ctx *ctx);
/* Copy the last error message of the calling thread into buf (NUL terminated). */
int coral_amd_last_error(char *buf, int32_t len);

/* (Re)factorize M - dt*a11*L for every wavenumber on the device
 * (factorize_operators, PL_IMEX_timestepping.f90:351-422). */
int coral_amd_set_dt(coral_amd_ctx *ctx, double dt);

/* State transfer in Coral's layout (see header comment).  kind 0 = kxky
 * system, 1 = mean system.  Upload zeroes dealiased modes and the kxky (0,0)
 * mode, which the reference keeps at zero (PL_IMEX_timestepping_marching.f90:32-46). */
int coral_amd_upload_state(coral_amd_ctx *ctx, int32_t kind, int32_t isys, const double *data);
int coral_amd_download_state(coral_amd_ctx *ctx, int32_t kind, int32_t isys, double *data);

/* march_forward(dt) repeated nsteps times with the factorized dt
 * (one_step_beyond, PL_IMEX_timestepping_marching.f90:3-48).  Asynchronous. */
int coral_amd_step(coral_amd_ctx *ctx, int32_t nsteps);
int coral_amd_synchronize(coral_amd_ctx *ctx);

/* Full variable `ifull` of the current field in physical space, C shape
 * [NXAA][NYAA][NZAA] (x slowest, z fastest). */
int coral_amd_get_physical(coral_amd_ctx *ctx, int32_t ifull, double *out);
/* CFL time step of the last evaluated stage (compute_cfl_based_timestep,
 * PL_IMEX_timestepping_marching.f90:200-254, cflFactor_along_z = 1). */
int coral_amd_get_cfl_dt(coral_amd_ctx *ctx, double *out);

/* Integer wavenumbers and 2/3-rule masks of Coral's slab layout
 * (PL_geometry.f90:179-222): kx_int[NX/2], ky_int[NYAA], dealias_x[NX/2],
 * dealias_y[NYAA] (1 = mode zeroed). */
int coral_amd_wavenumbers(coral_amd_ctx *ctx, int32_t *kx_int, int32_t *ky_int,
                          uint8_t *dealias_x, uint8_t *dealias_y);

/* Sizes: total unknowns of kxky system isys, its band (nl, nu) and the number
 * of kept wavenumbers (NX/2 * (NY-1)). */
int coral_amd_system_info(coral_amd_ctx *ctx, int32_t isys, int32_t *total, int32_t *nl,
                          int32_t *nu, int32_t *nkept);

/* Arithmetic of the factored evolution operator of kxky system isys: 1 when it
 * is factored in real form (each variable scaled by i^phi, each equation by
 * i^-psi, which makes every (i kx)^a (i ky)^b piece real), 0 when complex.
 * The reference factors the complex operator (zgbtrf, PL_algebra_z_3D.f90:245-276);
 * pivot rows and solutions agree, since the scaling leaves |Re|+|Im| of each
 * entry unchanged.  Setting CORAL_AMD_COMPLEX_FACTORS=1 before
 * coral_amd_create forces the complex form. */
int coral_amd_factor_kind(coral_amd_ctx *ctx, int32_t isys, int32_t *is_real);

/* Matrices factored per system (kind 0 = kxky, 1 = mean): the evolution
 * operator splits into nblk structurally independent blocks (even and odd
 * Chebyshev degrees for Coral's operators) of blk_n unknowns and band
 * (blk_nl, blk_nu), each factored and solved on its own; nblk = 1 when the
 * whole system is factored (blk_n = total).  The reference factors the whole
 * system (zgbtrf, PL_algebra_z_3D.f90:245-276): partial pivoting never selects
 * a structurally zero row, so every row receives the same eliminations in the
 * same order.  CORAL_AMD_NO_BLOCK_SPLIT=1 before coral_amd_create keeps whole
 * systems.
 * u_entries (may be NULL) receives the mean number of U entries per column the
 * solve reads after the last factorization: each column's record is cut to
 * the longest nonzero prefix it has over all wavenumbers (diagonal included). */
int coral_amd_factor_blocks(coral_amd_ctx *ctx, int32_t kind, int32_t isys, int32_t *nblk, int32_t *blk_n,
                            int32_t *blk_nl, int32_t *blk_nu, double *u_entries);

/* Physical values of coupled variable ivar of system isys in the current state:
 * the volumes Coral writes to CheckPoints/QuickSave.{kxky,zero}.sysNNN.varNNN.rollingR
 * (export_checkpoints, PL_IMEX_timestepping_output.f90:388-455).  kind 0: C
 * shape [NXAA][NYAA][NZAA] (file order: z fastest, then y, then x; the (0,0)
 * mode excluded); kind 1: the NZAA-point profile of a mean variable. */
int coral_amd_get_variable_physical(coral_amd_ctx *ctx, int32_t kind, int32_t isys, int32_t ivar, double *out);
/* Inverse: forward transform, solve of the square-truncated Galerkin stencil
 * and insertion of the variable's Galerkin coefficients into the state
 * (import_quicksave, PL_IMEX_timestepping_output.f90:309-385).  Other variables
 * keep their coefficients; the kxky (0,0) mode stays zero. */
int coral_amd_set_variable_physical(coral_amd_ctx *ctx, int32_t kind, int32_t isys, int32_t ivar,
                                    const double *in);

/* Horizontal averages of the current state at the NZAA Gauss points, the
 * quantities behind Coral's timeseries (output_global_quantities,
 * PL_IMEX_timestepping_output.f90:5-231: zSlice objects are these values,
 * volume objects their weighted sums): full_avg [nfull][NZAA], quad_avg
 * [nquad][NZAA] (either may be NULL).  The averages are the (kx, ky) = (0, 0)
 * modes, taken after the z transform (full variables) and after the x/y
 * transforms of the products (quadratic variables). */
int coral_amd_horizontal_averages(coral_amd_ctx *ctx, double *full_avg, double *quad_avg);

/* Unit seams on host buffers.
 * Backward transform of a spectral field, C shape [NX/2][NYAA][NZ] complex, to
 * physical C shape [NXAA][NYAA][NZAA] (c2r_transform, slabs.fftw3mpi/transforms.f90:155-225);
 * dealiased slots of the input are ignored.  Forward is its inverse on the kept
 * modes (r2c_transform, :227-301); dealiased output slots are set to zero. */
/* Quadratic variables of the current state, as the step forms them: full
 * variables on the 3/2 grid, their products (marching.f90:279-300), forward
 * transforms (r2c_transform, slabs.fftw3mpi/transforms.f90:227-301) and the
 * z-integral removal (PL_IMEX_timestepping.f90:223-235).  out: complex,
 * [nquad][NX/2][NYAA][NZ] (the reference's spectral layout), dealiased ky
 * slots zero. */
int coral_amd_products(coral_amd_ctx *ctx, double *out);
int coral_amd_bwd_transform(coral_amd_ctx *ctx, const double *spec, double *phys);
int coral_amd_fwd_transform(coral_amd_ctx *ctx, const double *phys, double *spec);

/* Batched complex banded LU with partial pivoting and solve, on the device, for
 * nsys matrices given in LAPACK band storage (ldab = 2kl+ku+1, column j of
 * system s at ab[(s*n + j)*ldab], complex interleaved): the zgbtrf/zgbtrs
 * calls of PL_algebra_z_3D.f90:189-276.  x[s*n + i] is overwritten with the
 * solution; info[s] receives the LAPACK info of each factorization.  Bands
 * with kl <= 16 and kl + ku <= 47 (Coral's operators: 16 and 43) go through
 * the factor records, U-record trimming and 16-lane solve kernel of the time
 * step; other bands through a generic one-wavefront-per-system solve. */
int coral_amd_banded_solve(int32_t device, int32_t n, int32_t kl, int32_t ku, int32_t nsys,
                           const double *ab, double *x, int32_t *info);

/* Allocation-free variant of coral_amd_banded_solve.  It has no counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  ab, x and info are DEVICE pointers on `device`: ab (LAPACK band
 * storage, as for coral_amd_banded_solve) is overwritten with its LU factors,
 * as zgbtrf overwrites AB; x holds the right-hand sides on entry and the
 * solutions on return; info[s] receives the LAPACK info of system s.  The work
 * space (multipliers, U records and pivots; coral_amd_banded_solve_noalloc_bytes
 * gives its size) lives in one fixed device array of
 * CORAL_AMD_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; a
 * batch that needs more is rejected with an error (split it).  The work is
 * enqueued on `stream` (a hipStream_t, NULL for the null stream) and the call
 * returns without waiting for it.  The first call on a device looks up the
 * array's device address (make that call outside graph capture); no call
 * allocates.  Every call on one GPU uses the same array, so calls whose work
 * overlaps in time on one GPU race.  The solve is the one-wavefront-per-system
 * kernel, so results equal coral_amd_banded_solve's to rounding. */
#define CORAL_AMD_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int coral_amd_banded_solve_noalloc(int32_t device, int32_t n, int32_t kl, int32_t ku, int32_t nsys, double *ab,
                                   double *x, int32_t *info, void *stream);
/* Bytes of work space coral_amd_banded_solve_noalloc needs for a batch (0 for
 * invalid dimensions). */
size_t coral_amd_banded_solve_noalloc_bytes(int32_t n, int32_t kl, int32_t ku, int32_t nsys);

/* Solve (M - dt*a11*L) x = rhs for every kept wavenumber of kxky system isys
 * with the factors of the last coral_amd_set_dt, through the factor layout and
 * solve kernel the time step uses: the backsolve of
 * zOperator_3d_1coupled_T%backsolve (PL_algebra_z_3D.f90:189-243), called as
 * backsolve_to_aux at PL_IMEX_timestepping_marching.f90:57-58.  rhs and x are
 * complex in the state layout (Fortran (total, NYAA, NX/2)); dealiased modes
 * of rhs are ignored and set to zero in x; the kxky (0,0) system is the
 * identity (factorize_operators, PL_IMEX_timestepping.f90:387-406), so x
 * equals rhs there.  The state is left unchanged (the stage buffers are used
 * as work space). */
int coral_amd_system_solve(coral_amd_ctx *ctx, int32_t isys, const double *rhs, double *x);

/* Assembly only, no device touched: the dense operator of kxky system isys at
 * integer wavenumbers (kx_int, ky_int), i.e. the per-wavenumber CSR matrix the
 * reference builds in PL_IMEX_timestepping.f90:1500-1574 after the change of
 * basis (PL_algebra_z_3D.f90:40-53).  which = 0 mass M, 1 stiffness L.  out
 * (complex, row-major total x total) may be NULL to query total/nl/nu only. */
int coral_amd_assembled_operator(const coral_amd_problem *prob, int32_t isys, int32_t which, int32_t kx_int,
                                 int32_t ky_int, double *out, int32_t *total, int32_t *nl, int32_t *nu);

/* Precision of the time step (no counterpart in the reference, which steps in
 * FP64 throughout).  mode 0: FP64 everywhere (default).  mode 1: the
 * transforms and the quadratic products run in FP32: extraction rounds the
 * Chebyshev coefficients of the full variables to FP32, the z transforms are
 * FP32 MFMA products with compensated (TwoSum) accumulation, the y and x FFTs
 * and the products are FP32; the NL sums into K_hat, the right-hand sides, the
 * banded factors and solves and the implicit update stay FP64.  refine: FP64
 * residual-refinement sweeps after the banded solves (e = A x - rhs, A f = e,
 * x -= f with the same factors): 0 none, 1 after the last stage of each step,
 * 2 after every stage.  Mode 1 needs wavefront FFT plans for NXAA and NYAA
 * (96, 144, 192, 384, 768) and an FP32 MFMA plan for NZ, and allocates FP32
 * work arrays (about half the size of the FP64 ones) on first use.  The
 * coral_amd_products seam follows the mode; the other seams stay FP64. */
int coral_amd_set_precision(coral_amd_ctx *ctx, int32_t mode, int32_t refine);

/* enable=1: after each step, replace the kx = 0 plane of every kxky state by
 * its Hermitian part, c(0, ky) = (c(0, ky) + conj c(0, -ky)) / 2.  No
 * counterpart in the reference, whose steps (and the default here, enable=0)
 * carry the anti-Hermitian part along.  The physical fields never see that
 * part: the x c2r drops the imaginary part of kx = 0 after the y transform
 * (c2r_transform, slabs.fftw3mpi/transforms.f90:204-210), so the Hermitian NL
 * terms never act on it and it evolves under the linear operator alone.  Where
 * that operator has unstable kx = 0 modes (supercritical convection) it grows
 * from rounding level until rounding leaks it back into the fields, which
 * ends the run (the 96x96x48 rotating case of
 * etc/benchmarks/02_RRBC_julien_JFM96 after about 15,000 steps).  The
 * projection changes no physical field. */
int coral_amd_set_kx0_projection(coral_amd_ctx *ctx, int32_t enable);

/* enable=1: coral_amd_step replays one step as a HIP graph (captured at the
 * next step, dropped by coral_amd_set_dt) instead of launching its kernels one
 * by one; same results bit for bit, less launch overhead on small grids.  No
 * counterpart in the reference (its time loop calls march_forward,
 * plane_layer_main.f90:157).  While timers are enabled steps run eagerly. */
int coral_amd_use_graphs(coral_amd_ctx *ctx, int32_t enable);

/* Learned subgrid closure of a coarse run (no counterpart in the reference).
 * The network of coral_amd/closure.py (per horizontal plane: a lift from the
 * nfull full variables on each point's 3 x 3 x 3 neighbourhood (periodic in x
 * and y, clamped at the walls) and the height to `width` channels, `layers` residual
 * layers of a GELU MLP of `hidden` channels plus a spectral convolution over
 * the plane's Fourier modes |kx|, |ky| <= modes, a projection to the nquad
 * quadratic variables) maps the full variables of each stage to corrections
 * of the quadratic variables, whose spectra are added to theirs before the NL
 * terms are assembled (every stage, and in coral_amd_products).  weights: the
 * flat FP32 parameters in that layout (nweights of them); in_mean / in_std
 * (cin = 27 nfull + 1 rounded up to a multiple of 8: the neighbourhood's
 * values, the height, zero columns) normalise the inputs, out_mean / out_std
 * (nquad) scale the outputs back.  width 32 or 64, hidden 128 or 256.  GEMMs
 * run in BF16 on MFMA with FP32 accumulation.  NULL removes the closure.
 * FP64 transform chain only (precision mode 0). */
typedef struct {
    int32_t width, hidden, layers, modes;
    const float *weights;
    int64_t nweights;
    const float *in_mean, *in_std;
    const float *out_mean, *out_std;
} coral_amd_closure;
int coral_amd_set_closure(coral_amd_ctx *ctx, const coral_amd_closure *closure);

/* Boundary-value extraction and vertical decomposition of full variable
 * ifull (the recipe's extract_value_at and baroclinic_or_barotropic,
 * fullVars_recipe_T, PL_equations.f90:114-118; applied where
 * compute_full_variables_in_physical_space applies them,
 * PL_IMEX_timestepping_marching.f90:544-577): extract 0 none, 1 BotSurf (the
 * z-constant field of the variable's value at the bottom wall), 2 TopSurf;
 * zdecomp 0 none, 1 barotropic (only its vertical mean kept), 2 baroclinic
 * (its vertical mean removed).  Set after coral_amd_create. */
int coral_amd_set_full_variable_filter(coral_amd_ctx *ctx, int32_t ifull, int32_t extract, int32_t zdecomp);

/* Fixed physical values of full variable ifull (a penalisation variable:
 * init_penalisation, PL_IMEX_timestepping_marching.f90:778-822, and the
 * penalisation guard of compute_full_variables_in_physical_space, :515):
 * NXAA x NYAA x NZAA values in the layout of coral_amd_get_physical (z
 * fastest, then y, then x), or NULL to return the variable to its spectral
 * definition.  Quadratic variables with such a factor are formed in
 * physical space, as the reference forms all of them (:279-300).  Not for the
 * first three full variables (the CFL bound's velocities).  FP64 transform
 * chain only. */
int coral_amd_set_fixed_physical(coral_amd_ctx *ctx, int32_t ifull, const double *phys);

/* Horizontal Smagorinsky viscosity (the --withHorizontalSmagorinsky run;
 * PL_IMEX_timestepping_marching.f90:514-541, 241-253): full variable ifull
 * (the last one, nuSmago; -1 switches it off) takes at every stage the
 * z-profile prefactor * sqrt(previous profile + the horizontal mean of the
 * summed squares of full variables grads[0..ngrads-1], ngrads <= 4: the
 * reference's dxUfull, dyUfull, dxVfull, dyVfull); the profile starts at
 * zero.  coral_amd_get_cfl_dt then also bounds dt by
 * 0.95 / ((2 pi)^2 max nuSmago ((NXAA/Lx)^2 + (NYAA/Ly)^2)).  FP64 transform
 * chain only. */
int coral_amd_set_smagorinsky(coral_amd_ctx *ctx, int32_t ifull, const int32_t *grads, int32_t ngrads,
                              double prefactor);

/* Sources of the mean systems.  A term (an add_rhs_source atom, built by
 * build_zero_sources_in_full_recipe, PL_equations.f90:1156-1230) adds
 * dsca Iz^iz sourceNN to equation eqn of mean system isys (0-based), with
 * source_index = NN in 1..99; coral_amd_clear_source_terms removes all terms.
 * coral_amd_set_source defines source NN from its NZAA values on the
 * Gauss-Chebyshev grid z_j = center + gap/2 cos(pi (2j+1) / (2 NZAA)),
 * j = 0..NZAA-1 (top wall first; the sourceDef array of
 * PL_user_sources_definitions.f90 and add_source_array,
 * PL_IMEX_timestepping_marching.f90:754-766), or removes it (NULL).  Every
 * stage adds trunc(I^iz)[order:, :] (dsca c) to the equation's rows of K_hat,
 * c the first NZ Chebyshev coefficients of the source (REDFT10 / NZAA, the
 * first halved; PL_IMEX_timestepping_marching.f90:399-416, 845-866).
 * Stepping fails while a term names a source with no definition. */
int coral_amd_add_source_term(coral_amd_ctx *ctx, int32_t isys, int32_t eqn, int32_t iz, double dsca,
                              int32_t source_index);
int coral_amd_clear_source_terms(coral_amd_ctx *ctx);
int coral_amd_set_source(coral_amd_ctx *ctx, int32_t source_index, const double *physical);

/* Launch shapes of a run (no counterpart in the reference): xpass_wpb is the
 * largest number of wavefronts per workgroup of the x pass (1..4; 0 = the most
 * that fit its LDS; 768-point lines take two wavefronts each, so they run on
 * two at least), z_plain = 1 runs the z transforms as vector GEMMs
 * instead of MFMA products (results equal to rounding).  The ensemble tuner
 * sets them per run. */
int coral_amd_set_launch_shape(coral_amd_ctx *ctx, int32_t xpass_wpb, int32_t z_plain);
int coral_amd_get_launch_shape(coral_amd_ctx *ctx, int32_t *xpass_wpb, int32_t *z_plain);

/* Ensembles of independent runs (no counterpart in the reference, which runs
 * one problem per executable).  Each run is a context of its own (its recipe,
 * grid, operators, dt and state; contexts stay owned by the caller, must
 * outlive the ensemble and keep working on their own).  The runs are spread
 * over nbranch branches; each branch's step (its runs one after the other) is
 * one captured HIP graph replayed on a stream of its own, so one
 * coral_amd_ensemble_step advances every run by one time step with nbranch
 * graph launches, and the branches overlap on the device.  Each run's results
 * are bit for bit those of its own coral_amd_step with the same launch
 * shapes.  The graphs are recaptured when a run's dt, precision, projection or
 * launch shapes change or when a run was stepped on its own.  All runs must be
 * on one device and have timers off. */
typedef struct coral_amd_ensemble coral_amd_ensemble;
int coral_amd_ensemble_create(coral_amd_ctx *const *runs, int32_t nrun, coral_amd_ensemble **out);
int coral_amd_ensemble_destroy(coral_amd_ensemble *ens);
/* nbranch parallel branches; assign[r] the branch of run r (NULL: longest
 * step first onto the least loaded branch, by the tuner's step times).
 * nbranch <= 0 selects the lockstep schedule: one graph per ensemble step in
 * which, stage by stage, every run's NL terms and right-hand sides (spread
 * over max(1, -nbranch) parallel branches) are followed by the stage's kxky
 * solves of all runs as one batched launch per kernel variant (same results
 * bit for bit); it needs the same IMEX scheme on every run and no residual
 * refinement. */
int coral_amd_ensemble_set_schedule(coral_amd_ensemble *ens, int32_t nbranch, const int32_t *assign);
/* The tuner: for every run, each launch shape (x pass 4, 2 or 1 wavefronts per
 * workgroup; z transforms on MFMA or as vector GEMMs) timed over trial_steps
 * graph replays, the fastest kept; then 1, 2, 4, ... 16 branches with
 * longest-first schedules, and the lockstep schedule with its stage fronts
 * over as many branches, timed as ensemble steps, the fastest kept.  Every
 * run's state is restored afterwards.  Returns the branch count (negative:
 * lockstep over that many branches) and the ensemble's milliseconds per step. */
int coral_amd_ensemble_tune(coral_amd_ensemble *ens, int32_t trial_steps, int32_t *nbranch, double *ms_per_step);
int coral_amd_ensemble_step(coral_amd_ensemble *ens, int32_t nsteps);
int coral_amd_ensemble_synchronize(coral_amd_ensemble *ens);

/* Kernels of the learned subgrid closure (no counterpart in the reference,
 * which has no closure), on DEVICE pointers, enqueued on `stream` (a
 * hipStream_t, NULL for the null stream).  BF16 arrays hold 16-bit BF16
 * patterns; conversions round to nearest even.
 * coral_amd_nn_gemm: C[b] = alpha op(A[b]) op(B[b]) + beta C[b] for b < batch
 * on MFMA (BF16 operands, FP32 accumulation and C): op(A) is M x K (A stored M
 * x K row-major with leading dimension lda, or K x M when transA), op(B) K x N
 * (B stored K x N, or N x K when transB); sA, sB, sC are the batch strides in
 * elements (0 shares an operand).  Products with few output tiles and a long
 * K are split along K: work then holds coral_amd_nn_gemm_work(M, N, K, batch)
 * floats (0: no split, work may be NULL); the partials are added in a fixed
 * order, so results do not depend on timing. */
int coral_amd_nn_gemm(int32_t transA, int32_t transB, int32_t M, int32_t N, int32_t K, float alpha, const void *A,
                      int32_t lda, int64_t sA, const void *B, int32_t ldb, int64_t sB, float beta, float *C,
                      int32_t ldc, int64_t sC, int32_t batch, float *work, int64_t nwork, void *stream);
int64_t coral_amd_nn_gemm_work(int32_t M, int32_t N, int32_t K, int32_t batch);
/* dst (cols x rows) = src (rows x cols) transposed, BF16 */
int coral_amd_nn_transpose_bf16(const void *src, int32_t rows, int32_t cols, void *dst, void *stream);
/* One MLP branch of a layer, fused, for R rows of `width` channels:
 * H += bf16(gelu(Hin16 W1 + b1)) W2 + b2 and Hout16 = bf16(H) (Hout16 may be
 * Hin16), the hidden activations kept on chip; W1t = W1^T (hidden x width)
 * and W2t = W2^T (width x hidden) in BF16.  Zc (R x hidden, Hin16 W1 + b1)
 * and A16c (R x hidden, the BF16 activations) are written for a backward pass
 * when not NULL.  width 32 or 64, hidden 128 or 256.  Same results as
 * coral_amd_nn_gemm, _bias_act (GELU), _gemm, _bias_act, _axpy. */
int coral_amd_nn_mlp(int64_t R, int32_t width, int32_t hidden, const void *Hin16, const void *W1t, const float *b1,
                     const void *W2t, const float *b2, float *H, void *Hout16, float *Zc, void *A16c,
                     const void *basis16, const void *CsT16, int32_t N2, int32_t nbp, void *H16T, void *stream);
/* With basis16 (the Fourier basis, N2 x nbp BF16, modes contiguous, zero past
 * nb) and CsT16 (per plane the scaled modes transposed, width x nbp, from
 * coral_amd_nn_mode_scale_t), coral_amd_nn_mlp also adds the spectral branch's
 * synthesis: H <- (B Cs + H) + (U + b2), the bits of the unfused GEMM with
 * beta = 1 followed by the MLP; needs N2 % 32 == 0 and nbp % 32 == 0 (NULL
 * basis16: no spectral term).  H16T, when not NULL, receives Hout16 in the
 * tiled transposed layout (R % 32 == 0): row r, column c at ((r / 32) width +
 * c) 32 + r % 32, which coral_amd_nn_tile_transpose also writes.
 * coral_amd_nn_mode_scale_t: out16T[p][c][k] =
 * bf16(c[p][k][c] d[k][c]) for k < nb, zero up to nbp.
 * coral_amd_nn_modes: the mode projections Cm[p] (nb x width) = B^T H16[p] of
 * P planes of N2 points from BT (nbp x N2, the basis transposed, zero rows
 * past nb) and HT (H16 in the tiled transposed layout), on MFMA, chunk partials in
 * work (coral_amd_nn_modes_work floats) added in a fixed order; for the
 * shapes coral_amd_nn_modes_supported accepts. */
int coral_amd_nn_mode_scale_t(const float *c, const float *d, void *out16T, int32_t P, int32_t nb, int32_t width,
                              int32_t nbp, void *stream);
int coral_amd_nn_tile_transpose(const void *src, int64_t R, int32_t width, void *dst, void *stream);
/* C (M x N, row-major) = A (M x K, lda) B (K x N, ldb) + bias (per column),
 * C16 = bf16(C), C16T = bf16(C) in the tiled transposed layout (M % 32 == 0);
 * C16 and C16T optional: the bits of coral_amd_nn_gemm, _bias_act and
 * _tile_transpose in one kernel. */
int coral_amd_nn_gemm_bias(int32_t M, int32_t N, int32_t K, const void *A, int32_t lda, const void *B, int32_t ldb,
                           const float *bias, float *C, void *C16, void *C16T, void *stream);
int32_t coral_amd_nn_modes_supported(int32_t nbp, int32_t N2, int32_t width);
int64_t coral_amd_nn_modes_work(int32_t nbp, int32_t N2, int32_t width, int32_t P);
int coral_amd_nn_modes(const void *BT, int32_t nbp, int32_t nb, int32_t N2, const void *HT, int32_t width, int32_t P,
                       float *Cm, float *work, int64_t nwork, void *stream);
int coral_amd_nn_cast_bf16(const float *x, void *y, int64_t n, void *stream);
/* z += bias (per column of ncol); out = act(z) (act 0: identity, 1: GELU, tanh
 * form), out16 its BF16 copy (either may be NULL) */
int coral_amd_nn_bias_act(float *z, const float *bias, int32_t ncol, int64_t n, int32_t act, float *out, void *out16,
                          void *stream);
int coral_amd_nn_gelu_bwd(const float *z, const float *da, float *dz, void *dz16, int64_t n, void *stream);
/* column sums of x [nrow][ncol] into out (out = sums + beta out), deterministic;
 * work holds coral_amd_nn_colsum_work(nrow, ncol) floats */
int64_t coral_amd_nn_colsum_work(int64_t nrow, int32_t ncol);
int coral_amd_nn_colsum(const float *x, int64_t nrow, int32_t ncol, float *out, float *work, float beta, void *stream);
/* out[p][e] = c[p][e] d[e] for p < P, e < ne; its gradient dd[e] = sum_p c[p][e] g[p][e] (+ beta dd[e]) */
int coral_amd_nn_mode_scale(const float *c, const float *d, float *out, void *out16, int32_t P, int32_t ne,
                            void *stream);
int coral_amd_nn_mode_scale_grad(const float *c, const float *g, float *dd, int32_t P, int32_t ne, float beta,
                                 void *stream);
int coral_amd_nn_axpy(float *y, const float *x, float alpha, int64_t n, void *y16, void *stream);
/* grad = scale (y - t) (and its BF16 copy); partial[b] = sum of (y - t)^2 over
 * block b of 256 elements (coral_amd_nn_mse_partials(n) blocks) */
int64_t coral_amd_nn_mse_partials(int64_t n);
int coral_amd_nn_mse(const float *y, const float *t, float *grad, void *grad16, float scale, int64_t n, float *partial,
                     void *stream);
/* Adam with bias-corrected moments at step t (1-based) on FP32 weights; w16 their BF16 copy */
int coral_amd_nn_adam(float *w, const float *g, float *m, float *v, void *w16, int64_t n, float lr, float b1, float b2,
                      float eps, int32_t t, void *stream);

/* Per-kernel timing with HIP events on the context's stream.  enable=1 resets
 * and starts accumulating; names are filled by coral_amd_timer_read. */
int coral_amd_timers_enable(coral_amd_ctx *ctx, int32_t enable);
/* Returns the number of kernels timed; for i < n: total ms and launch count. */
int coral_amd_timer_read(coral_amd_ctx *ctx, int32_t i, char *name, int32_t name_len,
                         double *total_ms, int32_t *count);

#ifdef __cplusplus
}
#endif
#endif
