/*
 * lynx_chefsi.h -- C ABI of the MI355X-native Chebyshev-filtered subspace
 * iteration (CheFSI) library, liblynx_chefsi.so.
 *
 * The library replaces the hot path of LYNX's real-space DFT eigensolver:
 * the Hamiltonian apply (FD stencil + local potential + nonlocal projectors)
 * and the CheFSI solve built on it.  Every entry point takes plain pointers
 * and sizes.  Arrays are column-major: a block of wavefunctions is (Nd x ncol)
 * with leading dimension ld >= Nd, and the grid index is x fastest
 * (loc = i + j*nx + k*nx*ny).  Complex data is interleaved (re, im) doubles
 * or floats.
 *
 * Errors: every call returns a lynx_status; the message of the last failure
 * on a context is available from lynx_last_error().  No call exits the
 * process.  Calls on one context are serialised on its HIP stream; contexts
 * are independent.
 */
#ifndef LYNX_CHEFSI_H
#define LYNX_CHEFSI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lynx_ctx lynx_ctx;
typedef int lynx_status;

enum lynx_status_code {
    LYNX_OK = 0,
    LYNX_E_ARG = 1,          /* invalid argument or size                     */
    LYNX_E_HIP = 2,          /* HIP runtime / rocBLAS / rocSOLVER failure    */
    LYNX_E_CHOLESKY = 3,     /* overlap matrix not positive definite         */
    LYNX_E_EIG = 4,          /* dense eigensolver did not converge           */
    LYNX_E_NONFINITE = 5,    /* NaN or Inf produced                          */
    LYNX_E_IO = 6,           /* pseudopotential file unreadable or malformed */
    LYNX_E_STATE = 7,        /* call made before the required set-up         */
    LYNX_E_UNSUPPORTED = 8   /* valid input outside what this build handles */
};

/* Storage / arithmetic type of a wavefunction block. */
enum lynx_dtype {
    LYNX_F64 = 0,            /* real double (Gamma point)                    */
    LYNX_C128 = 1,           /* complex double (k-point, Bloch)              */
    LYNX_F32 = 2,            /* real float                                   */
    LYNX_C64 = 3,            /* complex float                                */
    LYNX_BF16 = 4            /* real bfloat16 storage, float arithmetic:
                                lynx_h_apply and lynx_cheb_filter only, on
                                orthogonal cells with an even row length    */
};

enum lynx_mem { LYNX_MEM_HOST = 0, LYNX_MEM_DEVICE = 1 };
enum lynx_cell { LYNX_CELL_ORTHOGONAL = 0, LYNX_CELL_NONORTHOGONAL = 1 };

/* ---------------------------------------------------------------- context */

/* Replaces the process-wide GPUContext singleton (GPUContext.cuh:699-709):
 * one context per caller, owning a HIP stream, device workspaces and the
 * operator data.  device_id = -1 creates a host-only context: grid, stencil
 * and projector set-up and their getters work, GPU work returns
 * LYNX_E_STATE. */
lynx_status lynx_ctx_create(lynx_ctx **out, int device_id);
void lynx_ctx_destroy(lynx_ctx *ctx);
const char *lynx_last_error(const lynx_ctx *ctx);
const char *lynx_status_string(lynx_status s);
/* The hipStream_t on which the context enqueues all work. */
lynx_status lynx_ctx_stream(const lynx_ctx *ctx, void **stream);
/* Blocks until all work queued on the context's stream has finished. */
lynx_status lynx_ctx_synchronize(lynx_ctx *ctx);

/* ------------------------------------------------------------ grid (a1) */

/* Grid, lattice and FD stencil.  Replaces Lattice(latvec, cell_type)
 * (Lattice.cpp:8-54), FDGrid (FDGrid.cpp:6-35) and FDStencil(order, grid,
 * lattice) (FDStencil.cpp:16-130).  latvec holds the three lattice vectors as
 * rows (Bohr); boundaries are periodic in all directions. */
lynx_status lynx_set_grid(lynx_ctx *ctx, int nx, int ny, int nz, const double latvec[9],
                          int cell_type, int fd_order);
/* Mesh spacing, volume element and FDStencil::max_eigval_half_lap(). */
lynx_status lynx_get_grid_info(const lynx_ctx *ctx, double h[3], double *dV,
                               double *max_eig_half_lap);
/* Scaled coefficients, each 3 x (fd_order/2 + 1), row-major by axis:
 * D1 = D1_coeff_{x,y,z}, D2 = D2_coeff_{x,y,z}, D2m = D2_coeff_{xy,xz,yz}. */
lynx_status lynx_get_stencil(const lynx_ctx *ctx, double *D1, double *D2, double *D2m);

/* ------------------------------------------------------ nonlocal (a16) */

/* Reads psp8 files and builds the Kleinman-Bylander projectors of all atom
 * images whose rc sphere touches the grid.  Replaces
 * Pseudopotential::load_psp8 (Pseudopotential.cpp:28-333),
 * Crystal::compute_nloc_influence (Crystal.cpp:136-255) and
 * NonlocalProjector::setup (NonlocalProjector.cpp:13-119).
 * Atoms must be listed type-major (all atoms of type 0 first, ...).
 * coords: 3 per atom; fractional if coords_fractional, else Cartesian. */
lynx_status lynx_set_atoms(lynx_ctx *ctx, int ntypes, const char *const *psp8_paths,
                           int natom, const int *atom_type, const double *coords,
                           int coords_fractional);

/* Direct form: projectors already built (e.g. by the caller's own
 * NonlocalProjector).  One entry per influence image, in accumulation order:
 *   img_atom[i]   physical atom the image belongs to (0..n_atom-1)
 *   img_ndc[i]    grid points in the image's rc sphere
 *   img_nproj[i]  projectors of that atom
 *   gpos          concatenated flat grid indices, sum(img_ndc) entries
 *   chi           concatenated (ndc x nproj) column-major blocks
 *   img_shift     3 per image, periodic shift in lattice-aligned coordinates
 *   ip_displ      n_atom+1 projector offsets of the physical atoms
 *   gamma         ip_displ[n_atom] KB energies, one per projector
 * Mirrors the GPU flattening of NonlocalProjector (Hamiltonian.cu:342-495). */
lynx_status lynx_set_nonlocal(lynx_ctx *ctx, int n_img, const int *img_atom,
                              const int *img_ndc, const int *img_nproj, const int *gpos,
                              const double *chi, const double *img_shift, int n_atom,
                              const int *ip_displ, const double *gamma);
/* Local-pseudopotential set-up of the atoms of lynx_set_atoms, built on the
 * host on first use: pseudocharge b = -1/(4 pi) lap V_J summed over atom
 * images, reference charge b_ref of the smooth -Z/r potential, correction
 * potential Vc = V_ref - V_J, the isolated-atom valence densities scaled to
 * Nelectron, and the model core density (zero without NLCC), each Nd
 * doubles (any may be NULL); scalars[6] = Eself, Ec, integral of b,
 * Nelectron, has-NLCC flag, number of atom boxes.  Replaces
 * Electrostatics::compute_pseudocharge / compute_Vloc / compute_Ec /
 * compute_atomic_density / compute_core_density (Electrostatics.cpp:139-520)
 * with the boxes of Crystal::compute_atom_influence (Crystal.cpp:45-134).
 * Works on host-only contexts. */
lynx_status lynx_get_electrostatics(lynx_ctx *ctx, double *b, double *b_ref, double *vc, double *rho_atom,
                                    double *rho_core, double *scalars);
lynx_status lynx_get_nonlocal_sizes(const lynx_ctx *ctx, int *n_img, long *n_gpos,
                                    long *n_chi, int *n_atom, int *n_proj);
lynx_status lynx_get_nonlocal(const lynx_ctx *ctx, int *img_atom, int *img_ndc,
                              int *img_nproj, int *gpos, double *chi, double *img_shift,
                              int *ip_displ, double *gamma);

/* ------------------------------------------------- potential, k-point */

/* Effective potential, Nd doubles (host or device pointer); the context
 * keeps its own device copy (FP64, and an FP32 copy rounded to nearest for
 * the single-precision paths).  A NaN or infinite entry returns
 * LYNX_E_NONFINITE and leaves the context without a potential: calls that
 * need one return LYNX_E_STATE until the next successful call.  Replaces
 * upload_Veff_sync (EigenSolver.cu:1028-1033). */
lynx_status lynx_set_veff(lynx_ctx *ctx, const double *veff, int mem);
/* k-point in the lattice-aligned reciprocal basis (k_i = 2 pi k_red_i / L_i),
 * as passed to Hamiltonian::apply_kpt (Hamiltonian.hpp:78-80) and
 * NonlocalProjector::set_kpoint.  Used by the complex dtypes. */
lynx_status lynx_set_kpoint(lynx_ctx *ctx, const double kpt[3]);

/* ----------------------------------------------------- operator (a2-a7) */

/* y = -1/2 lap psi + (Veff + c) psi + V_nl psi for ncol columns.
 * Replaces Hamiltonian::apply (Hamiltonian.hpp:64-65) for LYNX_F64/F32 and
 * Hamiltonian::apply_kpt (Hamiltonian.hpp:78-80) for LYNX_C128/C64.
 * ld_psi, ld_y >= Nd (0 means Nd).  y is overwritten. */
lynx_status lynx_h_apply(lynx_ctx *ctx, int dtype, const void *psi, int ld_psi, void *y,
                         int ld_y, int ncol, double c, int mem);

/* lynx_h_apply on device buffers without allocating: no hipMalloc, hipFree
 * or allocating library call, for callers that cannot allocate (code under
 * HIP graph capture, serving loops, harnesses that forbid it).  No
 * counterpart in the reference.  psi and y are device pointers; the work is
 * enqueued on the context's stream and the call returns without waiting
 * (lynx_ctx_synchronize waits).  The nonlocal projector coefficients go to a
 * 32 MiB scratch array that each GPU holds once for the process: ncol times
 * atoms times padded projectors (8, 16 or 32) times 8 bytes (16 complex; 4
 * and 8 in single precision) must fit, else LYNX_E_ARG before any launch.
 * The array is shared by every call on one GPU, from every context:
 * overlapping calls on one GPU (two streams, two host threads) race.  The
 * launch shape is the forced one, or the one lynx_h_apply or the filter
 * tuned for this grid, dtype and column count, else a default shape. */
lynx_status lynx_h_apply_noalloc(lynx_ctx *ctx, int dtype, const void *psi, int ld_psi, void *y, int ld_y,
                                 int ncol, double c);

/* ------------------------------------------------------ filter (a8) */

/* Degree-m scaled Chebyshev filter Y = p_m(H) X on [lambda_c, eig_max],
 * normalised at eig_min.  Replaces EigenSolver::chebyshev_filter
 * (EigenSolver.cpp:558-619, GPU chebyshev_filter_gpu EigenSolver.cu:755-783).
 * x is read-only; y receives the result. */
lynx_status lynx_cheb_filter(lynx_ctx *ctx, int dtype, const void *x, int ld_x, void *y,
                             int ld_y, int ncol, int degree, double lambda_c,
                             double eig_min, double eig_max, int mem);

/* -------------------------------------------------- subspace (a9-a12) */

/* S = dV X^H X, Cholesky S = R^H R, X <- X R^-1.  EigenSolver.cpp:621-641.
 * Returns LYNX_E_CHOLESKY when S is not positive definite. */
lynx_status lynx_orthogonalize(lynx_ctx *ctx, int dtype, void *x, int ld, int ncol, int mem);
/* Hs = dV X^H H X, Hermitian part.  EigenSolver.cpp:643-660.  hs: ncol x ncol
 * column-major host array (complex interleaved for C128). */
lynx_status lynx_project_hamiltonian(lynx_ctx *ctx, int dtype, const void *x, int ld,
                                     int ncol, void *hs, int mem);

/* -------------------------------------------------------- solve (a13) */

/* One CheFSI pass: filter, Cholesky-QR orthonormalisation, Rayleigh-Ritz,
 * rotation.  psi (in/out) holds nband columns with leading dimension ld;
 * eigvals (host, nband doubles) receives the Ritz values in ascending order;
 * lambda_c_out (host, may be NULL) receives eigvals[nband-1] + 0.1.
 * Replaces EigenSolver::solve (EigenSolver.hpp:38-41, EigenSolver.cpp:738-832)
 * for LYNX_F64 and EigenSolver::solve_kpt (EigenSolver.hpp:49-53) for
 * LYNX_C128. */
lynx_status lynx_chefsi_solve(lynx_ctx *ctx, int dtype, void *psi, int ld, int nband,
                              double *eigvals, double lambda_c, double eig_min,
                              double eig_max, int degree, double *lambda_c_out, int mem);

/* ------------------------------------------------------ Lanczos (a14) */

/* Spectral bounds of H by Lanczos from a glibc srand(seed)/rand() start
 * vector: returns (lambda_min - 0.1, lambda_max * 1.01).  Replaces
 * EigenSolver::lanczos_bounds (EigenSolver.cpp:724-732, Lanczos.hpp:104-198);
 * the reference seeds with 1.  iters (may be NULL) gets the step count. */
lynx_status lynx_lanczos_bounds(lynx_ctx *ctx, int dtype, double *eig_min, double *eig_max,
                                double tol, int max_iter, unsigned seed, int *iters);

/* ------------------------------------------------- initial psi (a17) */

/* Initial wavefunctions: nband columns of nd values (leading dimension ld,
 * 0 means nd), uniform in [-0.5, 0.5] from glibc's srand(seed)/rand()
 * stream, column by column; complex entries draw the real part, then the
 * imaginary part.  LYNX_F32 / LYNX_C64 blocks receive the same values
 * rounded.  Replaces Wavefunction::randomize / randomize_kpt
 * (Wavefunction.cpp:53-76); the reference's SCF seeds spin s with
 * s*1000 + 1 (SCF.cpp:863-881).  The stream is drawn from a private state
 * (the caller's rand() sequence is not advanced).  ctx may be NULL for host
 * blocks; device blocks are uploaded on the context's stream. */
lynx_status lynx_randomize(lynx_ctx *ctx, int dtype, void *psi, int nd, int ld, int nband, unsigned seed,
                           int mem);

/* ------------------------------------------------------ diagnostics */

/* Host-only helpers, usable without a GPU (no context needed). */
/* FDStencil::compute_weights (FDStencil.cpp:23-44): w1, w2 of fd_order/2+1. */
lynx_status lynx_fd_weights(int fd_order, double *w1, double *w2);
/* psp8 header summary: zion, lmax, lloc, nproj_per_atom, rc_max. */
/* Monkhorst-Pack k-point grid k1 x k2 x k3 with shift, reduced by time
 * reversal (replaces KPoints::generate, src/core/KPoints.cpp:26-90).
 * kpt_cart: 3 per point, 2 pi k_red / L per lattice direction (the basis
 * lynx_set_kpoint takes); kpt_red: reduced coordinates; weights sum to 1.
 * *nkpt receives the count; pass capacity 0 and NULL arrays to query it. */
lynx_status lynx_kpoints(int k1, int k2, int k3, const double shift[3], const double latvec[9], int cell,
                         double *kpt_cart, double *kpt_red, double *weights, int capacity, int *nkpt);
lynx_status lynx_psp8_info(const char *path, double *zion, int *lmax, int *lloc,
                           int *nproj_per_atom, double *rc_max);

/* Timing of the most recent filter call on the context: total and
 * per-kernel-class milliseconds measured with HIP events on its stream.
 * times[0] total, [1] stencil steps, [2] nonlocal gather, [3] nonlocal
 * scatter; launches[0..3] the matching launch counts. */
lynx_status lynx_last_filter_timing(const lynx_ctx *ctx, double times[4], int launches[4]);
/* Enable (1) or disable (0) HIP-event timing of filter calls. */
lynx_status lynx_set_timing(lynx_ctx *ctx, int enable);
/* Filter calls as HIP graphs (1, the default) or direct launches (0).  With
 * graphs, the second call with the same buffers, column count and set-up
 * captures the degree's launches; later calls upload their step scalars to
 * a device table and replay the graph (no per-launch host work).  No
 * reference counterpart. */
lynx_status lynx_set_filter_graphs(lynx_ctx *ctx, int enable);
/* The practical HBM ceiling for the FP64 stencil step's access pattern on
 * this device: a streaming kernel that reads x and Y_(k-2) and writes y
 * (nd x ncol doubles each, 16-byte accesses) and re-reads one nd-sized Veff
 * column from L2, best of three grid sizes.  *gbps counts the step's
 * algorithmic bytes (3 x 8 x nd x ncol + 8 x nd) per second.  Allocates and
 * frees its own buffers.  No reference counterpart (a measurement aid). */
lynx_status lynx_stream_ceiling(lynx_ctx *ctx, long nd, int ncol, double *gbps);

/* Precision of the pass inside lynx_chefsi_solve.  FP32 filters a
 * single-precision copy of psi (complex64 for complex solves) and keeps the
 * Nd-sized steps in single precision: the overlap and projected-Hamiltonian
 * products (partial sums added in FP64), H applied to the orthonormalised
 * block, and the two rotations (by R^-1 and by the Ritz vectors, rounded to
 * FP32).  The N x N steps run in FP64: Cholesky, triangular inverse and the
 * dense eigensolve.  psi is returned in FP64.  BF16 is FP32 with the filter's
 * blocks stored in bfloat16 (float arithmetic, FP32 Veff and projectors),
 * for real solves on orthogonal cells with an even row length; otherwise it
 * runs as FP32.  No reference counterpart: the reference runs the pass in
 * FP64 only (EigenSolver.cpp:558-681). */
enum lynx_filter_precision { LYNX_FILTER_FP64 = 0, LYNX_FILTER_FP32 = 1, LYNX_FILTER_BF16 = 2 };
lynx_status lynx_set_filter_precision(lynx_ctx *ctx, int precision);
lynx_status lynx_get_filter_precision(const lynx_ctx *ctx, int *precision);
/* Single-precision Gram products of the reduced-precision pass (the overlap
 * dV X^T X and the projection dV X^T H X of real blocks).  LYNX_GRAM_FP32
 * runs them as FP32 GEMMs (rocBLAS) with FP64 chunk sums.  LYNX_GRAM_BF16X3
 * runs them as split-BF16x3 products on the matrix cores: each operand split
 * into a bfloat16 head and tail, hi.hi + hi.lo + lo.hi with FP32
 * accumulation per K chunk, chunks summed in FP64 (about 4e-6 relative
 * against FP64; FP32 GEMMs about 2e-7); the rotations stay FP32 GEMMs.
 * LYNX_GRAM_BF16X3_ROT also runs the pass's two rotations (by the inverse
 * Cholesky factor and by the Ritz vectors) as split-BF16x3 products from
 * the FP64 factors, the second written straight to the FP64 psi: faster per
 * pass, but their 2^-17 relative error in psi can lengthen the
 * reduced-precision phase of an SCF.  Complex blocks, and real blocks
 * whose columns are not 16-byte aligned, use FP32 GEMMs either way.  No
 * reference counterpart: the reference forms these products in FP64
 * (EigenSolver.cpp:621-660). */
enum lynx_gram_precision { LYNX_GRAM_FP32 = 0, LYNX_GRAM_BF16X3 = 1, LYNX_GRAM_BF16X3_ROT = 2 };
lynx_status lynx_set_gram_precision(lynx_ctx *ctx, int mode);
lynx_status lynx_get_gram_precision(const lynx_ctx *ctx, int *mode);
/* c = dV a^H b (n x n, FP64 for real dtypes, complex128 for complex ones,
 * written to host memory) for two blocks of n columns (leading dimensions
 * >= Nd, in host or device memory per mem) on the context's grid, through
 * the product the solve uses for dtype: for FP64 the overlap (a == b) on a
 * hand-written FP64 matrix-core kernel that computes the upper tiles only
 * (the solve's projection, whose Hermitian part it takes, too), other FP64 /
 * complex128 products as GEMMs; for single precision the FP32
 * GEMMs or split-BF16x3 products selected by lynx_set_gram_precision.  The overlap product of EigenSolver::orthogonalize
 * and project_hamiltonian (EigenSolver.cpp:621-660). */
lynx_status lynx_gram(lynx_ctx *ctx, int dtype, const void *a, int lda, const void *b, int ldb, int n, void *c,
                      int mem);
/* c = a q for a block of n columns (Nd rows, leading dimensions >= Nd; a
 * and c in host or device memory per mem, in dtype's precision) and an
 * n x n matrix q in host memory, FP64 (real dtypes) or complex128 (complex
 * ones), column-major: the rotation of EigenSolver::rotate_orbitals and of
 * the Cholesky step (EigenSolver.cpp:622-641, 674-681), through the product
 * the solve uses for dtype: FP64 / complex128 GEMMs; FP32 GEMMs of q
 * rounded to the block's precision; or, for real FP32 blocks with
 * LYNX_GRAM_BF16X3_ROT, split-BF16x3 matrix-core products of the block and
 * the FP64 q (the reduced-precision pass's rotations under that setting). */
lynx_status lynx_rotate(lynx_ctx *ctx, int dtype, const void *a, int lda, const double *q, int n, void *c, int ldc,
                        int mem);
/* Number of lynx_set_veff calls made on the context so far (accepted or
 * refused).  A caller that skips re-uploading an unchanged potential keeps
 * the generation of its own last upload and uploads again when another
 * caller (the eigen stage, another wrapper) has set a potential since.  No
 * reference counterpart. */
lynx_status lynx_get_veff_generation(const lynx_ctx *ctx, long *generation);
/* Number of CheFSI passes over one block completed on the context
 * (lynx_chefsi_solve calls, and each batch unit of lynx_batch_solve on it).
 * No reference counterpart (a diagnostic). */
lynx_status lynx_solve_count(const lynx_ctx *ctx, long *count);
/* Counters of the FP32 passes of this context.  *count: passes redone in
 * FP64 from the unchanged psi, because the filtered block was too
 * ill-conditioned for single precision (Cholesky factor diagonal spread above
 * 100x, overlap not positive definite, or Ritz values outside the spectral
 * bounds), as in the first passes from a random start.  *cholqr2 (may be
 * null): passes whose factor spread its diagonal by 10-100x and that took a
 * second Cholesky-QR of the orthonormalised block.  No reference
 * counterpart. */
lynx_status lynx_mixed_fallbacks(const lynx_ctx *ctx, long *count, long *cholqr2);
/* Grid point counts nx, ny, nz of the context. */
lynx_status lynx_get_grid_dims(const lynx_ctx *ctx, int dims[3]);

/* Launch shape of the fused stencil kernel.  By default the library times
 * the candidate shapes on the first call for each (grid, dtype, ncol) and
 * caches the fastest (its autotuner; no reference counterpart).
 * lynx_set_stencil_config forces a shape: threads per workgroup 256 or 512,
 * points_div 1 or 2 (points per thread = base / points_div), z_chunks >= 1,
 * vector_runs 0..3 (real types: each thread owns 16-byte runs of x points;
 * with 2 the z-direction terms are accumulated forward, with 3 as 2 with
 * loads issued two planes ahead instead of one; adding 4 starts every band
 * row at a lane multiple of 16, so 16-byte LDS accesses of one cycle never
 * span two rows; ignored
 * for complex types, non-orthogonal cells, and grids or buffers where nx or
 * the leading dimensions are not multiples of the run or the pointers are
 * not 16-byte aligned); (0, 0, 0, 0) returns to tuning.  Rows of more than
 * 116 points (up to 244) run on 512-thread shapes only, without vector runs
 * (a 256-point LDS row pitch), and not for complex128 blocks on
 * non-orthogonal cells (LYNX_E_UNSUPPORTED); batched units with such rows
 * run in 512-thread groups of the same pitch, under the same exclusion.
 * lynx_get_stencil_config
 * reports the shape in use and, in tune_ms (capacity *n_tune on input, count
 * on output), the timed candidates in order vector_runs 0..3 x row alignment
 * 0..1 (none for 0) x threads
 * 256..512 x points_div 1..2 x z_chunks 1..4 (shapes that do not fit a grid
 * are skipped). */
lynx_status lynx_set_stencil_config(lynx_ctx *ctx, int threads, int points_div, int z_chunks,
                                    int vector_runs);
lynx_status lynx_get_stencil_config(const lynx_ctx *ctx, int dtype, int ncol, int *threads,
                                    int *points_div, int *z_chunks, int *vector_runs, double *tune_ms,
                                    int *n_tune);

/* ---- SCF eigen stage --------------------------------------------------------
 * The per-iteration contract of SCF::solve_eigenproblem (SCF.cpp:310-467)
 * with the spectral bounds of SCF::estimate_spectral_bounds
 * (SCF.cpp:883-935): the caller's SCF loop (density, potential, mixing)
 * calls lynx_scf_eig_step once per SCF iteration with the current Veff and
 * residual.  Iteration 0 starts a run: Lanczos bounds per spin at the first
 * k-point, then rho_trigger CheFSI passes; later iterations run
 * nchefsi; after each pass eig_min[s] <- eig(s, k0)[0] (iterations > 0),
 * lambda_cutoff <- max_s eig(s, k0)[N-1] + 0.1, and the occupations and Fermi
 * level are recomputed (Occupation.cpp:53-176).  With fp64_switch_residual
 * > 0 the pass runs in reduced precision (low_precision: LYNX_FILTER_FP32,
 * the default when 0, or LYNX_FILTER_BF16; see lynx_set_filter_precision) in
 * iteration 0 and while scf_residual exceeds it, and in FP64 afterwards. */
typedef struct lynx_scf_eig lynx_scf_eig;
typedef struct {
    int nspin;                    /* 1 or 2 */
    int nkpt;                     /* k-points of this context */
    int complex_kpt;              /* 1: complex Bloch blocks (k-point runs), 0: real (Gamma) */
    int nband;
    int cheb_degree;
    int rho_trigger;              /* SCFParams::rho_trigger (4 in the reference) */
    int nchefsi;                  /* SCFParams::nchefsi (1) */
    double nelectron;
    double beta;                  /* 1 / (k_B T) in 1/Ha */
    int smearing;                 /* 0 Fermi-Dirac, 1 Gaussian */
    double fp64_switch_residual;  /* 0: FP64 filter throughout */
    int low_precision;            /* before the switch: LYNX_FILTER_FP32 (or 0) or LYNX_FILTER_BF16 */
    int gram_precision;           /* Gram products of the reduced-precision passes: LYNX_GRAM_FP32 (0),
                                     LYNX_GRAM_BF16X3 (split-BF16x3 on the matrix cores) or
                                     LYNX_GRAM_BF16X3_ROT (the rotations too) */
} lynx_scf_eig_params;
lynx_status lynx_scf_eig_create(lynx_scf_eig **driver, lynx_ctx *ctx, const lynx_scf_eig_params *params,
                                const double *kpts, const double *kpt_weights);
lynx_status lynx_scf_eig_destroy(lynx_scf_eig *driver);
const char *lynx_scf_eig_last_error(const lynx_scf_eig *driver);
/* veff: nspin x Nd doubles (host); psi[s * nkpt + k]: blocks of nband
 * columns (ld, mem as lynx_chefsi_solve), updated in place; eig and occ:
 * nspin x nkpt x nband; *fermi the Fermi level; *precision_used the filter
 * precision of this step (LYNX_FILTER_*).  Passes over several complex
 * k-points on device blocks solve each spin's k-points together through
 * one lynx_batch_solve (one batch per spin, created on first use, owned
 * by the driver): FP64
 * passes as lynx_chefsi_solve runs them, reduced-precision passes with the
 * filter in complex64 and the subspace steps in complex128 (gram_precision
 * does not apply to them).  Host blocks loop over the k-points with
 * lynx_chefsi_solve. */
lynx_status lynx_scf_eig_step(lynx_scf_eig *driver, int scf_iter, double scf_residual, const double *veff,
                              void *const *psi, int ld, int mem, double *eig, double *occ, double *fermi,
                              int *precision_used);
/* Fermi level and occupations of nspin x nkpt x nband eigenvalues with
 * k-point weights (summing to 1) and the spin factor 2 for nspin = 1:
 * the root of sum w f(e - Ef) = nelectron, f Fermi-Dirac (smearing 0) or
 * 0.5 erfc(beta x) (1).  Host-only; replaces Occupation::compute
 * (Occupation.cpp:86-176).  occ (may be NULL) receives f(e - Ef). */
lynx_status lynx_fermi_level(int nspin, int nkpt, int nband, const double *eig, const double *kpt_weights,
                             double nelectron, double beta, int smearing, double *occ, double *fermi);
/* lynx_scf_eig_step with Veff in host (LYNX_MEM_HOST) or device
 * (LYNX_MEM_DEVICE) memory. */
lynx_status lynx_scf_eig_step_veff(lynx_scf_eig *driver, int scf_iter, double scf_residual, const double *veff,
                                   int veff_mem, void *const *psi, int ld, int mem, double *eig, double *occ,
                                   double *fermi, int *precision_used);
/* Current lambda_cutoff and per-spin eig_min / eig_max (nspin each). */
lynx_status lynx_scf_eig_state(const lynx_scf_eig *driver, double *lambda_cutoff, double *eig_min,
                               double *eig_max);

/* ---- SCF ------------------------------------------------------------------
 * The self-consistent field loop around the CheFSI eigen stage, on the
 * GPU in FP64: electron density from the orbital blocks, exchange-
 * correlation, the Poisson solve for the electrostatic potential, Pulay
 * mixing of the density with the Kerker preconditioner, and the total
 * energy.  Replaces SCF::run (SCF.cpp:110-291) with initialize_scf
 * (SCF.cpp:937-1028), compute_new_density, compute_scf_energy,
 * check_convergence and mix_and_update (SCF.cpp:469-791);
 * ElectronDensity::compute (ElectronDensity.cpp:55-140);
 * EffectivePotential::compute (EffectivePotential.cpp:44-176);
 * PoissonSolver::solve (PoissonSolver.cpp:43-166); Mixer::mix
 * (Mixer.cpp:116-266) with KerkerPreconditioner::apply
 * (KerkerPreconditioner.cpp:37-95); Energy::compute_all (Energy.cpp:140-200).
 * The context needs lynx_set_grid and lynx_set_atoms (psp8 files); the
 * driver owns the orbitals (device), drawn as Wavefunction::randomize does
 * with seed spin * 1000 + 1, and the density starts from the isolated-atom
 * densities.  Density mixing only; collinear spin; LDA (PZ81 or PW92) or
 * GGA PBE, written from the published forms (the reference calls libxc). */
enum lynx_xc { LYNX_XC_LDA_PZ = 0, LYNX_XC_LDA_PW = 1, LYNX_XC_GGA_PBE = 2 };
/* The two constant-coefficient periodic solves of the SCF tail, Poisson
 * (-lap phi = 4 pi (rho + b)) and Kerker ((-lap + kTF^2) z = f), by
 *   LYNX_SOLVE_SPECTRAL: the FD operator is diagonal in the discrete Fourier
 *     basis of the periodic grid, so each solve is one real-to-complex FFT,
 *     a division by the stencil's own symbol and one inverse FFT (rocFFT):
 *     exact to rounding, for orthogonal and non-orthogonal cells;
 *   LYNX_SOLVE_AAR: the reference's alternating Anderson-Richardson
 *     iteration with a Jacobi preconditioner (PoissonSolver.cpp:43-166,
 *     KerkerPreconditioner.cpp:37-95) to poisson_tol / precond_tol.
 * Both reach the same discrete solution; AAR stops at its tolerance. */
enum lynx_linear_solver { LYNX_SOLVE_SPECTRAL = 0, LYNX_SOLVE_AAR = 1 };
typedef struct lynx_scf lynx_scf;
typedef struct {
    int nspin;                    /* 1 or 2 */
    int nkpt;                     /* k-points (kpts and weights given at create) */
    int complex_kpt;              /* 1: complex Bloch blocks; 0: real (Gamma point, nkpt 1) */
    int nband;
    int cheb_degree;
    int rho_trigger;              /* CheFSI passes in the first iteration (4 in the reference) */
    int nchefsi;                  /* passes per later iteration (1) */
    double beta;                  /* 1 / (k_B T) in 1/Ha */
    int smearing;                 /* 0 Fermi-Dirac, 1 Gaussian */
    int xc;                       /* LYNX_XC_* */
    double tol;                   /* density residual ||rho_out - rho_in|| / ||rho_out|| */
    int max_iter, min_iter;       /* 100 and 2 in the reference */
    double poisson_tol;           /* <= 0: tol / 100 (ParameterDefaults.cpp:41-43) */
    double precond_tol;           /* <= 0: h_eff^2 1e-3 (ParameterDefaults.cpp:45-47) */
    int mixing_history;           /* Pulay history (7) */
    double mixing_param;          /* 0.3 */
    int kerker;                   /* 1: Kerker preconditioner (the reference's default), 0: none */
    double fp64_switch_residual;  /* as lynx_scf_eig_params: reduced-precision passes above it */
    int low_precision;            /* LYNX_FILTER_FP32 or LYNX_FILTER_BF16 */
    int linear_solver;            /* LYNX_SOLVE_SPECTRAL (0) or LYNX_SOLVE_AAR (1) */
    int gram_precision;           /* as lynx_scf_eig_params */
} lynx_scf_params;
typedef struct {
    double etotal, eband, exc, ehart, eself, ec, entropy;   /* Ha (Energy::compute_all) */
    double fermi;                 /* Ha */
    double scf_error;             /* last density residual */
    int iterations;
    int converged;
    int natom;
    double seconds;               /* host wall time of lynx_scf_run */
} lynx_scf_result;
lynx_status lynx_scf_create(lynx_scf **scf, lynx_ctx *ctx, const lynx_scf_params *params, const double *kpts,
                            const double *kpt_weights);
lynx_status lynx_scf_destroy(lynx_scf *scf);
const char *lynx_scf_last_error(const lynx_scf *scf);
/* Starting density of the next lynx_scf_run in place of the atomic
 * densities, e.g. a converged density read back from a restart file
 * (host pointers, nd = Nx Ny Nz): the total density, and for two spins the
 * magnetization up - down (NULL: the total split evenly); each spin's
 * density is (total +- mag) / 2 floored at zero.  rho_total NULL restores
 * the atomic start.  Replaces SCF::set_initial_density (SCF.cpp:83-107) as
 * used for DensityIO restarts (SCF.cpp:742-761) and by pylynx's
 * Calculator::set_initial_density (python/src/Calculator.hpp:96-101). */
lynx_status lynx_scf_set_initial_density(lynx_scf *scf, const double *rho_total, int nd, const double *mag);
/* Runs the SCF loop from the atomic densities (or the density given by
 * lynx_scf_set_initial_density) and random orbitals to convergence (or
 * max_iter). */
lynx_status lynx_scf_run(lynx_scf *scf, lynx_scf_result *result);
/* Per-iteration record of the last run (capacity cap): total energy, density
 * residual, wall seconds, filter precision used; *n the iteration count. */
lynx_status lynx_scf_history(const lynx_scf *scf, int cap, double *etotal, double *error, double *seconds,
                             int *precision, int *n);
/* Eigenvalues and occupations (nspin x nkpt x nband), the input density of
 * the last iteration (nspin x Nd: total for one spin, up and down for two)
 * and Veff (nspin x Nd); any pointer may be NULL. */
lynx_status lynx_scf_get(const lynx_scf *scf, double *eig, double *occ, double *rho, double *veff);
/* The electrostatic potential phi and the XC energy density of the last
 * potential build (Nd each); any pointer may be NULL. */
lynx_status lynx_scf_get_fields(const lynx_scf *scf, double *phi, double *exc, double *vxc);
/* Atomic forces (Ha/Bohr, Cartesian, 3 per atom in the order of
 * lynx_set_atoms) of the state the last lynx_scf_run left (its final phi,
 * Vxc, orbitals and occupations): local (pseudocharge) term, nonlocal
 * (projector) term with <chi|psi> and <chi|grad psi> on the GPU, and the
 * core-correction XC term; components along the lattice directions mapped
 * to Cartesian (non-orthogonal cells), then the mean force removed.
 * Replaces Forces::compute (Forces.cpp:25-184, 186-420, 509-800,
 * 1108-1288).  parts (optional, 9 per atom): the local, nonlocal and
 * core-correction terms (Cartesian, before the mean is removed), each
 * 3 x natom.  LYNX_E_STATE before a completed run. */
lynx_status lynx_scf_forces(lynx_scf *scf, double *forces, double *parts);
/* One potential build from a given density (nspin x Nd as lynx_scf_get),
 * without the eigen stage: XC, Poisson (warm start from the previous phi)
 * and Veff; for testing the SCF tail alone. */
lynx_status lynx_scf_build_potential(lynx_scf *scf, const double *rho);

/* ---- Batched filter ------------------------------------------------------
 * A batch filters many independent units together: unit u applies the
 * operator of context ctxs[u] (grid, atoms and Veff set) at k-point
 * kpts[3u..3u+2] (lattice-aligned basis, as lynx_set_kpoint; NULL or ignored
 * for real dtypes) to its own block.  A context may serve several units
 * (the k-points of one structure); contexts with different grids mix
 * freely (a heterogeneous batch).  Each Chebyshev step is one launch per
 * kernel for all units, and the whole filter is replayed from a HIP graph
 * captured on the first call with the same buffers and degree.  The
 * reference filters k-points and structures one after another
 * (SCF.cpp:391-411 calling EigenSolver::solve_kpt, EigenSolver.cpp:834-842
 * per k-point); this is the batched form of lynx_cheb_filter. */
typedef struct lynx_batch lynx_batch;
lynx_status lynx_batch_create(lynx_batch **batch, int nunit, lynx_ctx *const *ctxs, const double *kpts);
lynx_status lynx_batch_destroy(lynx_batch *batch);
const char *lynx_batch_last_error(const lynx_batch *batch);
/* Per-unit blocks: x[u], y[u] with leading dimensions ldx[u], ldy[u] (0 or
 * NULL: Nd of the unit), ncol[u] columns, and per-unit filter bounds. */
lynx_status lynx_batch_filter(lynx_batch *batch, int dtype, void *const *x, const long *ldx, void *const *y,
                              const long *ldy, const int *ncol, int degree, const double *lambda_cutoff,
                              const double *eig_min, const double *eig_max, int mem);
/* 1 (default): capture the filter into a HIP graph and replay it; 0: launch eagerly. */
lynx_status lynx_batch_set_graph(lynx_batch *batch, int enable);
/* Filter precision of lynx_batch_solve: LYNX_FILTER_FP64 (default) or
 * LYNX_FILTER_FP32, the filter on single-precision copies of the blocks
 * (float / complex64) widened back before the FP64 / complex128 subspace
 * steps: the reduced-precision pass, batched (lynx_set_filter_precision
 * is its single-context counterpart; no reference counterpart, the
 * reference filters in FP64). */
lynx_status lynx_batch_set_filter_precision(lynx_batch *batch, int precision);
/* Device time of the last lynx_batch_filter (HIP events around the graph
 * launch or the eager launches) and the number of kernel launches it holds. */
lynx_status lynx_batch_last_timing(const lynx_batch *batch, double *ms, int *launches);
/* One CheFSI pass for every unit of the batch: lynx_chefsi_solve on each
 * unit's block psi[u] (device memory, leading dimension ldpsi[u], ncol[u]
 * orbitals), done together.  The filter is the batch's graph-replayed
 * filter; H times the orthonormalised blocks is one batched step; the
 * Gram products and rotations are batched GEMMs and the Cholesky
 * factorisations, triangular inverses and subspace eigensolves batched
 * rocSOLVER calls, over the units of equal (Nd, ncol, ldpsi).  dtype is
 * LYNX_F64 or LYNX_C128.  eigvals receives sum(ncol) values, unit after
 * unit, ascending within a unit; lambda_cutoff_out[u] (optional) the
 * unit's eig[ncol - 1] + 0.1.  Replaces the per-(spin, k-point) loop of
 * solve / solve_kpt calls (SCF.cpp:391-411, EigenSolver.cpp:738-842). */
lynx_status lynx_batch_solve(lynx_batch *batch, int dtype, void *const *psi, const long *ldpsi, const int *ncol,
                             double *eigvals, const double *lambda_cutoff, const double *eig_min,
                             const double *eig_max, int degree, double *lambda_cutoff_out);

#ifdef __cplusplus
}
#endif
#endif /* LYNX_CHEFSI_H */
