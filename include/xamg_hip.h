/*
 * libxamg_hip.so -- thin C ABI over the MI355X (gfx950) solve phase of XAMG.
 *
 * The reference XAMG has no GPU path; its solve phase is the C++ template stack
 *   XAMG_solver_solve2 (api/c/xamg_c_api.cpp:522-531)
 *     -> base_solver<F,NV>::solve (src/solvers/solver.h:289-311)
 *       -> MergedPBiCGStab / PBiCGStab / BiCGStab / MergedBiCGStab::solve
 *          with the MultiGrid preconditioner (src/solvers/detail/mg.inl:166-334)
 *         -> blas2::Ax_y / Axpy / SGS (src/blas2/blas2.h:185-326) and the merged BLAS-1
 *            groups (src/blas/detail/specific/blas_merged_bicgstab.inl:33-257).
 * Each entry point below names the reference interface it replaces.  All functions
 * return 0 on success and a negative code on failure (xh_last_error() gives the
 * message); nothing aborts.  Plain pointers and sizes only.  Vectors are the
 * reference's NV-interleaved layout x[i*nv + k] (src/primitives/vector/vector.h:509-526).
 * Supported nv: 1, 2, 4, 8, 16, 32, 64.
 *
 * Threading: one host thread per solver object; objects are not thread-safe.
 */
#ifndef XAMG_HIP_H_
#define XAMG_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xh_matrix_s *xh_matrix;
typedef struct xh_hier_s *xh_hier;
typedef struct xh_solver_s *xh_solver;
typedef struct xh_batch_s *xh_batch;

/* error codes */
#define XH_OK 0
#define XH_EINVAL -1
#define XH_EHIP -2
#define XH_ENOTSUP -3
#define XH_ENOMEM -4

/* method ids (the reference's registry strings, src/solvers/solver.h:408-443) */
#define XH_IDENTITY 0
#define XH_JACOBI 1
#define XH_HSGS 2
#define XH_CHEBYSHEV 3
#define XH_DIRECT 4
#define XH_MULTIGRID 5
#define XH_BICGSTAB 6
#define XH_PBICGSTAB 7
#define XH_MERGED_BICGSTAB 8
#define XH_MERGED_PBICGSTAB 9
#define XH_PCG 10
#define XH_RBICGSTAB 11
#define XH_MERGED_RBICGSTAB 12
#define XH_NONE -1

/* Runtime init: replaces XAMG::init's device part (src/init.h:54-76; the reference's
 * sys/gpu_wrapper.cu only enumerates devices).  Selects the HIP device. */
int xh_init(int device);
int xh_finalize(void);
const char *xh_last_error(void);
/* device name, compute units, HBM bytes */
int xh_device_info(char *name, int name_len, int *compute_units, double *hbm_bytes);

/* Device memory helpers (the reference copies user buffers on upload/download,
 * api/c/xamg_c_api.cpp:350-373). */
int xh_malloc(void **dptr, size_t bytes);
int xh_free(void *dptr);
int xh_memcpy_h2d(void *dst, const void *src, size_t bytes);
int xh_memcpy_d2h(void *dst, const void *src, size_t bytes);
int xh_synchronize(void);
/* dst[0:count] = src[0:count] (fp64) with the library's copy kernel, then synchronize;
 * a kernel of known traffic (8*count read + 8*count written) for counter calibration */
int xh_copy_device(double *dst, const double *src, size_t count);
/* the same copy launched `reps` times on the library stream, timed with HIP events on that
 * stream: average milliseconds per launch (the measured streaming ceiling of the device) */
int xh_copy_device_timed(double *dst, const double *src, size_t count, int reps,
                         double *ms_per_launch);
/* dptr[0:bytes] = 0 on the library stream, then synchronize (XAMG_vector_set_val_d(0)) */
int xh_fill_zero(void *dptr, size_t bytes);

/* CSR upload: replaces XAMG_csr_matrix_alloc/fill + XAMG_matrix_construct
 * (api/c/xamg_c_api.cpp:404-480).  Row/col are u32 (XAMG_LIMIT_TYPES_USAGE),
 * values f64; prec 0 keeps f64, 1 stores f32 (matrix_block::assemble with reduced_prec,
 * src/primitives/matrix/matrix.h:79-93).  index_mode: 0 auto (u16 panel-relative
 * columns when every panel of >= 64 rows spans <= 65536 columns), 1 u32, 2 u16 (fails
 * if impossible).  Nonzeros keep their stored order: each row sum is accumulated in that
 * order, as the reference does. */
int xh_matrix_create(int64_t nrows, int64_t ncols, int64_t nnz, const uint32_t *row,
                     const uint32_t *col, const double *val, int prec, int index_mode,
                     xh_matrix *out);
/* index_bits: 32 or 16; panel_rows: rows per u16 panel (0 for u32) */
int xh_matrix_info(xh_matrix A, int64_t *nrows, int64_t *ncols, int64_t *nnz, int *prec,
                   int *index_bits, int64_t *panel_rows, size_t *device_bytes);
int xh_matrix_destroy(xh_matrix A);

/* SpMV on device vectors: y = A x (accumulate 0, backend::csr_Ax_y,
 * src/blas2/detail/csr_matvec.inl:149-260) or y += A x (accumulate 1, backend::csr_Axpy,
 * :33-144).  x, y have the matrix's precision. */
int xh_spmv(xh_matrix A, int nv, const void *dx, void *dy, int accumulate);
/* Same launch repeated `reps` times on the library stream, timed with HIP events on
 * that stream: average milliseconds per launch. */
int xh_spmv_timed(xh_matrix A, int nv, const void *dx, void *dy, int accumulate, int reps,
                  double *ms_per_launch);
/* Row schedule of the SpMV launches (no counterpart in the reference, which walks rows
 * in order on each core: csr_matvec.inl:33-260).  A square matrix whose rows couple
 * mostly to rows up to `fold` away (a structured-grid operator: fold = nx*ny) is
 * processed as an (nrows/fold) x fold array in column strips, one or a few per XCD,
 * so each x row is reused from L2; the per-row arithmetic, and so y, is unchanged.
 * xh_matrix_create detects fold (the most frequent largest positive column offset, if
 * it is the largest one of at least a quarter of the rows and >= 256).
 * xh_matrix_fold: set >= 0 overrides it (0: plain row order), set < 0 only queries. */
int xh_matrix_fold(xh_matrix A, int64_t set, int64_t *fold);

/* Structured-grid view of a square matrix (set at upload when it applies): rows are the
 * points of an nx x ny x nz grid in natural order (row = (z*ny + y)*nx + x) and every
 * entry couples grid neighbours (|dx|, |dy|, |dz| <= 1; 7- and 27-point operators).  Such
 * a matrix runs the SpMV as a sweep along x over tiles of grid lines with the x rows of
 * the tile and its halo lines staged in LDS.  No reference counterpart (the reference's csr_Ax_y is row-parallel, src/blas2/
 * detail/csr_matvec.inl:149-260); the arithmetic is that of the row-parallel kernel.
 * xh_matrix_grid: dims = {nx, ny, nz}, zeros when the matrix has no grid view.
 * xh_csr_grid: the same detection on a host CSR; code
 * (nullable, nnz bytes) receives 9(dz+1) + 3(dy+1) + (dx+1) per entry or 255, coded the
 * number of coded entries.  xh_grid_schedule: the work items of one launch, 4 values
 * each (first line y, first line z, x begin, x end). */
int xh_matrix_grid(xh_matrix A, int32_t *dims);
/* The 7-point stencil layout (no reference counterpart): a matrix with a grid view whose
 * every row holds exactly the diagonal and all its in-grid neighbours, in the order -z,
 * -y, -x, +x, +y, +z (the reference generators' order), also stores its values as 8
 * slots per row {C, -z, -y, -x, +x, +y, +z, unused}; with "spmv_st7_enable" its SpMVs for
 * nv >= 8 sweep GY x GZ grid lines per wave along x, reusing x(c-1), x(c) from registers
 * and the rows of the wave's own lines through cross-lane moves (same products, same
 * order: identical results).  Knobs "spmv_st7_waves" (waves per launch, default 3072),
 * "spmv_st7_min_seg" (default 32).
 * xh_matrix_layouts: flags bit 0 grid view, bit 1 stencil layout, bit 2 u16 columns. */
int xh_matrix_layouts(xh_matrix A, int32_t *flags);
int xh_csr_grid(int64_t nrows, const uint32_t *row, const uint32_t *col, int32_t *dims,
                uint8_t *code, int64_t *coded);
int xh_grid_schedule(int32_t nx, int32_t ny, int32_t nz, int prec, int nv, int dots,
                     int64_t *grid, int64_t *nitems, int64_t *ipw, int32_t *items);
/* the detection rule above on a host CSR matrix */
int xh_csr_fold(int64_t nrows, int64_t ncols, const uint32_t *row, const uint32_t *col,
                int64_t *fold);
/* The launch shape of an SpMV over nrows rows: workgroups, tiles, tiles per workgroup,
 * and (rows != NULL) the [begin, end) row range of every tile, 2*ntiles entries, in
 * processing order.  dots != 0: a launch that writes dot-product partials. */
int xh_spmv_schedule(int64_t nrows, int64_t nnz, int64_t fold, int prec, int nv, int dots,
                     int64_t *grid, int64_t *ntiles, int64_t *tpw, int64_t *rows);
/* Launch-shape knobs: "spmv_strip_budget" (bytes, default 8 MiB), "spmv_tile_passes"
 * (default 2), "spmv_min_passes" (default 1: in row order, at least this many passes of
 * rows per workgroup), "dot_groups" (default 1024, a multiple of 8 up to 4096: workgroups
 * of a launch that writes dot-product partials), "grid_dot_groups" (default 2048: the same
 * for the structured-grid SpMV), "vec_dot_groups" (default 2048: the same for the fused
 * vector updates), "spmv_fold_enable" (0/1), "spmv_fold_min_bytes" (default 2 MiB: fold only
 * when 2*fold*nv*sizeof(value) exceeds it), "spmv_grid_enable" (the
 * structured-grid SpMV of matrices with a grid view: 0 never, 1 always, 2 = default: when
 * the rows average 12 or more entries, e.g. 27-point operators), "spmv_grid_items" (default 2048),
 * "cheb_fuse_zero" (0/1, default 1: a Chebyshev smoother from a zero x runs its first two
 * steps as one SpMV that gathers b and forms the first step's values per entry,
 * bit-identical to two launches), "hsgs_offblock_csr" (0/1, default 1, read at solver
 * setup: a hybrid GS level also keeps the entries outside each row's core block as a
 * matrix of their own and computes the off-block residual as b - (that matrix) x,
 * bit-identical to skipping the in-block columns of A), "spmv_wide_batches" (0/1,
 * default 1: SpMVs at nv 8 and 32 over matrices of few rows and 32 or more entries per row
 * on average, at most 512 workgroups, issue 16 x loads per batch instead of 8; the sums
 * and their order are unchanged),
 * "spmv_grid_min_seg" (default 32), "unet_fuse_up" (0/1, default 1: U-Net inference reads
 * the decoder's upsampled inputs through the convolution instead of writing them),
 * "unet_conv_tile" (0/1, default 1: forward convolutions with 16 or 32 output channels
 * on volumes of w % 64 == h % 4 == d % 2 == 0, and with 48 on w % 32 == 0, read their
 * input from LDS-staged tiles). */
int xh_set_option(const char *key, int64_t value);

/* Compressed block: the reference's compact-index CSR block, as csr_matrix::compress
 * (src/primitives/matrix/detail/csr_matrix.inl:86-170) leaves it.  Empty rows are
 * dropped (row i of the block is row row_ind[i] of y, row_ind ascending, NULL when no row
 * was dropped) and empty columns are dropped (column c is column col_ind[c] + col_offset
 * of the full operand, col_ind ascending by construction).  col holds nnz compact column
 * ids of col_bytes = 1, 2 or 4 bytes each (the widths define_int_type picks,
 * src/bit_operations.h:119-129); they are decoded in the SpMM kernel.  xh_spmv on such a
 * matrix reads x as the compact vector of ncols rows (the reference's halo layout,
 * csr_matvec.inl:121-140) and writes rows row_ind[i] of y (y has y_rows rows): y += A x
 * (csr_Axpy) or y = A x on those rows only (csr_Ax_y, :227-249; other rows untouched).
 * The other solver operations do not take compressed blocks. */
int xh_matrix_create_compressed(int64_t y_rows, int64_t nrows, int64_t ncols, int64_t nnz,
                                const uint32_t *row, const void *col, int col_bytes,
                                const uint32_t *row_ind, const uint32_t *col_ind,
                                int64_t col_offset, int64_t x_rows, const double *val, int prec,
                                xh_matrix *out);
/* x_c[c][k] = x_full[col_ind[c] + col_offset][k] for the compact columns of a compressed
 * block (the decode col_ind[col[j]] + block_col_offset,
 * src/primitives/matrix/csr_matrix.h:290-297, applied to the operand); x_full has x_rows
 * rows, x_c has ncols rows, both nv-interleaved in the matrix precision */
int xh_matrix_gather(xh_matrix A, int nv, const void *dx_full, void *dx_c);

/* Smooth right-hand sides on the device (no reference counterpart; the learned-initial-
 * guess workload): db[p*nv + k] = sum over q < count[k] of
 * a exp(-|x_p - c|^2 / (2 w^2)) with (cx, cy, cz, w, a) = params[(k*max_bumps + q)*5 ..],
 * x_p the cell centre (x + 0.5, y + 0.5, z + 0.5), cells outside the box of half-width
 * ceil(4w) around c skipped (the parameters of xamg_bump_params, the definition of
 * xamg_rhs_bumps) */
int xh_eval_bumps(int nx, int ny, int nz, int nv, int max_bumps, const int32_t *count,
                  const double *params, double *db);
/* xh_eval_bumps without any device allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops).  The counts and bump parameters are staged in a static device array of
 * 164,096 bytes per GPU (held until the process ends) instead of hipMalloc'd buffers, so
 * nv and max_bumps must be at most 64 (XH_EINVAL otherwise).  Same results as
 * xh_eval_bumps.  All calls on one GPU share the staging array: calls that overlap on one
 * GPU (several host threads) race. */
int xh_eval_bumps_noalloc(int nx, int ny, int nz, int nv, int max_bumps,
                          const int32_t *count, const double *params, double *db);

/* y = D x, D dense row-major n x n (backend::dense_Ax_y, dense_matvec.inl:72-121) */
int xh_dense_matvec(int64_t n, int nv, int prec, const void *dD, const void *dx, void *dy);

/* most hierarchy levels with a role of their own (the reference's default
 * mg_max_levels is 25, src/param/xamg_params_details.h:186) */
#define XH_MAX_LEVELS 32

/* One parameter role of the reference param dictionary
 * (src/param/xamg_params_details.h:79-236). */
typedef struct {
    int32_t method;
    int32_t max_iters;
    float abs_tol, rel_tol;
    int32_t conv_check;
    float relax_factor;
    int32_t poly_order;
    float spectrum_fraction;
} xh_role;

typedef struct {
    xh_role solver, precond, pre, post, coarse;
    int32_t mg_cycle;        /* 0 V, 1 W, 2 F */
    int32_t mg_prec;         /* 0 fp64; 1 mg_reduced_precision (levels >= 1 fp32);
                                2 fp32 preconditioner including level 0 */
    int32_t hsgs_ncores;     /* HSGS block count when hsgs_block_rows == 0 */
    int32_t hsgs_block_rows; /* > 0: per-level blocks ceil(n/hsgs_block_rows) */
    int32_t coarse_fp64;     /* mg_prec 2: coarsest direct solve in fp64 */
    int32_t index_mode;      /* column index width of the finest operator, as above */
    int32_t graph;           /* 1: replay one captured iteration as a hipGraph */
    int32_t true_res_period; /* > 0: every k-th iteration the recurrence residual is
                                replaced by b - A x (fp64) and rho = (r, r0) recomputed;
                                the convergence test is unchanged.  No reference
                                counterpart (0, the default, is the reference behaviour);
                                not for MergedBiCGStab, which updates x after r */
    int32_t reserved[3];
    /* Per-level smoother roles: MultiGrid::setup builds the smoothers of level l from
     * params.get("pre_smoother", l) and params.get("post_smoother", l), the role lists
     * with the "key=value@levX-Y" overrides of that level applied (src/solvers/detail/
     * mg.inl:113-131, api/c/xamg_c_api.cpp:256-260).  Levels l < nlev_roles use
     * pre_lev[l] / post_lev[l]; the other levels use pre / post.  0: every level uses
     * pre / post.  The coarse-grid solver has one role (params.get("coarse_grid_solver")
     * takes no level). */
    int32_t nlev_roles;
    xh_role pre_lev[XH_MAX_LEVELS], post_lev[XH_MAX_LEVELS];
} xh_params;

/* Hierarchy (the output of hypre::get_full_hierarchy, hypre_hierarchy.inl:31-52):
 * level l holds A_l and, for l < nlevels-1, P_l (n_l x n_{l+1}) and R_l = P_l^T.
 * Arrays are copied; stored nonzero order is kept.  The HSGS smoother accumulates the
 * off-block part of a row in stored order, which equals the reference's block order
 * when rows are stored diagonal first and then by ascending column (as the host AMG
 * setup and the reference generators store them). */
int xh_hier_create(int nlevels, xh_hier *out);
int xh_hier_set_level(xh_hier H, int level, int64_t n, int64_t nnzA, const uint32_t *rowA,
                      const uint32_t *colA, const double *valA, int64_t nc, int64_t nnzP,
                      const uint32_t *rowP, const uint32_t *colP, const double *valP,
                      int64_t nnzR, const uint32_t *rowR, const uint32_t *colR,
                      const double *valR, double min_eig, double max_eig);
int xh_hier_destroy(xh_hier H);

/* Solver construction: replaces construct_solver_hierarchy<F,NV>
 * (src/solvers/solver.h:481-500).  Uploads the hierarchy once, in the precision the
 * parameters ask for, and allocates every device buffer for `nv` right-hand sides. */
int xh_solver_create(xh_hier H, const xh_params *prm, int nv, xh_solver *out);
/* host vectors b, x (x holds x0 on entry, the solution on return):
 * replaces XAMG_solver_solve2 (api/c/xamg_c_api.cpp:522-531) */
int xh_solver_solve(xh_solver S, const double *b, double *x);
/* device vectors (inputs resident in HBM): the timed path */
int xh_solver_solve_device(xh_solver S, const double *db, double *dx);
/* solver_stats (src/solvers/solver.h:48-66) */
int xh_solver_stats(xh_solver S, int32_t *iters, int32_t *converged, double *abs_res,
                    double *rel_res);
/* sqrt of the tested residual per convergence check: hist[check*nv + k] */
int xh_solver_history(xh_solver S, double *hist, int32_t max_checks, int32_t *nchecks);
/* one preconditioner application x = M^{-1} b on device vectors (x starts at zero) */
int xh_solver_precond_apply(xh_solver S, const double *db, double *dx);
/* x <- the finest level's pre-smoother applied `sweeps` times to A x = b from the given x
 * (MultiGrid preconditioner with an fp64 finest level; Chebyshev, Jacobi or HSGS role):
 * the pre-smoothing step of MultiGrid::MG_cycle (mg.inl:197-199) on its own.  No reference
 * counterpart as an entry point; used to damp the high-frequency error of a learned
 * initial guess before the solve. */
int xh_solver_smooth(xh_solver S, const double *db, double *dx, int32_t sweeps);
/* Per-kernel accounting of a solve: the reference's perf counters (src/io/perf.h:34-104
 * count the bytes each operation reads and writes per solve) restated per GPU launch.
 * Runs the init step and exactly `iters` iterations from x = dx (tolerances ignored, no
 * hipGraph: every launch is enqueued on its own, bracketed by two HIP events on the
 * solver's stream) and sums, per (kernel, level), the launches, their algorithmic HBM
 * bytes (each operand read once, each result written once) and their event-timed
 * milliseconds.  level: hierarchy level of a V-cycle launch, -1 for the outer Krylov
 * method.  Entries are sorted by time; *nstats receives their total count (only the first
 * max_stats are written).  The solver's statistics afterwards are those of this run. */
typedef struct {
    char kernel[80];
    int32_t level;
    int32_t launches;
    double bytes;
    double ms;
} xh_kernel_stat;
int xh_solver_profile(xh_solver S, const double *db, double *dx, int32_t iters,
                      xh_kernel_stat *stats, int32_t max_stats, int32_t *nstats);
/* renew the solver-role parameters (tolerances, max_iters, convergence_check) of an
 * existing solver: base_solver::renew_params (src/solvers/solver.h:277-288); the method
 * itself cannot change */
int xh_solver_renew(xh_solver S, const xh_role *solver);
/* device memory held by the solver (matrices + vectors) */
int xh_solver_info(xh_solver S, int *nlevels, size_t *device_bytes);
int xh_solver_destroy(xh_solver S);

/* Batched independent systems (no reference counterpart: the reference solves one system
 * per call).  nsys CSR systems of different sizes, system s with nv[s] in {1,2,4,8,16}
 * RHS, solved together by one solver configuration, each exactly as a single solve of
 * that system: prm->solver BiCGStab, PBiCGStab or MergedPBiCGStab; prm->precond Chebyshev
 * (order 1..16 on D^-1/2 A D^-1/2, bounds [0, lmax[s]] with the role's spectrum
 * fraction, chebyshev.inl:50-206) or Identity.  Every iteration is one sequence of
 * launches over all systems (one hipGraph when prm->graph), each launch a fixed grid
 * walking (system, row tile) work items; convergence is tracked per system. */
int xh_batch_create(int nsys, const int64_t *nrows, const int32_t *nv,
                    const uint32_t *const *row, const uint32_t *const *col,
                    const double *const *val, const double *lmax, const xh_params *prm,
                    xh_batch *out);
/* b of system s (n_s x nv_s, interleaved); every solve starts from x = 0 */
int xh_batch_set_rhs(xh_batch B, int s, const double *b);
int xh_batch_solve(xh_batch B);
int xh_batch_get_x(xh_batch B, int s, double *x);
/* iterations, per-RHS convergence / residuals of system s, and the time from the start of
 * the last xh_batch_solve until system s finished (device clock, ms) */
int xh_batch_stats(xh_batch B, int s, int32_t *iters, int32_t *conv, double *abs_res,
                   double *rel_res, double *latency_ms);
/* Autotuner: for each nv class present (1, 2, 4, 8, 16) times `reps` batched SpMVs per
 * candidate layout -- RHS per lane (the nv tile width of a lane: 1, or 2 for nv >= 2, so
 * nv or nv/2 lanes per row) times work-item size (1..32 passes of 256*w/nv rows) -- and
 * keeps the fastest; tile_passes (nullable) receives the 5 pass choices.
 * xh_batch_set_tiles sets the passes, xh_batch_set_layout passes and RHS per lane,
 * xh_batch_get_layout reads both.  The layout changes only the grouping of the dot
 * partial sums, not any RHS's arithmetic. */
int xh_batch_autotune(xh_batch B, int reps, int32_t *tile_passes);
/* Times `reps` launches of the batched Chebyshev step (y = D^-1/2 (c r + D^-1/2 A u), the
 * preconditioner's SpMV) over every work item with every system active; returns the mean
 * microseconds per launch and its algorithmic bytes (matrix, row pointers, D^-1/2, work
 * items, u and r read, y written once).  Saves and restores the solver state; overwrites
 * a work vector, so call it between solves.  No reference counterpart (measurement). */
int xh_batch_spmv_timed(xh_batch B, int reps, double *us_per_launch, double *alg_bytes);
int xh_batch_set_tiles(xh_batch B, const int32_t *tile_passes);
int xh_batch_set_layout(xh_batch B, const int32_t *tile_passes, const int32_t *rhs_per_lane);
int xh_batch_get_layout(xh_batch B, int32_t *tile_passes, int32_t *rhs_per_lane);
/* on = 1: every later xh_batch_solve is one cooperative (persistent) launch: the init step
 * and the iterations run on the device, the launch sequence of an iteration as phases
 * separated by grid-wide barriers, until every system is done or max_iters; the host
 * issues one launch and waits once.  Same work items and arithmetic as the per-iteration
 * launches: identical results.  on = 0 (default): one launch sequence (or hipGraph) per
 * iteration, the host polling the all-done flag. */
int xh_batch_set_persistent(xh_batch B, int on);
int xh_batch_destroy(xh_batch B);

#ifdef __cplusplus
}
#endif

#endif
