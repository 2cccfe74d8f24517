/*
 * XAMG solve-phase oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A CPU restatement of the reference XAMG solve phase, used as the checker for the HIP
 * implementation.  Loops over rows and vector elements run on OpenMP threads (each element
 * is computed by one thread, so results do not depend on the thread count); every
 * reduction keeps the reference's order: sequential partial sums per emulated core
 * (dot_ncores, split_range blocks), added in core order.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or as the timed
 * CPU baseline), never as the thing that is shipped or measured on the GPU.
 *
 * Every function states the reference file:line it restates.  Arithmetic follows
 * the reference's generic build (g++ -O3 on x86-64, no FMA contraction): this file is
 * compiled with -ffp-contract=off so that every a*b+c rounds twice, as there.
 *
 * Parity pins (what checks each function) are listed in the header comment of
 * xamg_oracle.cpp next to each group.
 */
#ifndef XAMG_ORACLE_H_
#define XAMG_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* method ids -- same numbering as the product's xh_method enum */
enum {
    XO_IDENTITY = 0,
    XO_JACOBI = 1,
    XO_HSGS = 2,
    XO_CHEBYSHEV = 3,
    XO_DIRECT = 4,
    XO_MULTIGRID = 5,
    XO_BICGSTAB = 6,
    XO_PBICGSTAB = 7,
    XO_MERGED_BICGSTAB = 8,
    XO_MERGED_PBICGSTAB = 9,
    XO_PCG = 10,
    XO_RBICGSTAB = 11,
    XO_MERGED_RBICGSTAB = 12,
    XO_NONE = -1
};

/* CSR with 32-bit indices (the reference default, XAMG_LIMIT_TYPES_USAGE) */
typedef struct {
    int64_t nrows, ncols, nnz;
    const uint32_t *row; /* nrows+1 */
    const uint32_t *col; /* nnz */
    const double *val;   /* nnz; cast to float on reduced-precision levels */
} xo_csr;

#define XO_MAX_LEVELS 32

/* one parameter role (solver / preconditioner / pre_smoother / post_smoother /
 * coarse_grid_solver), src/param/xamg_params_details.h:79-236 */
typedef struct {
    int32_t method;
    int32_t max_iters;
    float abs_tol, rel_tol; /* float32, as the reference stores them */
    int32_t conv_check;
    float relax_factor;
    int32_t poly_order;
    float spectrum_fraction;
} xo_role;

typedef struct {
    xo_role solver, precond, pre, post, coarse;
    int32_t mg_cycle;        /* 0 V, 1 W, 2 F (mg_cycle key) */
    int32_t mg_prec;         /* 0 all fp64; 1 mg_reduced_precision=1 (levels>=1 fp32);
                                2 fp32 preconditioner including level 0 */
    int32_t hsgs_ncores;     /* emulated core count for the HSGS block split (>=1) */
    int32_t hsgs_block_rows; /* >0: per-level block count ceil(n/block_rows) */
    int32_t dot_ncores;      /* emulated core count for partial sums of dots (>=1) */
    int32_t coarse_fp64;     /* mg_prec==2: coarsest direct solve kept in fp64 */
    int32_t true_res_period; /* >0: every k-th iteration r = b - A x, rho = (r, r0)
                                (no reference counterpart; 0 = reference behaviour) */
    /* per-level smoother roles: level l < nlev_roles smooths with pre_lev[l] /
     * post_lev[l] (params.get("pre_smoother", l), src/solvers/detail/mg.inl:113-131) */
    int32_t nlev_roles;
    xo_role pre_lev[XO_MAX_LEVELS], post_lev[XO_MAX_LEVELS];
} xo_params;

typedef struct {
    int32_t nlevels;
    const xo_csr *A; /* [nlevels] */
    const xo_csr *P; /* [nlevels-1] */
    const xo_csr *R; /* [nlevels-1] */
    const double *min_eig, *max_eig; /* [nlevels] */
} xo_hier;

/* y = A x  (csr_Ax_y)  and  y += A x (csr_Axpy); x, y interleaved [n][nv] */
void xo_spmv_d(const xo_csr *A, int nv, const double *x, double *y, int accumulate);
void xo_spmv_f(const xo_csr *A, int nv, const float *x, float *y, int accumulate);

/* y = D x, dense row-major (dense_Ax_y) */
void xo_dense_matvec_d(int64_t n, int64_t m, const double *D, int nv, const double *x, double *y);
void xo_dense_matvec_f(int64_t n, int64_t m, const float *D, int nv, const float *x, float *y);

/* Gauss-Jordan inverse without pivoting (inverse.inl) */
void xo_inverse_d(int64_t n, const double *A, double *Ainv);
void xo_inverse_f(int64_t n, const float *A, float *Ainv);

/* Chebyshev polynomial coefficients (chebyshev.inl:50-121; order>4 by recurrence) */
int xo_cheby_coeffs(double min_eig, double max_eig, int order, double fraction, double *coeffs);

/* split_range / core ranges (misc.h:46-55, part.h:115-135) */
uint64_t xo_split_range(uint64_t size, int nblocks);
void xo_core_ranges(uint64_t size, int nblocks, uint64_t *offsets /* nblocks+1 */);

/* HSGS one sweep pair on a single level, double precision, with given block count */
void xo_hsgs_d(const xo_csr *A, int nv, int nblocks, float relax, int sweeps, const double *b,
               double *x, int x_is_zero);

/* Chebyshev smoother application (chebyshev.inl:124-206), double */
void xo_chebyshev_d(const xo_csr *A, int nv, double min_eig, double max_eig, int order,
                    float fraction, const double *b, double *x, int x_is_zero);

/* Jacobi smoother without convergence check (jacobi.inl:107-115), double */
void xo_jacobi_d(const xo_csr *A, int nv, float relax, int sweeps, const double *b, double *x,
                 int x_is_zero);

/* reference compress() of a CSR block (csr_matrix.inl:86-170): returns the compact
 * block; widths chosen by define_int_type (bit_operations.h:119-129) */
int xo_compress(const xo_csr *A, int64_t block_ncols, uint32_t *row_out /* nrows+1 */,
                uint32_t *col_out /* nnz */, uint32_t *row_ind /* nrows */,
                uint32_t *col_ind /* block_ncols */, int64_t *nrows_out, int64_t *ncols_out);
int xo_define_int_type(uint64_t range); /* 0:u8 1:u16 2:u32 3:u64 */

/* 7-point cube generator (examples/common/generator/cube.inl:28-174), single block */
int64_t xo_cube_nnz(int64_t nx, int64_t ny, int64_t nz);
void xo_cube(int64_t nx, int64_t ny, int64_t nz, uint32_t *row, uint32_t *col, double *val);

/* full solve; x holds x0 on entry.  returns 0, or <0 on a reference assert */
int xo_solve(const xo_hier *H, const xo_params *prm, int nv, const double *b, double *x,
             int32_t *iters, int32_t *converged, double *abs_res, double *rel_res);

/* same, but also records the per-iteration residual history res[it*nv+k]
 * (sqrt of the recurrence residual) up to max_hist iterations */
int xo_solve_hist(const xo_hier *H, const xo_params *prm, int nv, const double *b, double *x,
                  int32_t *iters, int32_t *converged, double *abs_res, double *rel_res,
                  double *hist, int32_t max_hist);

/* the same solve split into setup (hierarchy copies, smoother set-up, coarse inverse)
 * and solve, so that the solve alone can be timed */
typedef struct xo_solver_s *xo_solver;
int xo_solver_create(const xo_hier *H, const xo_params *prm, xo_solver *out);
int xo_solver_solve(xo_solver S, int nv, const double *b, double *x, int32_t *iters,
                    int32_t *converged, double *abs_res, double *rel_res);
void xo_solver_free(xo_solver S);

/* OpenMP threads of the oracle's parallel loops; results do not depend on them (the dot
 * partition is dot_ncores, not the thread count) */
int xo_set_threads(int n);
int xo_get_threads(void);

/* one application of the MultiGrid preconditioner to b (x = M^{-1} b, x starts at 0) */
int xo_mg_apply(const xo_hier *H, const xo_params *prm, int nv, const double *b, double *x);

#ifdef __cplusplus
}
#endif

#endif
