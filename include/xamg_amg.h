/*
 * libxamg.so -- host-side setup: the multigrid hierarchy the reference gets from hypre
 * BoomerAMG (src/hypre/detail/hypre_hierarchy.inl:31-52, hypre_base.inl:337-403), the
 * reference's test-matrix generator (examples/common/generator/cube.inl:28-174) and the
 * right-hand-side generator of the benchmarks.  Setup runs once, on the host.
 * All functions return 0 on success, < 0 on failure.
 */
#ifndef XAMG_AMG_H_
#define XAMG_AMG_H_

#include <stddef.h>
#include <stdint.h>

#include "xamg_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xamg_amg_s *xamg_amg;

/* the MultiGrid keys of src/param/xamg_params_details.h:179-209 that the classical
 * setup honours, with the reference defaults in comments */
typedef struct {
    float strength;     /* mg_strength_threshold 0.5 */
    float max_row_sum;  /* mg_max_row_sum 0.9 */
    int32_t coarse_size;/* mg_coarse_matrix_size 100 */
    int32_t max_levels; /* mg_max_levels 25 */
    float trunc_factor; /* mg_trunc_factor 0.25 */
    int32_t pmax;       /* mg_Pmax_elements 4 */
    int32_t second_pass;/* Ruge-Stueben second pass (0) */
    int32_t agg_num_levels; /* mg_agg_num_levels: levels coarsened aggressively (10) */
    int32_t num_paths;  /* mg_num_paths (3) */
    float agg_trunc_factor; /* mg_agg_trunc_factor 0.25 */
    int32_t agg_pmax;   /* mg_agg_Pmax_elements 4 */
    int32_t eig_iters;  /* power iterations for Chebyshev bounds (10) */
    uint64_t eig_seed;
    int32_t coarsen_type;    /* mg_coarsening_type: 6 Falgout (Ruge-Stueben on one
                                process, the default), 8 PMIS */
    int32_t interp_type;     /* mg_interpolation_type: 0 classical (default), 6 extended+i */
    int32_t agg_interp_type; /* mg_agg_interpolation_type: 4 multipass (default) */
    /* keys of levels 0..nlev_keys-1 (the "key=value@levX-Y" overrides of a per-level
     * hierarchy, src/hypre/detail/hypre_hierarchy.inl:54-108); other levels use the keys
     * above.  lev_keys may be NULL when nlev_keys is 0. */
    int32_t nlev_keys;
    const struct xamg_amg_level_keys_s *lev_keys;
} xamg_amg_params;

typedef struct xamg_amg_level_keys_s {
    float strength, max_row_sum, trunc_factor;
    int32_t pmax, coarsen_type, interp_type, agg_interp_type;
    float agg_trunc_factor;
    int32_t agg_pmax, num_paths;
} xamg_amg_level_keys;

void xamg_amg_default_params(xamg_amg_params *p);
int xamg_amg_setup(int64_t n, const uint32_t *row, const uint32_t *col, const double *val,
                   const xamg_amg_params *p, xamg_amg *out);
int xamg_amg_nlevels(xamg_amg H);
/* which: 0 A_l, 1 P_l, 2 R_l.  Pointers stay valid until xamg_amg_free. */
int xamg_amg_level(xamg_amg H, int level, int which, int64_t *nrows, int64_t *ncols,
                   int64_t *nnz, const uint32_t **row, const uint32_t **col,
                   const double **val);
int xamg_amg_eigs(xamg_amg H, int level, double *min_eig, double *max_eig);
/* re-estimate lambda_max of D^-1/2 A D^-1/2 on every level by `iters` power iterations */
int xamg_amg_estimate_eigs(xamg_amg H, int iters, uint64_t seed);
/* The steps of one coarsening level, for tests of what the algorithms fix (A square,
 * n rows): strength flags per nonzero (1 strong), the C/F splitting (1 C, -1 F) of
 * coarsen_type with the given PMIS seed, and the interpolation P (n x nc, interp_type
 * 0, 4 or 6) of a splitting, returned like a level of a hierarchy (free with
 * xamg_amg_free; which = 1 of level 0). */
int xamg_amg_strength(int64_t n, const uint32_t *row, const uint32_t *col, const double *val,
                      float theta, float max_row_sum, uint8_t *strong);
int xamg_amg_split(int64_t n, const uint32_t *row, const uint32_t *col, const double *val,
                   const uint8_t *strong, int coarsen_type, uint64_t seed, int8_t *cf);
int xamg_amg_interp(int64_t n, const uint32_t *row, const uint32_t *col, const double *val,
                    const uint8_t *strong, const int8_t *cf, int interp_type, float trunc,
                    int pmax, xamg_amg *out);
/* The parameters XAMG_solver_create would pass on for the XAMG_* parameter set `pid`
 * (include/xamg_c_api.h): the solver-side roles with the per-level smoother roles of
 * the "@levX-Y" overrides resolved (params.get("pre_smoother", l),
 * src/solvers/detail/mg.inl:113-131), and the AMG setup keys, per level for a per-level
 * hierarchy (src/hypre/detail/hypre_hierarchy.inl:54-108).  lev_keys (nullable) receives
 * setup->nlev_keys entries, at most XH_MAX_LEVELS; setup->lev_keys is set to NULL.
 * Returns what XAMG_solver_create would return for the parameters (e.g. -3 for an
 * unsupported coarsening type). */
int xamg_param_resolve(size_t pid, xh_params *prm, xamg_amg_params *setup,
                       xamg_amg_level_keys *lev_keys);
/* copy into a device-side hierarchy object of libxamg_hip */
int xamg_amg_to_device(xamg_amg H, xh_hier *out);
int xamg_amg_free(xamg_amg H);

/* 7-point cube Laplacian (cube.inl:84-146), one block; nnz of the result first */
int64_t xamg_cube_nnz(int64_t nx, int64_t ny, int64_t nz);
int xamg_cube(int64_t nx, int64_t ny, int64_t nz, uint32_t *row, uint32_t *col, double *val);
/* RHS entries uniform in [-1,1]: b[i] = 2*u(splitmix64(seed, i)) - 1, i = row*nv + k */
int xamg_rhs_uniform(uint64_t seed, int64_t count, double *b);

/* Variable-coefficient operators of the benchmark configurations (no reference
 * counterpart beyond the constant-coefficient cube):
 * k[i] = exp(sigma * z_i), z_i standard normal (counter-based, stream `seed`); k = 1
 * when sigma = 0 */
int xamg_lognormal_field(uint64_t seed, double sigma, int64_t count, double *k);
/* -div(k grad u), cell-centred finite volumes on nx*ny*nz cells, 7-point, harmonic-mean
 * face coefficients q_d*2k_Pk_Q/(k_P+k_Q), Dirichlet ghost cells (boundary face adds
 * 2 q_d k_P); row layout and diagonal order of cube.inl:84-146, so k = 1 gives the cube
 * matrix exactly.  kcell has nx*ny*nz entries. */
int64_t xamg_varcoef7_nnz(int64_t nx, int64_t ny, int64_t nz);
int xamg_varcoef7(int64_t nx, int64_t ny, int64_t nz, const double *kcell, uint32_t *row,
                  uint32_t *col, double *val);
/* -div(k grad u), trilinear finite elements on the interior nodes of an (nx+1)*(ny+1)*(nz+1)
 * cell grid of the unit cube (nx*ny*nz unknowns, Dirichlet nodes eliminated), k one value
 * per cell (kcell has (nx+1)*(ny+1)*(nz+1) entries, x fastest); every existing 27-point
 * neighbour is stored (face couplings are exactly 0), diagonal first, then ascending
 * columns; nnz = (3nx-2)(3ny-2)(3nz-2) */
int64_t xamg_q1fem27_nnz(int64_t nx, int64_t ny, int64_t nz);
int xamg_q1fem27(int64_t nx, int64_t ny, int64_t nz, const double *kcell, uint32_t *row,
                 uint32_t *col, double *val);
/* smooth RHS: b[p*nv+k] = sum of m in [min_bumps, max_bumps] Gaussian bumps
 * a*exp(-|x-c|^2/(2w^2)) (cell units, centres uniform over the grid, widths uniform in
 * [min_width, max_width], amplitudes standard normal), cut off below exp(-8) */
int xamg_rhs_bumps(uint64_t seed, int64_t nx, int64_t ny, int64_t nz, int nv, int min_bumps,
                   int max_bumps, double min_width, double max_width, double *b);
/* the bumps of xamg_rhs_bumps: count[k] per RHS, params[(k*max_bumps + q)*5 + {cx, cy, cz,
 * w, a}] (for evaluation on the device, xh_eval_bumps) */
int xamg_bump_params(uint64_t seed, int64_t nx, int64_t ny, int64_t nz, int nv, int min_bumps,
                     int max_bumps, double min_width, double max_width, int32_t *count,
                     double *params);

#ifdef __cplusplus
}
#endif

#endif
