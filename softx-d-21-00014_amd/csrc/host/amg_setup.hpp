// Host-side classical AMG setup: builds the multigrid hierarchy that the reference
// obtains from hypre BoomerAMG (src/hypre/detail/hypre_hierarchy.inl:31-52,
// src/hypre/detail/hypre_base.inl:337-403).  Runs once per matrix, before the solve
// phase; its output (A_l, P_l, R_l = P_l^T, eigenvalue estimates) is the input of the
// GPU solve.
#pragma once

#include <cstdint>
#include <vector>

namespace xamg {
namespace amg {

struct CsrHost {
    int64_t nrows = 0, ncols = 0;
    std::vector<uint32_t> row, col;
    std::vector<double> val;
    int64_t nnz() const { return (int64_t)col.size(); }
};

// Coarsening and interpolation schemes (hypre BoomerAMG numbering, the values the
// reference's mg_coarsening_type / mg_interpolation_type / mg_agg_interpolation_type keys
// take, src/param/xamg_params_details.h:196-214)
enum CoarsenType : int { COARSEN_FALGOUT = 6, COARSEN_PMIS = 8 };
enum InterpType : int { INTERP_CLASSICAL = 0, INTERP_MULTIPASS = 4, INTERP_EXT_I = 6 };

// the keys that may differ per level ("key=value@levX-Y" on the preconditioner list with
// hypre_per_level_hierarchy, src/hypre/detail/hypre_hierarchy.inl:54-108)
struct LevelKeys {
    float strength = 0.5f;
    float max_row_sum = 0.9f;
    float trunc_factor = 0.25f;
    int pmax = 4;
    int coarsen_type = COARSEN_FALGOUT;
    int interp_type = INTERP_CLASSICAL;
    int agg_interp_type = INTERP_MULTIPASS;
    float agg_trunc_factor = 0.25f;
    int agg_pmax = 4;
    int num_paths = 3;
};

struct SetupParams {
    float strength = 0.5f;    // mg_strength_threshold
    float max_row_sum = 0.9f; // mg_max_row_sum
    int coarse_size = 100;    // mg_coarse_matrix_size
    int max_levels = 25;      // mg_max_levels
    float trunc_factor = 0.25f; // mg_trunc_factor
    int pmax = 4;             // mg_Pmax_elements
    int second_pass = 0;      // RS second pass (F-F strong pairs share a C point)
    int agg_num_levels = 10;  // mg_agg_num_levels: levels coarsened aggressively
    int num_paths = 3;        // mg_num_paths: paths of length <= 2 for a long-range link
    float agg_trunc_factor = 0.25f; // mg_agg_trunc_factor
    int agg_pmax = 4;         // mg_agg_Pmax_elements
    int eig_iters = 10;       // power iterations for the Chebyshev bounds
    uint64_t eig_seed = 12345;
    int coarsen_type = COARSEN_FALGOUT; // mg_coarsening_type (6: Falgout = RS on one process)
    int interp_type = INTERP_CLASSICAL; // mg_interpolation_type
    int agg_interp_type = INTERP_MULTIPASS; // mg_agg_interpolation_type
    uint64_t pmis_seed = 2747;  // random measure tie-break of PMIS
    // per-level keys: level l < per_level.size() uses per_level[l], the others the keys above
    std::vector<LevelKeys> per_level;
    LevelKeys keys(int level) const;
};

// the coarsening and interpolation types build() implements
bool coarsen_supported(int type);
bool interp_supported(int type);
bool agg_interp_supported(int type);

struct Hierarchy {
    std::vector<CsrHost> A, P, R;
    std::vector<double> min_eig, max_eig;
    int nlevels() const { return (int)A.size(); }
};

// rows reordered in place: diagonal first, then ascending column
void sort_rows_diag_first(CsrHost &m);
CsrHost transpose(const CsrHost &m);
CsrHost spgemm(const CsrHost &a, const CsrHost &b);
void build(const CsrHost &A0, const SetupParams &p, Hierarchy &H);

// The steps of one level, exposed for tests of what the algorithms fix:
// strength flags per nonzero; C/F splitting (1 = C, -1 = F); interpolation
std::vector<uint8_t> strength_of(const CsrHost &A, float theta, float max_row_sum);
std::vector<int8_t> coarsen(const CsrHost &A, const std::vector<uint8_t> &S, int type,
                            int second_pass, uint64_t seed);
CsrHost interpolate(const CsrHost &A, const std::vector<uint8_t> &S,
                    const std::vector<int8_t> &cf, int type, float trunc, int pmax);
// lambda_max of D^{-1/2} A D^{-1/2} by power iteration; lambda_min reported as 0
void estimate_eigs(const CsrHost &A, int iters, uint64_t seed, double &mn, double &mx);

} // namespace amg
} // namespace xamg
