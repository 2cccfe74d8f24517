// Host side of the drop-in: the reference's XAMG_* C API (api/c/xamg_c_api.cpp:66-716)
// over the MI355X solve phase of libxamg_hip.so, the parameter dictionary with the
// reference keys and defaults (src/param/xamg_params_details.h:79-236), the classical
// AMG setup C ABI (include/xamg_amg.h) and the cube generator.
#include "xamg_amg.h"
#include "xamg_c_api.h"
#include "xamg_hip.h"

#include "amg_setup.hpp"

#include <sys/time.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

using xamg::amg::CsrHost;
using xamg::amg::Hierarchy;
using xamg::amg::LevelKeys;
using xamg::amg::SetupParams;

namespace {
thread_local std::string g_api_err;
} // namespace

namespace xamg {
int api_fail(int code, const std::string &m) {
    g_api_err = m;
    return code;
}
} // namespace xamg

namespace {

using xamg::api_fail;

uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

} // namespace

// ---------------------------------------------------------------------------
// AMG setup C ABI
// ---------------------------------------------------------------------------

struct xamg_amg_s {
    Hierarchy H;
};

extern "C" {

void xamg_amg_default_params(xamg_amg_params *p) {
    SetupParams d;
    p->strength = d.strength;
    p->max_row_sum = d.max_row_sum;
    p->coarse_size = d.coarse_size;
    p->max_levels = d.max_levels;
    p->trunc_factor = d.trunc_factor;
    p->pmax = d.pmax;
    p->second_pass = d.second_pass;
    p->agg_num_levels = d.agg_num_levels;
    p->num_paths = d.num_paths;
    p->agg_trunc_factor = d.agg_trunc_factor;
    p->agg_pmax = d.agg_pmax;
    p->eig_iters = d.eig_iters;
    p->eig_seed = d.eig_seed;
    p->coarsen_type = d.coarsen_type;
    p->interp_type = d.interp_type;
    p->agg_interp_type = d.agg_interp_type;
    p->nlev_keys = 0;
    p->lev_keys = nullptr;
}

int xamg_amg_setup(int64_t n, const uint32_t *row, const uint32_t *col, const double *val,
                   const xamg_amg_params *p, xamg_amg *out) {
    if (n <= 0 || !row || !col || !val || !out)
        return api_fail(-1, "xamg_amg_setup: bad arguments");
    CsrHost A;
    A.nrows = A.ncols = n;
    A.row.assign(row, row + n + 1);
    A.col.assign(col, col + row[n]);
    A.val.assign(val, val + row[n]);
    for (uint32_t c : A.col)
        if ((int64_t)c >= n)
            return api_fail(-1, "xamg_amg_setup: column index out of range");
    SetupParams sp;
    if (p) {
        sp.strength = p->strength;
        sp.max_row_sum = p->max_row_sum;
        sp.coarse_size = p->coarse_size;
        sp.max_levels = std::max(1, (int)p->max_levels);
        sp.trunc_factor = p->trunc_factor;
        sp.pmax = p->pmax;
        sp.second_pass = p->second_pass;
        sp.agg_num_levels = p->agg_num_levels;
        sp.num_paths = std::max(1, (int)p->num_paths);
        sp.agg_trunc_factor = p->agg_trunc_factor;
        sp.agg_pmax = p->agg_pmax;
        sp.eig_iters = p->eig_iters;
        sp.eig_seed = p->eig_seed;
        sp.coarsen_type = p->coarsen_type;
        sp.interp_type = p->interp_type;
        sp.agg_interp_type = p->agg_interp_type;
        if (p->nlev_keys < 0 || (p->nlev_keys > 0 && !p->lev_keys))
            return api_fail(-1, "xamg_amg_setup: bad per-level keys");
        for (int l = 0; l < p->nlev_keys; ++l) {
            const xamg_amg_level_keys &q = p->lev_keys[l];
            LevelKeys k;
            k.strength = q.strength;
            k.max_row_sum = q.max_row_sum;
            k.trunc_factor = q.trunc_factor;
            k.pmax = q.pmax;
            k.coarsen_type = q.coarsen_type;
            k.interp_type = q.interp_type;
            k.agg_interp_type = q.agg_interp_type;
            k.agg_trunc_factor = q.agg_trunc_factor;
            k.agg_pmax = q.agg_pmax;
            k.num_paths = std::max(1, (int)q.num_paths);
            sp.per_level.push_back(k);
        }
    }
    for (int l = -1; l < (int)sp.per_level.size(); ++l) {
        const LevelKeys k = sp.keys(l);
        if (!xamg::amg::coarsen_supported(k.coarsen_type) ||
            !xamg::amg::interp_supported(k.interp_type) ||
            !xamg::amg::agg_interp_supported(k.agg_interp_type))
            return api_fail(-3, "xamg_amg_setup: unsupported coarsening or interpolation type");
    }
    auto *h = new xamg_amg_s();
    xamg::amg::build(A, sp, h->H);
    *out = h;
    return 0;
}

int xamg_amg_nlevels(xamg_amg H) { return H ? H->H.nlevels() : -1; }

int xamg_amg_level(xamg_amg H, int level, int which, int64_t *nrows, int64_t *ncols,
                   int64_t *nnz, const uint32_t **row, const uint32_t **col,
                   const double **val) {
    if (!H || level < 0 || level >= H->H.nlevels() || which < 0 || which > 2)
        return api_fail(-1, "xamg_amg_level: bad level");
    if (which > 0 && level + 1 >= H->H.nlevels())
        return api_fail(-1, "xamg_amg_level: the coarsest level has no P/R");
    const CsrHost &m = which == 0 ? H->H.A[level] : (which == 1 ? H->H.P[level] : H->H.R[level]);
    if (nrows)
        *nrows = m.nrows;
    if (ncols)
        *ncols = m.ncols;
    if (nnz)
        *nnz = m.nnz();
    if (row)
        *row = m.row.data();
    if (col)
        *col = m.col.data();
    if (val)
        *val = m.val.data();
    return 0;
}

int xamg_amg_eigs(xamg_amg H, int level, double *min_eig, double *max_eig) {
    if (!H || level < 0 || level >= H->H.nlevels())
        return api_fail(-1, "xamg_amg_eigs: bad level");
    if (min_eig)
        *min_eig = H->H.min_eig[level];
    if (max_eig)
        *max_eig = H->H.max_eig[level];
    return 0;
}

int xamg_amg_estimate_eigs(xamg_amg H, int iters, uint64_t seed) {
    if (!H)
        return api_fail(-1, "null hierarchy");
    for (int l = 0; l < H->H.nlevels(); ++l)
        xamg::amg::estimate_eigs(H->H.A[l], iters, seed, H->H.min_eig[l], H->H.max_eig[l]);
    return 0;
}

int xamg_amg_to_device(xamg_amg H, xh_hier *out) {
    if (!H || !out)
        return api_fail(-1, "xamg_amg_to_device: bad arguments");
    const int nl = H->H.nlevels();
    xh_hier D = nullptr;
    int rc = xh_hier_create(nl, &D);
    if (rc)
        return api_fail(rc, xh_last_error());
    for (int l = 0; l < nl; ++l) {
        const CsrHost &A = H->H.A[l];
        const bool last = l + 1 == nl;
        const CsrHost *P = last ? nullptr : &H->H.P[l];
        const CsrHost *R = last ? nullptr : &H->H.R[l];
        rc = xh_hier_set_level(D, l, A.nrows, A.nnz(), A.row.data(), A.col.data(), A.val.data(),
                               P ? P->ncols : 0, P ? P->nnz() : 0, P ? P->row.data() : nullptr,
                               P ? P->col.data() : nullptr, P ? P->val.data() : nullptr,
                               R ? R->nnz() : 0, R ? R->row.data() : nullptr,
                               R ? R->col.data() : nullptr, R ? R->val.data() : nullptr,
                               H->H.min_eig[l], H->H.max_eig[l]);
        if (rc) {
            xh_hier_destroy(D);
            return api_fail(rc, xh_last_error());
        }
    }
    *out = D;
    return 0;
}

namespace {
int host_csr(int64_t n, const uint32_t *row, const uint32_t *col, const double *val,
             CsrHost &A) {
    if (n <= 0 || !row || !col || !val)
        return -1;
    A.nrows = A.ncols = n;
    A.row.assign(row, row + n + 1);
    A.col.assign(col, col + row[n]);
    A.val.assign(val, val + row[n]);
    for (uint32_t c : A.col)
        if ((int64_t)c >= n)
            return -1;
    return 0;
}
} // namespace

int xamg_amg_strength(int64_t n, const uint32_t *row, const uint32_t *col, const double *val,
                      float theta, float max_row_sum, uint8_t *strong) {
    CsrHost A;
    if (host_csr(n, row, col, val, A) || !strong)
        return api_fail(-1, "xamg_amg_strength: bad arguments");
    std::vector<uint8_t> S = xamg::amg::strength_of(A, theta, max_row_sum);
    std::copy(S.begin(), S.end(), strong);
    return 0;
}

int xamg_amg_split(int64_t n, const uint32_t *row, const uint32_t *col, const double *val,
                   const uint8_t *strong, int coarsen_type, uint64_t seed, int8_t *cf) {
    CsrHost A;
    if (host_csr(n, row, col, val, A) || !strong || !cf)
        return api_fail(-1, "xamg_amg_split: bad arguments");
    if (!xamg::amg::coarsen_supported(coarsen_type))
        return api_fail(-3, "xamg_amg_split: unsupported coarsening type");
    std::vector<uint8_t> S(strong, strong + A.nnz());
    std::vector<int8_t> c = xamg::amg::coarsen(A, S, coarsen_type, 0, seed);
    std::copy(c.begin(), c.end(), cf);
    return 0;
}

int xamg_amg_interp(int64_t n, const uint32_t *row, const uint32_t *col, const double *val,
                    const uint8_t *strong, const int8_t *cf, int interp_type, float trunc,
                    int pmax, xamg_amg *out) {
    CsrHost A;
    if (host_csr(n, row, col, val, A) || !strong || !cf || !out)
        return api_fail(-1, "xamg_amg_interp: bad arguments");
    if (!xamg::amg::interp_supported(interp_type) && interp_type != 4)
        return api_fail(-3, "xamg_amg_interp: unsupported interpolation type");
    std::vector<uint8_t> S(strong, strong + A.nnz());
    std::vector<int8_t> c(cf, cf + n);
    auto *h = new xamg_amg_s();
    h->H.A.push_back(A);
    h->H.P.push_back(xamg::amg::interpolate(A, S, c, interp_type, trunc, pmax));
    h->H.R.push_back(xamg::amg::transpose(h->H.P[0]));
    CsrHost Ac;
    Ac.nrows = Ac.ncols = h->H.P[0].ncols;
    Ac.row.assign(Ac.nrows + 1, 0);
    h->H.A.push_back(Ac);
    h->H.min_eig.assign(2, 0.0);
    h->H.max_eig.assign(2, 0.0);
    *out = h;
    return 0;
}

int xamg_amg_free(xamg_amg H) {
    delete H;
    return 0;
}

int64_t xamg_cube_nnz(int64_t nx, int64_t ny, int64_t nz) {
    int64_t nnz = 0;
    for (int64_t kk = 0; kk < nz; ++kk)
        for (int64_t jj = 0; jj < ny; ++jj)
            for (int64_t ii = 0; ii < nx; ++ii)
                nnz += 1 + (kk > 0) + (jj > 0) + (ii > 0) + (ii < nx - 1) + (jj < ny - 1) +
                       (kk < nz - 1);
    return nnz;
}

int xamg_cube(int64_t nx, int64_t ny, int64_t nz, uint32_t *row, uint32_t *col, double *val) {
    // generate_cube (examples/common/generator/cube.inl:84-146) for one block: row
    // entries are the diagonal, then -z, -y, -x, +x, +y, +z neighbours; a missing
    // neighbour (Dirichlet ghost cell) adds 2q to the diagonal
    if (nx < 1 || ny < 1 || nz < 1 || !row || !col || !val)
        return api_fail(-1, "xamg_cube: bad arguments");
    const double hx = 1.0 / nx, hy = 1.0 / ny, hz = 1.0 / nz;
    const double q1 = 1.0 / hx / hx, q2 = 1.0 / hy / hy, q3 = 1.0 / hz / hz;
    uint64_t c = 0;
    row[0] = 0;
    const int64_t n = nx * ny * nz;
    for (int64_t indx = 0; indx < n; ++indx) {
        const int64_t ii = indx % nx, jj = (indx / nx) % ny, kk = indx / (nx * ny);
        const uint64_t d = c;
        col[c] = (uint32_t)indx;
        val[c++] = 0.0;
        auto nb = [&](bool has, int64_t off, double q, int64_t dim) {
            if (has) {
                val[d] += q;
                col[c] = (uint32_t)(indx + off);
                val[c++] = -q;
            } else if (dim > 1) {
                val[d] += 2.0 * q;
            }
        };
        nb(kk > 0, -nx * ny, q3, nz);
        nb(jj > 0, -nx, q2, ny);
        nb(ii > 0, -1, q1, nx);
        nb(ii < nx - 1, 1, q1, nx);
        nb(jj < ny - 1, nx, q2, ny);
        nb(kk < nz - 1, nx * ny, q3, nz);
        row[indx + 1] = (uint32_t)c;
    }
    return 0;
}

int xamg_rhs_uniform(uint64_t seed, int64_t count, double *b) {
    if (count < 0 || (count > 0 && !b))
        return api_fail(-1, "xamg_rhs_uniform: bad arguments");
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; ++i) {
        const uint64_t z = splitmix64(seed * 0xD1B54A32D192ED03ull + (uint64_t)i);
        b[i] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
    return 0;
}

} // extern "C"

// ---------------------------------------------------------------------------
// parameter dictionary (src/param/params.h, xamg_params_details.h)
// ---------------------------------------------------------------------------

namespace {

using ParamList = std::map<std::string, std::string>;

// one level-ranged override "key=value@levX[-Y]" (examples/common/cmdline/cmdline.h:79-97):
// levels lo..hi, hi = INT32_MAX for "E" (the last level)
struct LevelOverride {
    std::string key, value;
    int lo = 0, hi = 0;
};

struct GlobalParams {
    std::map<std::string, ParamList> roles;
    std::map<std::string, std::vector<LevelOverride>> overrides;
};

// "value@levX", "value@levX-Y" or "value@levX-E" -> value and the level range
bool parse_level_value(const std::string &v, LevelOverride &o) {
    const size_t at = v.rfind("@lev");
    if (at == std::string::npos || at + 4 >= v.size())
        return false;
    o.value = v.substr(0, at);
    const std::string r = v.substr(at + 4);
    const size_t dash = r.find('-');
    auto num = [](const std::string &t, int &out) {
        if (t.empty() || t.find_first_not_of("0123456789") != std::string::npos)
            return false;
        out = std::atoi(t.c_str());
        return true;
    };
    if (!num(r.substr(0, dash), o.lo))
        return false;
    if (dash == std::string::npos) {
        o.hi = o.lo;
    } else {
        const std::string y = r.substr(dash + 1);
        if (y == "E")
            o.hi = INT32_MAX;
        else if (!num(y, o.hi))
            return false;
    }
    return !o.value.empty() && o.hi >= o.lo;
}

// Keys a level override may set, per role.  Smoother roles: the keys of the smoother
// families (xamg_params_details.h:146-236) the GPU smoothers read.  Preconditioner: the
// MultiGrid keys the host setup honours level by level (hypre_per_level_hierarchy,
// src/hypre/detail/hypre_hierarchy.inl:54-108).  Every other key or role is rejected.
bool override_key_ok(const std::string &role, const std::string &key) {
    static const char *smoother[] = {"method", "max_iters", "relax_factor", "polynomial_order",
                                     "spectrum_fraction", "convergence_check",
                                     "convergence_info", "abs_tolerance", "rel_tolerance"};
    static const char *mg[] = {"mg_coarsening_type", "mg_interpolation_type",
                               "mg_agg_interpolation_type", "mg_strength_threshold",
                               "mg_max_row_sum", "mg_trunc_factor", "mg_Pmax_elements",
                               "mg_agg_trunc_factor", "mg_agg_Pmax_elements", "mg_num_paths",
                               "mg_nonGalerkin_tol"};
    if (role == "pre_smoother" || role == "post_smoother") {
        for (const char *k : smoother)
            if (key == k)
                return true;
        return false;
    }
    if (role == "preconditioner") {
        for (const char *k : mg)
            if (key == k)
                return true;
    }
    return false;
}

std::string fmt_float(float v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.9g", (double)v);
    return b;
}

bool has(const ParamList &l, const std::string &k) { return l.find(k) != l.end(); }

void set_missing(ParamList &l, const std::string &k, const std::string &v) {
    if (!has(l, k))
        l[k] = v;
}

// set_family_defaults (xamg_params_details.h:146-236)
void family_defaults(ParamList &l, const std::string &list_name) {
    const std::string fam = has(l, "method") ? l["method"] : "";
    if (fam == "Identity" || fam == "Direct")
        return;
    if (fam == "Chebyshev") {
        set_missing(l, "convergence_check", "0");
        set_missing(l, "convergence_info", "0");
        set_missing(l, "polynomial_order", "2");
        set_missing(l, "spectrum_fraction", fmt_float(0.3f));
        return;
    }
    int iters = 0;
    float abs_tol = 0.f, rel_tol = 0.f;
    if (list_name == "solver") {
        abs_tol = 1.e-8f;
        rel_tol = 1.e-8f;
        set_missing(l, "convergence_check", "1");
    } else {
        set_missing(l, "convergence_check", "0");
    }
    set_missing(l, "convergence_info", "0");
    if (fam == "Jacobi" || fam == "HSGS") {
        iters = list_name == "solver" ? 100 : 2;
        set_missing(l, "relax_factor", fmt_float(0.99f));
    } else if (fam == "MultiGrid") {
        iters = list_name == "solver" ? 30 : 1;
        set_missing(l, "hypre_log", "1");
        set_missing(l, "hypre_per_level_hierarchy", "0");
        set_missing(l, "mg_reduced_precision", "0");
        set_missing(l, "mg_cycle", "0");
        set_missing(l, "mg_max_levels", "25");
        set_missing(l, "mg_coarse_matrix_size", "100");
        set_missing(l, "mg_strength_threshold", fmt_float(0.5f));
        set_missing(l, "mg_trunc_factor", fmt_float(0.25f));
        set_missing(l, "mg_num_paths", "3");
        set_missing(l, "mg_coarsening_type", "6");
        set_missing(l, "mg_interpolation_type", "0");
        set_missing(l, "mg_max_row_sum", fmt_float(0.9f));
        set_missing(l, "mg_Pmax_elements", "4");
        set_missing(l, "mg_nonGalerkin_tol", fmt_float(0.0f));
        set_missing(l, "mg_agg_num_levels", "10");
        set_missing(l, "mg_agg_interpolation_type", "4");
        set_missing(l, "mg_agg_trunc_factor", fmt_float(0.25f));
        set_missing(l, "mg_agg_P12_trunc_factor", fmt_float(0.0f));
        set_missing(l, "mg_agg_Pmax_elements", "4");
        set_missing(l, "mg_agg_P12max_elements", "4");
    } else {
        iters = list_name == "solver" ? 50 : 1;
    }
    set_missing(l, "abs_tolerance", fmt_float(abs_tol));
    set_missing(l, "rel_tolerance", fmt_float(rel_tol));
    set_missing(l, "max_iters", std::to_string(iters));
}

void set_defaults(GlobalParams &g) {
    // set_dictionary_defaults (xamg_params_details.h:131-144)
    bool mg = false;
    for (auto &r : g.roles)
        if (has(r.second, "method") && r.second["method"] == "MultiGrid")
            mg = true;
    if (mg) {
        if (!g.roles.count("pre_smoother"))
            g.roles["pre_smoother"]["method"] = "Jacobi";
        if (!g.roles.count("post_smoother"))
            g.roles["post_smoother"]["method"] = "Jacobi";
        if (!g.roles.count("coarse_grid_solver"))
            g.roles["coarse_grid_solver"]["method"] = "Direct";
    }
    for (auto &r : g.roles)
        family_defaults(r.second, r.first);
}

int method_id(const std::string &m) {
    static const std::map<std::string, int> ids = {
        {"Identity", XH_IDENTITY},         {"Jacobi", XH_JACOBI},
        {"HSGS", XH_HSGS},                 {"Chebyshev", XH_CHEBYSHEV},
        {"Direct", XH_DIRECT},             {"MultiGrid", XH_MULTIGRID},
        {"BiCGStab", XH_BICGSTAB},         {"PBiCGStab", XH_PBICGSTAB},
        {"MergedBiCGStab", XH_MERGED_BICGSTAB}, {"MergedPBiCGStab", XH_MERGED_PBICGSTAB},
        {"PCG", XH_PCG},                   {"RBiCGStab", XH_RBICGSTAB},
        {"MergedRBiCGStab", XH_MERGED_RBICGSTAB}};
    auto it = ids.find(m);
    return it == ids.end() ? -100 : it->second;
}

long geti(const ParamList &l, const std::string &k, long d) {
    auto it = l.find(k);
    return it == l.end() ? d : std::strtol(it->second.c_str(), nullptr, 10);
}

float getf(const ParamList &l, const std::string &k, float d) {
    auto it = l.find(k);
    return it == l.end() ? d : std::strtof(it->second.c_str(), nullptr);
}

int role_of(const ParamList &l, xh_role &r) {
    r.method = has(l, "method") ? method_id(l.at("method")) : XH_NONE;
    if (r.method == -100)
        return api_fail(-3, "unknown or out-of-scope method '" + l.at("method") + "'");
    r.max_iters = (int32_t)geti(l, "max_iters", 1);
    r.abs_tol = getf(l, "abs_tolerance", 0.f);
    r.rel_tol = getf(l, "rel_tolerance", 0.f);
    r.conv_check = (int32_t)geti(l, "convergence_check", 0);
    r.relax_factor = getf(l, "relax_factor", 0.99f);
    r.poly_order = (int32_t)geti(l, "polynomial_order", 2);
    r.spectrum_fraction = getf(l, "spectrum_fraction", 0.3f);
    return 0;
}

// ---- object tables (api/c/xamg_c_api.cpp:66-85) ----

struct Vector {
    std::vector<double> d;
    uint64_t size = 0;
    int nv = 0;
    uint64_t offset = 0;
};

struct Csr {
    int64_t nrows = 0, nnz = 0;
    std::vector<uint32_t> row, col;
    std::vector<double> val;
    int row_off = 0, col_off = 0;
};

struct Matrix {
    std::shared_ptr<Csr> csr;
};

struct Solver {
    xh_solver gpu = nullptr;
    xh_hier hier = nullptr;
    int nv = 0;
    size_t xid = (size_t)-1, bid = (size_t)-1;
    std::vector<int32_t> conv;
    int32_t iters = 0;
    std::vector<double> abs_res, rel_res;
    ~Solver() {
        if (gpu)
            xh_solver_destroy(gpu);
        if (hier)
            xh_hier_destroy(hier);
    }
};

std::vector<std::unique_ptr<GlobalParams>> g_params;
std::vector<std::unique_ptr<std::vector<std::pair<std::string, std::string>>>> g_ovs;
std::vector<int> g_parts;
std::vector<std::unique_ptr<Vector>> g_vecs;
std::vector<std::unique_ptr<Csr>> g_csrs;
std::vector<std::unique_ptr<Matrix>> g_mats;
std::vector<std::unique_ptr<Solver>> g_solvers;
int g_nv = 2;

template <typename T>
T *lookup(std::vector<std::unique_ptr<T>> &tab, size_t id) {
    return id < tab.size() ? tab[id].get() : nullptr;
}

bool nv_supported(int nv) {
    return nv == 1 || nv == 2 || nv == 4 || nv == 8 || nv == 16 || nv == 32 || nv == 64;
}

// coarsening / interpolation types and Galerkin products the host setup implements
int check_setup_keys(const LevelKeys &k, float non_galerkin_tol, const std::string &where) {
    if (!xamg::amg::coarsen_supported(k.coarsen_type))
        return api_fail(-3, "mg_coarsening_type " + std::to_string(k.coarsen_type) + where +
                                " is not supported (6 Falgout/RS, 8 PMIS)");
    if (!xamg::amg::interp_supported(k.interp_type))
        return api_fail(-3, "mg_interpolation_type " + std::to_string(k.interp_type) + where +
                                " is not supported (0 classical, 6 extended+i)");
    if (!xamg::amg::agg_interp_supported(k.agg_interp_type))
        return api_fail(-3, "mg_agg_interpolation_type " + std::to_string(k.agg_interp_type) +
                                where + " is not supported (4 multipass)");
    if (non_galerkin_tol != 0.f)
        return api_fail(-3, "mg_nonGalerkin_tol" + where +
                                " must be 0 (Galerkin coarse operators only)");
    return 0;
}

int build_xh_params(GlobalParams &g, xh_params &p, SetupParams &sp, bool &mg) {
    std::memset(&p, 0, sizeof p);
    if (!g.roles.count("solver"))
        return api_fail(-1, "parameter list has no 'solver' role");
    int rc;
    if ((rc = role_of(g.roles["solver"], p.solver)))
        return rc;
    // extension of this build (no reference key): periodic residual replacement
    p.true_res_period = (int32_t)geti(g.roles["solver"], "true_res_period", 0);
    p.precond.method = XH_NONE;
    p.pre.method = p.post.method = p.coarse.method = XH_NONE;
    mg = false;
    if (g.roles.count("preconditioner")) {
        ParamList &l = g.roles["preconditioner"];
        if ((rc = role_of(l, p.precond)))
            return rc;
        mg = p.precond.method == XH_MULTIGRID;
        p.mg_cycle = (int32_t)geti(l, "mg_cycle", 0);
        p.mg_prec = geti(l, "mg_reduced_precision", 0) ? 1 : 0;
        // extensions of this build (no reference key): fp32 preconditioner including the
        // finest level, HSGS block size, coarse solve in fp64, hipGraph replay
        if (geti(l, "mg_fp32_precond", 0))
            p.mg_prec = 2;
        p.hsgs_block_rows = (int32_t)geti(l, "hsgs_block_rows", 64);
        p.hsgs_ncores = (int32_t)geti(l, "hsgs_ncores", 1);
        p.coarse_fp64 = (int32_t)geti(l, "coarse_fp64", 1);
        p.graph = (int32_t)geti(l, "graph", 1);
        p.index_mode = (int32_t)geti(l, "index_mode", 0);
        sp.strength = getf(l, "mg_strength_threshold", 0.5f);
        sp.max_row_sum = getf(l, "mg_max_row_sum", 0.9f);
        sp.coarse_size = (int)geti(l, "mg_coarse_matrix_size", 100);
        sp.max_levels = (int)geti(l, "mg_max_levels", 25);
        sp.trunc_factor = getf(l, "mg_trunc_factor", 0.25f);
        sp.pmax = (int)geti(l, "mg_Pmax_elements", 4);
        sp.agg_num_levels = (int)geti(l, "mg_agg_num_levels", 10);
        sp.num_paths = std::max(1, (int)geti(l, "mg_num_paths", 3));
        sp.agg_trunc_factor = getf(l, "mg_agg_trunc_factor", 0.25f);
        sp.agg_pmax = (int)geti(l, "mg_agg_Pmax_elements", 4);
        sp.coarsen_type = (int)geti(l, "mg_coarsening_type", 6);
        sp.interp_type = (int)geti(l, "mg_interpolation_type", 0);
        sp.agg_interp_type = (int)geti(l, "mg_agg_interpolation_type", 4);
        if ((rc = check_setup_keys(sp.keys(-1), getf(l, "mg_nonGalerkin_tol", 0.f), "")))
            return rc;
        // level overrides of the setup keys take effect with a per-level hierarchy, as in
        // the reference, where get_full_hierarchy reads the list without a level
        // (hypre_hierarchy.inl:31-52) and get_per_level_hierarchy the list of each level
        // (:54-108)
        auto ov = g.overrides.find("preconditioner");
        if (geti(l, "hypre_per_level_hierarchy", 0) && ov != g.overrides.end() &&
            !ov->second.empty()) {
            int nlev = 0;
            for (auto &o : ov->second)
                nlev = std::max(nlev, std::min(o.hi, sp.max_levels - 1) + 1);
            for (int lev = 0; lev < nlev; ++lev) {
                ParamList pl = l;
                for (auto &o : ov->second)
                    if (o.lo <= lev && lev <= o.hi)
                        pl[o.key] = o.value;
                LevelKeys k;
                k.strength = getf(pl, "mg_strength_threshold", 0.5f);
                k.max_row_sum = getf(pl, "mg_max_row_sum", 0.9f);
                k.trunc_factor = getf(pl, "mg_trunc_factor", 0.25f);
                k.pmax = (int)geti(pl, "mg_Pmax_elements", 4);
                k.coarsen_type = (int)geti(pl, "mg_coarsening_type", 6);
                k.interp_type = (int)geti(pl, "mg_interpolation_type", 0);
                k.agg_interp_type = (int)geti(pl, "mg_agg_interpolation_type", 4);
                k.agg_trunc_factor = getf(pl, "mg_agg_trunc_factor", 0.25f);
                k.agg_pmax = (int)geti(pl, "mg_agg_Pmax_elements", 4);
                k.num_paths = std::max(1, (int)geti(pl, "mg_num_paths", 3));
                if ((rc = check_setup_keys(k, getf(pl, "mg_nonGalerkin_tol", 0.f),
                                           " on level " + std::to_string(lev))))
                    return rc;
                sp.per_level.push_back(k);
            }
        }
    } else {
        p.graph = 1;
        p.hsgs_block_rows = 64;
        p.hsgs_ncores = 1;
    }
    for (auto q : {std::make_pair("pre_smoother", &p.pre), std::make_pair("post_smoother", &p.post),
                   std::make_pair("coarse_grid_solver", &p.coarse)})
        if (g.roles.count(q.first) && (rc = role_of(g.roles[q.first], *q.second)))
            return rc;
    // per-level smoother roles: params.get("pre_smoother", l) (mg.inl:113-131) is the
    // role's list with the overrides whose range holds l applied in the order given, and
    // the defaults of the resulting family filled in
    int nlev = 0;
    for (const char *r : {"pre_smoother", "post_smoother"}) {
        auto ov = g.overrides.find(r);
        if (ov != g.overrides.end())
            for (auto &o : ov->second)
                nlev = std::max(nlev, std::min(o.hi, XH_MAX_LEVELS - 1) + 1);
    }
    for (int lev = 0; lev < nlev; ++lev) {
        for (auto q : {std::make_pair("pre_smoother", &p.pre_lev[lev]),
                       std::make_pair("post_smoother", &p.post_lev[lev])}) {
            ParamList pl = g.roles.count(q.first) ? g.roles[q.first] : ParamList();
            auto ov = g.overrides.find(q.first);
            if (ov != g.overrides.end())
                for (auto &o : ov->second)
                    if (o.lo <= lev && lev <= o.hi)
                        pl[o.key] = o.value;
            family_defaults(pl, q.first);
            if ((rc = role_of(pl, *q.second)))
                return rc;
        }
    }
    p.nlev_roles = nlev;
    // the GPU smoothers run without a convergence check (the reference default for
    // smoother lists, xamg_params_details.h:164-170)
    auto checked = [](const xh_role &r) { return r.method != XH_NONE && r.conv_check != 0; };
    bool bad = checked(p.pre) || checked(p.post) || checked(p.coarse);
    for (int lev = 0; lev < nlev; ++lev)
        bad = bad || checked(p.pre_lev[lev]) || checked(p.post_lev[lev]);
    if (mg && bad)
        return api_fail(-3, "convergence_check on a smoother or coarse-grid role is not "
                            "supported (smoothers run a fixed number of sweeps)");
    return 0;
}

} // namespace

extern "C" {

const char *XAMG_get_last_error() { return g_api_err.c_str(); }

int XAMG_init(int argc, char **argv, char *conf) {
    (void)argc;
    (void)argv;
    (void)conf;
    if (const char *e = std::getenv("XAMG_NV")) {
        int v = std::atoi(e);
        if (!nv_supported(v))
            return api_fail(-1, "XAMG_NV must be one of 1,2,4,8,16,32,64");
        g_nv = v;
    }
    int rc = xh_init(0);
    return rc ? api_fail(rc, xh_last_error()) : 0;
}

int XAMG_finalize() {
    g_solvers.clear();
    g_mats.clear();
    g_csrs.clear();
    g_vecs.clear();
    g_params.clear();
    g_ovs.clear();
    g_parts.clear();
    xh_finalize();
    return 0;
}

int XAMG_get_nv() { return g_nv; }
int XAMG_id_get_proc() { return 0; }
int XAMG_id_get_nprocs() { return 1; }
void *XAMG_id_get_comm_ptr() { return nullptr; }

int XAMG_param_create(XAMG_param_id *pid) {
    g_params.emplace_back(new GlobalParams());
    *pid = g_params.size() - 1;
    return 0;
}

static int param_put(char *role, char *key, const std::string &v, XAMG_param_id pid, int mode) {
    GlobalParams *g = lookup(g_params, pid);
    if (!g || !role || !key)
        return api_fail(-1, "bad parameter handle");
    ParamList &l = g->roles[role];
    // mode 0 add (keeps an existing value), 1 change (must exist), 2 forced change
    if (mode == 1 && !has(l, key))
        return api_fail(-1, std::string("change_value: key '") + key + "' not set");
    if (mode == 0 && has(l, key))
        return 0;
    l[key] = v;
    return 0;
}

int XAMG_param_add_value_s(char *r, char *k, char *v, XAMG_param_id pid) {
    return param_put(r, k, v ? v : "", pid, 0);
}
int XAMG_param_add_value_i(char *r, char *k, uint16_t v, XAMG_param_id pid) {
    return param_put(r, k, std::to_string(v), pid, 0);
}
int XAMG_param_add_value_d(char *r, char *k, float v, XAMG_param_id pid) {
    return param_put(r, k, fmt_float(v), pid, 0);
}
int XAMG_param_change_value_s(char *r, char *k, char *v, XAMG_param_id pid) {
    return param_put(r, k, v ? v : "", pid, 1);
}
int XAMG_param_change_value_i(char *r, char *k, uint16_t v, XAMG_param_id pid) {
    return param_put(r, k, std::to_string(v), pid, 1);
}
int XAMG_param_change_value_d(char *r, char *k, float v, XAMG_param_id pid) {
    return param_put(r, k, fmt_float(v), pid, 1);
}
int XAMG_param_forced_change_value_s(char *r, char *k, char *v, XAMG_param_id pid) {
    return param_put(r, k, v ? v : "", pid, 2);
}
int XAMG_param_forced_change_value_i(char *r, char *k, uint16_t v, XAMG_param_id pid) {
    return param_put(r, k, std::to_string(v), pid, 2);
}
int XAMG_param_forced_change_value_d(char *r, char *k, float v, XAMG_param_id pid) {
    return param_put(r, k, fmt_float(v), pid, 2);
}

int XAMG_param_override_create(XAMG_param_override_id *ovid) {
    g_ovs.emplace_back(new std::vector<std::pair<std::string, std::string>>());
    *ovid = g_ovs.size() - 1;
    return 0;
}

int XAMG_param_override_add_value(char *key, char *val, XAMG_param_override_id ovid) {
    auto *o = lookup(g_ovs, ovid);
    if (!o || !key || !val)
        return api_fail(-1, "bad override handle");
    o->emplace_back(key, val);
    return 0;
}

int XAMG_param_override_apply(char *role, XAMG_param_id pid, XAMG_param_override_id ovid) {
    auto *o = lookup(g_ovs, ovid);
    GlobalParams *g = lookup(g_params, pid);
    if (!o || !g || !role)
        return api_fail(-1, "bad override/parameter handle");
    // params::add_override (api/c/xamg_c_api.cpp:256-260): each value carries its level
    // range; the whole set is checked before any of it is stored
    std::vector<LevelOverride> parsed;
    for (auto &kv : *o) {
        LevelOverride lo;
        lo.key = kv.first;
        if (!parse_level_value(kv.second, lo))
            return api_fail(-1, "override '" + kv.first + "=" + kv.second +
                                    "': value must end in @levX, @levX-Y or @levX-E");
        if (!override_key_ok(role, lo.key))
            return api_fail(-3, std::string("override of '") + lo.key + "' on role '" + role +
                                    "' is not supported (level overrides apply to the "
                                    "pre_smoother / post_smoother keys and to the MultiGrid "
                                    "setup keys of the preconditioner)");
        parsed.push_back(lo);
    }
    auto &dst = g->overrides[role];
    dst.insert(dst.end(), parsed.begin(), parsed.end());
    return 0;
}

int XAMG_param_override_destroy(XAMG_param_override_id ovid) {
    if (ovid < g_ovs.size())
        g_ovs[ovid].reset();
    return 0;
}

void XAMG_param_set_default(XAMG_param_id pid) {
    if (GlobalParams *g = lookup(g_params, pid))
        set_defaults(*g);
}

int XAMG_param_print(XAMG_param_id pid) {
    GlobalParams *g = lookup(g_params, pid);
    if (!g)
        return api_fail(-1, "bad parameter handle");
    for (auto &r : g->roles) {
        std::printf("[%s]\n", r.first.c_str());
        for (auto &kv : r.second)
            std::printf("  %-28s %s\n", kv.first.c_str(), kv.second.c_str());
    }
    return 0;
}

int XAMG_param_destroy(XAMG_param_id pid) {
    if (pid < g_params.size())
        g_params[pid].reset();
    return 0;
}

int XAMG_part_create(XAMG_part_id *pid) {
    g_parts.push_back(0);
    *pid = g_parts.size() - 1;
    return 0;
}

int XAMG_part_construct(int size, XAMG_part_id pid) {
    if (pid >= g_parts.size())
        return api_fail(-1, "bad part handle");
    g_parts[pid] = size;
    return 0;
}

int XAMG_part_get_numa_size(XAMG_part_id pid) {
    return pid < g_parts.size() ? g_parts[pid] : -1;
}

int XAMG_part_destroy(XAMG_part_id) { return 0; }

int XAMG_vector_create(XAMG_vector_id *vid, vector_alloc_mode) {
    g_vecs.emplace_back(new Vector());
    *vid = g_vecs.size() - 1;
    return 0;
}

int XAMG_vector_offset(uint64_t off, XAMG_vector_id vid) {
    Vector *v = lookup(g_vecs, vid);
    if (!v)
        return api_fail(-1, "bad vector handle");
    v->offset = off;
    return 0;
}

int XAMG_vector_alloc_d(int size, uint16_t nv, XAMG_vector_id vid) {
    Vector *v = lookup(g_vecs, vid);
    if (!v || size < 0 || !nv_supported(nv))
        return api_fail(-1, "XAMG_vector_alloc_d: bad arguments (nv must be 1,2,...,64)");
    v->size = size;
    v->nv = nv;
    v->d.assign((size_t)size * nv, 0.0);
    return 0;
}

int XAMG_vector_set_part(XAMG_part_id, XAMG_vector_id) { return 0; }

int XAMG_vector_set_val_d(double val, XAMG_vector_id vid) {
    Vector *v = lookup(g_vecs, vid);
    if (!v)
        return api_fail(-1, "bad vector handle");
    std::fill(v->d.begin(), v->d.end(), val);
    return 0;
}

int XAMG_vector_upload_d(double *ptr, uint64_t block_size, uint64_t block_offset, uint16_t nv,
                         XAMG_vector_id vid) {
    Vector *v = lookup(g_vecs, vid);
    if (!v || !ptr || nv != v->nv || block_offset + block_size > v->size)
        return api_fail(-1, "XAMG_vector_upload_d: bad arguments");
    std::memcpy(v->d.data() + block_offset * nv, ptr, block_size * nv * sizeof(double));
    return 0;
}

int XAMG_vector_download_d(double *ptr, uint64_t block_size, uint64_t block_offset, uint16_t nv,
                           XAMG_vector_id vid) {
    Vector *v = lookup(g_vecs, vid);
    if (!v || !ptr || nv != v->nv || block_offset + block_size > v->size)
        return api_fail(-1, "XAMG_vector_download_d: bad arguments");
    std::memcpy(ptr, v->d.data() + block_offset * nv, block_size * nv * sizeof(double));
    return 0;
}

int XAMG_vector_upload_v(XAMG_vector_id in, XAMG_vector_id out) {
    Vector *a = lookup(g_vecs, in), *b = lookup(g_vecs, out);
    if (!a || !b || a->nv != b->nv || a->size + a->offset > b->size)
        return api_fail(-1, "XAMG_vector_upload_v: bad arguments");
    std::copy(a->d.begin(), a->d.end(), b->d.begin() + a->offset * a->nv);
    return 0;
}

int XAMG_vector_destroy(XAMG_vector_id vid) {
    if (vid < g_vecs.size())
        g_vecs[vid].reset();
    return 0;
}

int XAMG_csr_matrix_create(XAMG_csr_matrix_id *id) {
    g_csrs.emplace_back(new Csr());
    *id = g_csrs.size() - 1;
    return 0;
}

int XAMG_csr_matrix_alloc(uint64_t nrows, uint64_t nnz, XAMG_csr_matrix_id id) {
    Csr *c = lookup(g_csrs, id);
    if (!c)
        return api_fail(-1, "bad csr handle");
    c->nrows = (int64_t)nrows;
    c->nnz = (int64_t)nnz;
    c->row.assign(nrows + 1, 0);
    c->col.assign(nnz, 0);
    c->val.assign(nnz, 0.0);
    return 0;
}

int XAMG_csr_matrix_fill(int *row, int *col, double *val, XAMG_matrix_id id) {
    Csr *c = lookup(g_csrs, id);
    if (!c || !row || (c->nnz && (!col || !val)))
        return api_fail(-1, "XAMG_csr_matrix_fill: bad arguments");
    for (int64_t i = 0; i <= c->nrows; ++i)
        c->row[i] = (uint32_t)row[i];
    for (int64_t j = 0; j < c->nnz; ++j) {
        c->col[j] = (uint32_t)col[j];
        c->val[j] = val[j];
    }
    return 0;
}

int XAMG_csr_matrix_offset(int ro, int co, XAMG_matrix_id id) {
    Csr *c = lookup(g_csrs, id);
    if (!c)
        return api_fail(-1, "bad csr handle");
    c->row_off = ro;
    c->col_off = co;
    return 0;
}

uint32_t XAMG_csr_matrix_get_nrows(XAMG_csr_matrix_id id) {
    Csr *c = lookup(g_csrs, id);
    return c ? (uint32_t)c->nrows : 0;
}

int XAMG_csr_matrix_destroy(XAMG_csr_matrix_id id) {
    if (id < g_csrs.size())
        g_csrs[id].reset();
    return 0;
}

int XAMG_matrix_create(XAMG_matrix_id *mid) {
    g_mats.emplace_back(new Matrix());
    *mid = g_mats.size() - 1;
    return 0;
}

int XAMG_matrix_set_part(XAMG_part_id, XAMG_matrix_id) { return 0; }

int XAMG_matrix_construct(XAMG_csr_matrix_id cid, XAMG_matrix_id mid) {
    Csr *c = lookup(g_csrs, cid);
    Matrix *m = lookup(g_mats, mid);
    if (!c || !m)
        return api_fail(-1, "XAMG_matrix_construct: bad handles");
    if (c->row_off != 0 || c->col_off != 0)
        return api_fail(-3, "XAMG_matrix_construct: one process holds the whole matrix");
    m->csr = std::make_shared<Csr>(*c);
    return 0;
}

int XAMG_matrix_destroy(XAMG_matrix_id mid) {
    if (mid < g_mats.size())
        g_mats[mid].reset();
    return 0;
}

static int solver_create(XAMG_matrix_id mid, XAMG_param_id pid, int nv, size_t xid, size_t bid,
                         XAMG_solver_id *sid) {
    Matrix *m = lookup(g_mats, mid);
    GlobalParams *g = lookup(g_params, pid);
    if (!m || !m->csr || !g || !sid)
        return api_fail(-1, "XAMG_solver_create: bad handles");
    if (!nv_supported(nv))
        return api_fail(-1, "XAMG_solver_create: unsupported nv");
    xh_params p;
    SetupParams sp;
    bool mg = false;
    int rc = build_xh_params(*g, p, sp, mg);
    if (rc)
        return rc;
    const Csr &c = *m->csr;
    CsrHost A;
    A.nrows = A.ncols = c.nrows;
    A.row = c.row;
    A.col = c.col;
    A.val = c.val;
    Hierarchy H;
    if (mg) {
        xamg::amg::build(A, sp, H);
    } else {
        xamg::amg::sort_rows_diag_first(A);
        H.A.push_back(A);
        H.min_eig.assign(1, 0.0);
        H.max_eig.assign(1, 0.0);
        xamg::amg::estimate_eigs(H.A[0], sp.eig_iters, sp.eig_seed, H.min_eig[0], H.max_eig[0]);
    }
    auto s = std::make_unique<Solver>();
    s->nv = nv;
    s->xid = xid;
    s->bid = bid;
    rc = xh_hier_create(H.nlevels(), &s->hier);
    for (int l = 0; !rc && l < H.nlevels(); ++l) {
        const bool last = l + 1 == H.nlevels();
        const CsrHost &Al = H.A[l];
        rc = xh_hier_set_level(
            s->hier, l, Al.nrows, Al.nnz(), Al.row.data(), Al.col.data(), Al.val.data(),
            last ? 0 : H.P[l].ncols, last ? 0 : H.P[l].nnz(), last ? nullptr : H.P[l].row.data(),
            last ? nullptr : H.P[l].col.data(), last ? nullptr : H.P[l].val.data(),
            last ? 0 : H.R[l].nnz(), last ? nullptr : H.R[l].row.data(),
            last ? nullptr : H.R[l].col.data(), last ? nullptr : H.R[l].val.data(), H.min_eig[l],
            H.max_eig[l]);
    }
    if (!rc)
        rc = xh_solver_create(s->hier, &p, nv, &s->gpu);
    if (rc)
        return api_fail(rc, xh_last_error());
    s->conv.assign(nv, 0);
    s->abs_res.assign(nv, 0.0);
    s->rel_res.assign(nv, 0.0);
    g_solvers.push_back(std::move(s));
    *sid = g_solvers.size() - 1;
    return 0;
}

int xamg_param_resolve(size_t pid, xh_params *prm, xamg_amg_params *setup,
                       xamg_amg_level_keys *lev_keys) {
    GlobalParams *g = lookup(g_params, pid);
    if (!g || !prm)
        return api_fail(-1, "xamg_param_resolve: bad arguments");
    SetupParams sp;
    bool mg = false;
    int rc = build_xh_params(*g, *prm, sp, mg);
    if (rc)
        return rc;
    if (setup) {
        setup->strength = sp.strength;
        setup->max_row_sum = sp.max_row_sum;
        setup->coarse_size = sp.coarse_size;
        setup->max_levels = sp.max_levels;
        setup->trunc_factor = sp.trunc_factor;
        setup->pmax = sp.pmax;
        setup->second_pass = sp.second_pass;
        setup->agg_num_levels = sp.agg_num_levels;
        setup->num_paths = sp.num_paths;
        setup->agg_trunc_factor = sp.agg_trunc_factor;
        setup->agg_pmax = sp.agg_pmax;
        setup->eig_iters = sp.eig_iters;
        setup->eig_seed = sp.eig_seed;
        setup->coarsen_type = sp.coarsen_type;
        setup->interp_type = sp.interp_type;
        setup->agg_interp_type = sp.agg_interp_type;
        setup->nlev_keys = std::min((int)sp.per_level.size(), XH_MAX_LEVELS);
        setup->lev_keys = nullptr;
        for (int l = 0; lev_keys && l < setup->nlev_keys; ++l) {
            const LevelKeys &k = sp.per_level[l];
            lev_keys[l] = {k.strength,     k.max_row_sum,     k.trunc_factor, k.pmax,
                           k.coarsen_type, k.interp_type,     k.agg_interp_type,
                           k.agg_trunc_factor, k.agg_pmax, k.num_paths};
        }
    }
    return 0;
}

int XAMG_solver_create1(XAMG_matrix_id mid, XAMG_param_id pid, XAMG_solver_id *sid) {
    return solver_create(mid, pid, g_nv, (size_t)-1, (size_t)-1, sid);
}

int XAMG_solver_create2(XAMG_matrix_id mid, XAMG_param_id pid, XAMG_vector_id xid,
                        XAMG_vector_id bid, XAMG_solver_id *sid) {
    Vector *x = lookup(g_vecs, xid), *b = lookup(g_vecs, bid);
    if (!x || !b || x->nv != b->nv)
        return api_fail(-1, "XAMG_solver_create2: bad vectors");
    return solver_create(mid, pid, x->nv, xid, bid, sid);
}

int XAMG_solver_renew_params(XAMG_param_id pid, XAMG_solver_id sid) {
    // the reference example passes (solver, params) -- accept either order
    Solver *s = lookup(g_solvers, sid);
    GlobalParams *g = lookup(g_params, pid);
    if ((!s || !g) && lookup(g_solvers, pid) && lookup(g_params, sid)) {
        s = lookup(g_solvers, pid);
        g = lookup(g_params, sid);
    }
    if (!s || !g || !g->roles.count("solver"))
        return api_fail(-1, "XAMG_solver_renew_params: bad handles");
    xh_role r;
    int rc = role_of(g->roles["solver"], r);
    if (rc)
        return rc;
    rc = xh_solver_renew(s->gpu, &r);
    return rc ? api_fail(rc, xh_last_error()) : 0;
}

static int do_solve(Solver *s, Vector *x, Vector *b) {
    if (!x || !b || x->nv != s->nv || b->nv != s->nv || x->size != b->size)
        return api_fail(-1, "solve: vectors do not match the solver");
    int rc = xh_solver_solve(s->gpu, b->d.data(), x->d.data());
    if (rc)
        return api_fail(rc, xh_last_error());
    xh_solver_stats(s->gpu, &s->iters, s->conv.data(), s->abs_res.data(), s->rel_res.data());
    return 0;
}

int XAMG_solver_solve1(uint16_t nv, XAMG_solver_id sid) {
    Solver *s = lookup(g_solvers, sid);
    if (!s || nv != s->nv)
        return api_fail(-1, "XAMG_solver_solve1: bad solver or nv");
    return do_solve(s, lookup(g_vecs, s->xid), lookup(g_vecs, s->bid));
}

int XAMG_solver_solve2(uint16_t nv, XAMG_vector_id xid, XAMG_vector_id bid, XAMG_solver_id sid) {
    Solver *s = lookup(g_solvers, sid);
    if (!s || nv != s->nv)
        return api_fail(-1, "XAMG_solver_solve2: bad solver or nv");
    return do_solve(s, lookup(g_vecs, xid), lookup(g_vecs, bid));
}

int XAMG_solver_get_convergence_info(int *conv, uint16_t nv, XAMG_solver_id sid) {
    Solver *s = lookup(g_solvers, sid);
    if (!s || !conv || nv != s->nv)
        return api_fail(-1, "XAMG_solver_get_convergence_info: bad arguments");
    for (int k = 0; k < nv; ++k)
        conv[k] = s->conv[k];
    return 0;
}

int XAMG_solver_get_stats(XAMG_solver_id sid, int *iters, double *abs_res, double *rel_res) {
    Solver *s = lookup(g_solvers, sid);
    if (!s)
        return api_fail(-1, "bad solver handle");
    if (iters)
        *iters = s->iters;
    for (int k = 0; k < s->nv; ++k) {
        if (abs_res)
            abs_res[k] = s->abs_res[k];
        if (rel_res)
            rel_res[k] = s->rel_res[k];
    }
    return 0;
}

int XAMG_solver_destroy(XAMG_solver_id sid) {
    if (sid < g_solvers.size())
        g_solvers[sid].reset();
    return 0;
}

double XAMG_timer() {
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    return tv.tv_sec + 1e-6 * tv.tv_usec;
}

int XAMG_barrier() {
    int rc = xh_synchronize();
    return rc ? api_fail(rc, xh_last_error()) : 0;
}

} // extern "C"
