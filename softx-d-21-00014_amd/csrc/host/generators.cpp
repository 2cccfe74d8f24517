// Test-matrix and right-hand-side generators of the benchmark configurations, host side.
//
//  * xamg_varcoef7: -div(k grad u) on an nx x ny x nz cell grid of the unit cube,
//    cell-centred finite volumes, 7-point stencil, Dirichlet ghost cells.  k is one value
//    per cell, lognormal: k = exp(sigma * z), z standard normal from the counter-based
//    generator below.  The coefficient of an interior face is q_d * 2 k_P k_Q / (k_P + k_Q)
//    (harmonic mean, q_d = 1/h_d^2); a boundary face adds 2 q_d k_P to the diagonal.  Rows
//    are laid out like the reference's cube generator (examples/common/generator/cube.inl:
//    84-146): diagonal first, then -z, -y, -x, +x, +y, +z, the diagonal accumulated in that
//    order, so sigma = 0 reproduces that generator bit for bit.
//  * xamg_q1fem27: the same operator discretised with trilinear (Q1) finite elements on
//    the nodes of an (nx+1) x (ny+1) x (nz+1) cell grid, Dirichlet nodes eliminated
//    (nx x ny x nz unknowns), k one value per cell.  Each row stores all 27 neighbours
//    that exist (the face couplings of Q1 on a cube are exactly zero and are kept as
//    stored zeros), diagonal first, then ascending columns.
//  * xamg_rhs_bumps: smooth random right-hand sides, each a sum of Gaussian bumps (cut
//    off below exp(-8) of their peak).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../../include/xamg_amg.h"

namespace {

uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [0, 1): 53 random bits of stream `seed`, draw `i`
double unif(uint64_t seed, uint64_t i) {
    return (double)(splitmix64(seed * 0xD1B54A32D192ED03ull + i) >> 11) * 0x1.0p-53;
}

// standard normal, draw i of stream seed (Box-Muller on two uniforms)
double normal(uint64_t seed, uint64_t i) {
    const double u1 = 1.0 - unif(seed, 2 * i);   // (0, 1]
    const double u2 = unif(seed, 2 * i + 1);
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

} // namespace

extern "C" {

int xamg_lognormal_field(uint64_t seed, double sigma, int64_t count, double *k) {
    if (count < 0 || (count > 0 && !k))
        return -1;
    for (int64_t i = 0; i < count; ++i)
        k[i] = sigma == 0.0 ? 1.0 : std::exp(sigma * normal(seed, (uint64_t)i));
    return 0;
}

int64_t xamg_varcoef7_nnz(int64_t nx, int64_t ny, int64_t nz) {
    return xamg_cube_nnz(nx, ny, nz);
}

int xamg_varcoef7(int64_t nx, int64_t ny, int64_t nz, const double *kcell, uint32_t *row,
                  uint32_t *col, double *val) {
    if (nx < 1 || ny < 1 || nz < 1 || !kcell || !row || !col || !val)
        return -1;
    const double hx = 1.0 / nx, hy = 1.0 / ny, hz = 1.0 / nz;
    const double q1 = 1.0 / hx / hx, q2 = 1.0 / hy / hy, q3 = 1.0 / hz / hz;
    const int64_t n = nx * ny * nz;
    uint64_t c = 0;
    row[0] = 0;
    for (int64_t p = 0; p < n; ++p) {
        const int64_t ii = p % nx, jj = (p / nx) % ny, kk = p / (nx * ny);
        const double kp = kcell[p];
        const uint64_t d = c;
        col[c] = (uint32_t)p;
        val[c++] = 0.0;
        auto face = [&](bool inside, int64_t q, double qd, bool dim_gt1) {
            if (inside) {
                const double kq = kcell[q];
                const double f = qd * (2.0 * kp * kq / (kp + kq));
                val[d] += f;
                col[c] = (uint32_t)q;
                val[c++] = -f;
            } else if (dim_gt1) {
                val[d] += 2.0 * qd * kp;
            }
        };
        face(kk > 0, p - nx * ny, q3, nz > 1);
        face(jj > 0, p - nx, q2, ny > 1);
        face(ii > 0, p - 1, q1, nx > 1);
        face(ii < nx - 1, p + 1, q1, nx > 1);
        face(jj < ny - 1, p + nx, q2, ny > 1);
        face(kk < nz - 1, p + nx * ny, q3, nz > 1);
        row[p + 1] = (uint32_t)c;
    }
    return 0;
}

int64_t xamg_q1fem27_nnz(int64_t nx, int64_t ny, int64_t nz) {
    return (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
}

int xamg_q1fem27(int64_t nx, int64_t ny, int64_t nz, const double *kcell, uint32_t *row,
                 uint32_t *col, double *val) {
    // unknowns: interior nodes (i, j, l), 1 <= i <= nx (grid of nx+2 nodes per axis);
    // cells: (nx+1)^3, cell (a, b, c) spans nodes a..a+1; kcell indexed a + (nx+1)(b + (ny+1)c)
    if (nx < 1 || ny < 1 || nz < 1 || !kcell || !row || !col || !val)
        return -1;
    const double h = 1.0 / (double)(nx + 1);
    // Q1 stiffness on a cube of edge h: h/3 (same node), 0 (nodes differing in one
    // coordinate), -h/12 (two or three coordinates)
    auto kref = [&](int di, int dj, int dl) {
        const int m = (di != 0) + (dj != 0) + (dl != 0);
        return m == 0 ? h / 3.0 : (m == 1 ? 0.0 : -h / 12.0);
    };
    const int64_t cx = nx + 1, cy = ny + 1;
    uint64_t c = 0;
    row[0] = 0;
    for (int64_t l = 1; l <= nz; ++l)
        for (int64_t j = 1; j <= ny; ++j)
            for (int64_t i = 1; i <= nx; ++i) {
                const int64_t p = (i - 1) + nx * ((j - 1) + ny * (l - 1));
                // coefficient to neighbour (i+di, j+dj, l+dl): sum over the cells holding
                // both nodes
                double a[3][3][3];
                for (int dl = -1; dl <= 1; ++dl)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int di = -1; di <= 1; ++di) {
                            double s = 0.0;
                            for (int el = -1; el <= 0; ++el)
                                for (int ej = -1; ej <= 0; ++ej)
                                    for (int ei = -1; ei <= 0; ++ei) {
                                        // cell with lower corner (i+ei, j+ej, l+el) holds node
                                        // p; it holds the neighbour if the offset stays in it
                                        const int oi = di - ei, oj = dj - ej, ol = dl - el;
                                        if (oi < 0 || oi > 1 || oj < 0 || oj > 1 || ol < 0 ||
                                            ol > 1)
                                            continue;
                                        const int64_t ca = i + ei, cb = j + ej, cc = l + el;
                                        s += kcell[ca + cx * (cb + cy * cc)] * kref(di, dj, dl);
                                    }
                            a[dl + 1][dj + 1][di + 1] = s;
                        }
                col[c] = (uint32_t)p;
                val[c++] = a[1][1][1];
                for (int dl = -1; dl <= 1; ++dl)
                    for (int dj = -1; dj <= 1; ++dj)
                        for (int di = -1; di <= 1; ++di) {
                            if (!di && !dj && !dl)
                                continue;
                            const int64_t qi = i + di, qj = j + dj, ql = l + dl;
                            if (qi < 1 || qi > nx || qj < 1 || qj > ny || ql < 1 || ql > nz)
                                continue; // Dirichlet node, eliminated
                            col[c] = (uint32_t)((qi - 1) + nx * ((qj - 1) + ny * (ql - 1)));
                            val[c++] = a[dl + 1][dj + 1][di + 1];
                        }
                row[p + 1] = (uint32_t)c;
            }
    return 0;
}

int xamg_bump_params(uint64_t seed, int64_t nx, int64_t ny, int64_t nz, int nv, int min_bumps,
                     int max_bumps, double min_width, double max_width, int32_t *count,
                     double *params) {
    // RHS k: count[k] bumps, bump q described by params[(k*max_bumps + q)*5 + {cx, cy, cz,
    // w, a}] (centre in cell units, width, amplitude)
    if (nx < 1 || ny < 1 || nz < 1 || nv < 1 || min_bumps < 1 || max_bumps < min_bumps ||
        min_width <= 0 || max_width < min_width || !count || !params)
        return -1;
    for (int k = 0; k < nv; ++k) {
        const uint64_t s = seed * 1000003ull + (uint64_t)k;
        const int m = min_bumps + (int)(unif(s, 0) * (max_bumps - min_bumps + 1));
        count[k] = m;
        for (int q = 0; q < max_bumps; ++q) {
            double *P = params + ((size_t)k * max_bumps + q) * 5;
            const uint64_t o = 8 + 8 * (uint64_t)q;
            P[0] = unif(s, o) * nx;
            P[1] = unif(s, o + 1) * ny;
            P[2] = unif(s, o + 2) * nz;
            P[3] = min_width + unif(s, o + 3) * (max_width - min_width);
            P[4] = normal(s, o + 4);
        }
    }
    return 0;
}

int xamg_rhs_bumps(uint64_t seed, int64_t nx, int64_t ny, int64_t nz, int nv, int min_bumps,
                   int max_bumps, double min_width, double max_width, double *b) {
    // sum over the bumps of xamg_bump_params; a bump is cut off outside the box of
    // half-width ceil(4w) cells around its centre (it has fallen below exp(-8) there)
    if (!b)
        return -1;
    std::vector<int32_t> count(nv > 0 ? nv : 1);
    std::vector<double> P((size_t)(nv > 0 ? nv : 1) * (max_bumps > 0 ? max_bumps : 1) * 5);
    if (xamg_bump_params(seed, nx, ny, nz, nv, min_bumps, max_bumps, min_width, max_width,
                         count.data(), P.data()))
        return -1;
    std::memset(b, 0, sizeof(double) * nx * ny * nz * nv);
    for (int k = 0; k < nv; ++k)
        for (int q = 0; q < count[k]; ++q) {
            const double *Q = &P[((size_t)k * max_bumps + q) * 5];
            const double ccx = Q[0], ccy = Q[1], ccz = Q[2], w = Q[3], a = Q[4];
            const double inv = 1.0 / (2.0 * w * w);
            const int64_t R = (int64_t)std::ceil(4.0 * w);
            auto lo = [](double c0, int64_t r) { return std::max<int64_t>(0, (int64_t)c0 - r); };
            auto hi = [](double c0, int64_t r, int64_t m) {
                return std::min<int64_t>(m - 1, (int64_t)c0 + r);
            };
            for (int64_t z = lo(ccz, R); z <= hi(ccz, R, nz); ++z)
                for (int64_t y = lo(ccy, R); y <= hi(ccy, R, ny); ++y)
                    for (int64_t x = lo(ccx, R); x <= hi(ccx, R, nx); ++x) {
                        const double dx = (double)x + 0.5 - ccx;
                        const double dy = (double)y + 0.5 - ccy;
                        const double dz = (double)z + 0.5 - ccz;
                        const int64_t p = x + nx * (y + ny * z);
                        b[p * nv + k] += a * std::exp(-(dx * dx + dy * dy + dz * dz) * inv);
                    }
        }
    return 0;
}

} // extern "C"
