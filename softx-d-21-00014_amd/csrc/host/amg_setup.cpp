// Classical (Ruge-Stueben) AMG setup on the host.
//
// The reference takes its hierarchy from hypre 2.20.0 BoomerAMG (not in the reference
// tree and not available offline here).  This file restates the published classical
// algorithm with the reference's default keys (src/param/xamg_params_details.h:179-209):
//   strength      -a_ij >= theta * max_k(-a_ik) (sign relative to the diagonal),
//                 rows whose |row sum| > max_row_sum * |a_ii| have no strong links
//   coarsening    mg_coarsening_type 6 (Falgout; on one process its Ruge-Stueben
//                 part): first pass (measure = number of points that strongly depend on
//                 a point), optional second pass; 8 (PMIS, De Sterck, Yang and Heys, SIAM
//                 J. Matrix Anal. Appl. 27 (2006)): measure + a random tie-break in (0,1),
//                 a point whose measure beats every undecided strong neighbour's becomes
//                 C, the undecided points that strongly depend on a new C point become F,
//                 until every point is decided.  On the first mg_agg_num_levels levels
//                 aggressive coarsening: a second pass of the same scheme over the C
//                 points with the long-range strength graph (at least mg_num_paths paths
//                 of length <= 2)
//   interpolation mg_interpolation_type 0: modified classical; 6: extended+i (De Sterck,
//                 Falgout, Nolting and Yang, Numer. Linear Algebra Appl. 15 (2008));
//                 mg_agg_interpolation_type 4: multipass on aggressive levels
//   truncation    drop |w| < trunc_factor * max|w|, keep at most Pmax entries,
//                 rescale each row to its original sum
//   Galerkin      A_c = R A P with R = P^T (hypre_base.inl:400-401)
// Rows of every output matrix are stored diagonal first, then ascending column.
#include "amg_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

namespace xamg {
namespace amg {

void sort_rows_diag_first(CsrHost &m) {
#pragma omp parallel for schedule(dynamic, 1024)
    for (int64_t i = 0; i < m.nrows; ++i) {
        uint32_t b = m.row[i], e = m.row[i + 1];
        std::vector<std::pair<uint32_t, double>> r;
        r.reserve(e - b);
        for (uint32_t j = b; j < e; ++j)
            r.emplace_back(m.col[j], m.val[j]);
        std::sort(r.begin(), r.end(), [i](const std::pair<uint32_t, double> &x,
                                          const std::pair<uint32_t, double> &y) {
            bool dx = x.first == (uint32_t)i, dy = y.first == (uint32_t)i;
            if (dx != dy)
                return dx;
            return x.first < y.first;
        });
        for (uint32_t j = b; j < e; ++j) {
            m.col[j] = r[j - b].first;
            m.val[j] = r[j - b].second;
        }
    }
}

CsrHost transpose(const CsrHost &m) {
    // counting-sort transpose: entries of a row of the result appear in ascending
    // source-row order (src/primitives/matrix/operations/transpose.inl:31-94)
    CsrHost t;
    t.nrows = m.ncols;
    t.ncols = m.nrows;
    t.row.assign(t.nrows + 1, 0);
    for (uint32_t c : m.col)
        t.row[c + 1]++;
    for (int64_t i = 0; i < t.nrows; ++i)
        t.row[i + 1] += t.row[i];
    t.col.resize(m.col.size());
    t.val.resize(m.col.size());
    std::vector<uint32_t> pos(t.row.begin(), t.row.end() - 1);
    for (int64_t i = 0; i < m.nrows; ++i)
        for (uint32_t j = m.row[i]; j < m.row[i + 1]; ++j) {
            uint32_t c = m.col[j];
            t.col[pos[c]] = (uint32_t)i;
            t.val[pos[c]] = m.val[j];
            pos[c]++;
        }
    return t;
}

CsrHost spgemm(const CsrHost &a, const CsrHost &b) {
    CsrHost c;
    c.nrows = a.nrows;
    c.ncols = b.ncols;
    c.row.assign(a.nrows + 1, 0);
    const int64_t n = a.nrows;
    std::vector<std::vector<uint32_t>> rc(n);
    std::vector<std::vector<double>> rv(n);
#pragma omp parallel
    {
        std::vector<int64_t> mark(b.ncols, -1);
        std::vector<double> acc(b.ncols, 0.0);
        std::vector<uint32_t> cols;
#pragma omp for schedule(dynamic, 256)
        for (int64_t i = 0; i < n; ++i) {
            cols.clear();
            for (uint32_t ja = a.row[i]; ja < a.row[i + 1]; ++ja) {
                uint32_t k = a.col[ja];
                double av = a.val[ja];
                for (uint32_t jb = b.row[k]; jb < b.row[k + 1]; ++jb) {
                    uint32_t cc = b.col[jb];
                    if (mark[cc] != i) {
                        mark[cc] = i;
                        acc[cc] = av * b.val[jb];
                        cols.push_back(cc);
                    } else {
                        acc[cc] += av * b.val[jb];
                    }
                }
            }
            std::sort(cols.begin(), cols.end());
            rc[i] = cols;
            rv[i].resize(cols.size());
            for (size_t q = 0; q < cols.size(); ++q)
                rv[i][q] = acc[cols[q]];
        }
    }
    for (int64_t i = 0; i < n; ++i)
        c.row[i + 1] = c.row[i] + (uint32_t)rc[i].size();
    c.col.resize(c.row[n]);
    c.val.resize(c.row[n]);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        std::copy(rc[i].begin(), rc[i].end(), c.col.begin() + c.row[i]);
        std::copy(rv[i].begin(), rv[i].end(), c.val.begin() + c.row[i]);
    }
    return c;
}

namespace {

double diag_of(const CsrHost &A, int64_t i) {
    for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j)
        if (A.col[j] == (uint32_t)i)
            return A.val[j];
    return 0.0;
}

// strength pattern S: for every row, a flag per nonzero (1 = strong dependency)
std::vector<uint8_t> strength(const CsrHost &A, float theta, float max_row_sum) {
    std::vector<uint8_t> s(A.col.size(), 0);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < A.nrows; ++i) {
        double d = diag_of(A, i);
        double sgn = d < 0 ? -1.0 : 1.0;
        double mx = 0.0, rs = 0.0;
        for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j) {
            rs += A.val[j];
            if (A.col[j] == (uint32_t)i)
                continue;
            mx = std::max(mx, -sgn * A.val[j]);
        }
        if (mx <= 0.0)
            continue;
        if (max_row_sum < 1.0f && std::fabs(rs) > (double)max_row_sum * std::fabs(d))
            continue;
        for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j) {
            if (A.col[j] == (uint32_t)i)
                continue;
            if (-sgn * A.val[j] >= (double)theta * mx)
                s[j] = 1;
        }
    }
    return s;
}

enum : int8_t { UNDECIDED = 0, CPT = 1, FPT = -1 };

// Ruge-Stueben first pass with bucket lists keyed by the measure
std::vector<int8_t> rs_coarsen(const CsrHost &A, const std::vector<uint8_t> &S, int second_pass,
                               bool isolated_as_c = false) {
    const int64_t n = A.nrows;
    // S^T adjacency: who strongly depends on i
    std::vector<uint32_t> tr(n + 1, 0);
    for (int64_t i = 0; i < n; ++i)
        for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j)
            if (S[j])
                tr[A.col[j] + 1]++;
    for (int64_t i = 0; i < n; ++i)
        tr[i + 1] += tr[i];
    std::vector<uint32_t> tc(tr[n]);
    {
        std::vector<uint32_t> pos(tr.begin(), tr.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j)
                if (S[j])
                    tc[pos[A.col[j]]++] = (uint32_t)i;
    }
    std::vector<int8_t> cf(n, UNDECIDED);
    std::vector<int64_t> lam(n);
    int64_t maxlam = 0;
    for (int64_t i = 0; i < n; ++i) {
        lam[i] = tr[i + 1] - tr[i];
        bool has_dep = false;
        for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j)
            has_dep |= S[j] != 0;
        if (!has_dep && lam[i] == 0)
            cf[i] = isolated_as_c ? CPT : FPT; // isolated point
        maxlam = std::max(maxlam, lam[i]);
    }
    const int64_t nb = 2 * maxlam + n + 2; // measures can grow while F points are made
    std::vector<int64_t> head(nb, -1), nxt(n, -1), prv(n, -1);
    auto insert = [&](int64_t i) {
        int64_t b = lam[i];
        nxt[i] = head[b];
        prv[i] = -1;
        if (head[b] >= 0)
            prv[head[b]] = i;
        head[b] = i;
    };
    auto remove = [&](int64_t i) {
        int64_t b = lam[i];
        if (prv[i] >= 0)
            nxt[prv[i]] = nxt[i];
        else
            head[b] = nxt[i];
        if (nxt[i] >= 0)
            prv[nxt[i]] = prv[i];
        nxt[i] = prv[i] = -1;
    };
    for (int64_t i = 0; i < n; ++i)
        if (cf[i] == UNDECIDED)
            insert(i);
    int64_t top = nb - 1;
    while (true) {
        while (top >= 0 && head[top] < 0)
            --top;
        if (top < 0)
            break;
        int64_t i = head[top];
        if (lam[i] == 0) {
            // nobody depends on the remaining points: they become F points
            for (int64_t k = 0; k < n; ++k)
                if (cf[k] == UNDECIDED)
                    cf[k] = FPT;
            break;
        }
        remove(i);
        cf[i] = CPT;
        for (uint32_t q = tr[i]; q < tr[i + 1]; ++q) {
            int64_t j = tc[q];
            if (cf[j] != UNDECIDED)
                continue;
            remove(j);
            cf[j] = FPT;
            for (uint32_t jj = A.row[j]; jj < A.row[j + 1]; ++jj) {
                if (!S[jj])
                    continue;
                int64_t k = A.col[jj];
                if (cf[k] != UNDECIDED)
                    continue;
                remove(k);
                lam[k]++;
                if (lam[k] >= nb)
                    lam[k] = nb - 1;
                insert(k);
                top = std::max(top, lam[k]);
            }
        }
        for (uint32_t jj = A.row[i]; jj < A.row[i + 1]; ++jj) {
            if (!S[jj])
                continue;
            int64_t k = A.col[jj];
            if (cf[k] != UNDECIDED)
                continue;
            remove(k);
            if (lam[k] > 0)
                lam[k]--;
            insert(k);
        }
    }
    if (second_pass) {
        // F points i, j with a strong F-F link must share a strong C point
        std::vector<int64_t> mark(n, -1);
        for (int64_t i = 0; i < n; ++i) {
            if (cf[i] != FPT)
                continue;
            for (uint32_t jj = A.row[i]; jj < A.row[i + 1]; ++jj)
                if (S[jj] && cf[A.col[jj]] == CPT)
                    mark[A.col[jj]] = i;
            for (uint32_t jj = A.row[i]; jj < A.row[i + 1]; ++jj) {
                if (!S[jj])
                    continue;
                int64_t k = A.col[jj];
                if (cf[k] != FPT)
                    continue;
                bool common = false;
                for (uint32_t kk = A.row[k]; kk < A.row[k + 1]; ++kk)
                    if (S[kk] && cf[A.col[kk]] == CPT && mark[A.col[kk]] == i) {
                        common = true;
                        break;
                    }
                if (!common) {
                    cf[k] = CPT;
                    mark[k] = i;
                }
            }
        }
    }
    return cf;
}

// modified classical interpolation + truncation
CsrHost interpolation(const CsrHost &A, const std::vector<uint8_t> &S,
                      const std::vector<int8_t> &cf, float trunc_factor, int pmax,
                      int64_t &nc) {
    const int64_t n = A.nrows;
    std::vector<int64_t> cidx(n, -1);
    nc = 0;
    for (int64_t i = 0; i < n; ++i)
        if (cf[i] == CPT)
            cidx[i] = nc++;
    std::vector<std::vector<uint32_t>> rc(n);
    std::vector<std::vector<double>> rv(n);
#pragma omp parallel
    {
        std::vector<int64_t> mark(n, -1);
        std::vector<double> w(n, 0.0);
#pragma omp for schedule(dynamic, 512)
        for (int64_t i = 0; i < n; ++i) {
            if (cf[i] == CPT) {
                rc[i] = {(uint32_t)cidx[i]};
                rv[i] = {1.0};
                continue;
            }
            std::vector<uint32_t> Ci;
            double diag = 0.0;
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j) {
                uint32_t c = A.col[j];
                if (c == (uint32_t)i)
                    diag += A.val[j];
                else if (S[j] && cf[c] == CPT) {
                    Ci.push_back(c);
                    mark[c] = i;
                    w[c] = A.val[j];
                }
            }
            if (Ci.empty())
                continue; // F point with no strong C neighbour: empty row of P
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j) {
                uint32_t k = A.col[j];
                if (k == (uint32_t)i || (S[j] && cf[k] == CPT))
                    continue;
                if (!S[j]) {
                    diag += A.val[j]; // weak connection: lumped into the diagonal
                    continue;
                }
                // strong F neighbour k: distribute a_ik over C_i with weights abar_km
                double akk = diag_of(A, k);
                double den = 0.0;
                for (uint32_t q = A.row[k]; q < A.row[k + 1]; ++q) {
                    uint32_t m = A.col[q];
                    if (m < (uint32_t)n && mark[m] == i && A.val[q] * akk < 0)
                        den += A.val[q];
                }
                if (den == 0.0) {
                    diag += A.val[j];
                    continue;
                }
                for (uint32_t q = A.row[k]; q < A.row[k + 1]; ++q) {
                    uint32_t m = A.col[q];
                    if (mark[m] == i && A.val[q] * akk < 0)
                        w[m] += A.val[j] * A.val[q] / den;
                }
            }
            std::vector<std::pair<uint32_t, double>> row;
            for (uint32_t c : Ci)
                row.emplace_back(c, diag != 0.0 ? -w[c] / diag : 0.0);
            // truncation (trunc factor, then Pmax), each followed by a rescale to the
            // original row sum
            double sum0 = 0.0, mx = 0.0;
            for (auto &e : row) {
                sum0 += e.second;
                mx = std::max(mx, std::fabs(e.second));
            }
            if (trunc_factor > 0.0f) {
                std::vector<std::pair<uint32_t, double>> kept;
                for (auto &e : row)
                    if (std::fabs(e.second) >= (double)trunc_factor * mx)
                        kept.push_back(e);
                double s1 = 0.0;
                for (auto &e : kept)
                    s1 += e.second;
                if (s1 != 0.0 && kept.size() < row.size())
                    for (auto &e : kept)
                        e.second *= sum0 / s1;
                row.swap(kept);
            }
            if (pmax > 0 && (int)row.size() > pmax) {
                double s0 = 0.0;
                for (auto &e : row)
                    s0 += e.second;
                std::stable_sort(row.begin(), row.end(),
                                 [](const std::pair<uint32_t, double> &a,
                                    const std::pair<uint32_t, double> &b) {
                                     return std::fabs(a.second) > std::fabs(b.second);
                                 });
                row.resize(pmax);
                double s1 = 0.0;
                for (auto &e : row)
                    s1 += e.second;
                if (s1 != 0.0)
                    for (auto &e : row)
                        e.second *= s0 / s1;
            }
            std::sort(row.begin(), row.end());
            for (auto &e : row) {
                rc[i].push_back((uint32_t)cidx[e.first]);
                rv[i].push_back(e.second);
            }
        }
    }
    CsrHost P;
    P.nrows = n;
    P.ncols = nc;
    P.row.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i)
        P.row[i + 1] = P.row[i] + (uint32_t)rc[i].size();
    P.col.resize(P.row[n]);
    P.val.resize(P.row[n]);
    for (int64_t i = 0; i < n; ++i) {
        std::copy(rc[i].begin(), rc[i].end(), P.col.begin() + P.row[i]);
        std::copy(rv[i].begin(), rv[i].end(), P.val.begin() + P.row[i]);
    }
    return P;
}

// Second stage of aggressive coarsening: among the first-stage C points, j is a
// long-range strong dependency of i when there are at least num_paths paths of length
// <= 2 from i to j in the strength graph (i -> j, or i -> k -> j).
CsrHost second_strength(const CsrHost &A, const std::vector<uint8_t> &S,
                        const std::vector<int8_t> &cf, int num_paths,
                        std::vector<int64_t> &cidx) {
    const int64_t n = A.nrows;
    cidx.assign(n, -1);
    std::vector<int64_t> cpts;
    for (int64_t i = 0; i < n; ++i)
        if (cf[i] == CPT) {
            cidx[i] = (int64_t)cpts.size();
            cpts.push_back(i);
        }
    const int64_t nc = (int64_t)cpts.size();
    std::vector<std::vector<uint32_t>> rows(nc);
#pragma omp parallel
    {
        std::vector<int> cnt(nc, 0);
        std::vector<uint32_t> touched;
#pragma omp for schedule(dynamic, 256)
        for (int64_t ci = 0; ci < nc; ++ci) {
            const int64_t i = cpts[ci];
            touched.clear();
            auto add = [&](int64_t j) {
                const int64_t cj = cidx[j];
                if (cj < 0 || cj == ci)
                    return;
                if (cnt[cj]++ == 0)
                    touched.push_back((uint32_t)cj);
            };
            for (uint32_t q = A.row[i]; q < A.row[i + 1]; ++q) {
                if (!S[q])
                    continue;
                const int64_t k = A.col[q];
                add(k);
                for (uint32_t r = A.row[k]; r < A.row[k + 1]; ++r)
                    if (S[r])
                        add(A.col[r]);
            }
            std::sort(touched.begin(), touched.end());
            for (uint32_t cj : touched) {
                if (cnt[cj] >= num_paths)
                    rows[ci].push_back(cj);
                cnt[cj] = 0;
            }
        }
    }
    CsrHost G;
    G.nrows = G.ncols = nc;
    G.row.assign(nc + 1, 0);
    for (int64_t ci = 0; ci < nc; ++ci)
        G.row[ci + 1] = G.row[ci] + (uint32_t)rows[ci].size();
    G.col.resize(G.row[nc]);
    for (int64_t ci = 0; ci < nc; ++ci)
        std::copy(rows[ci].begin(), rows[ci].end(), G.col.begin() + G.row[ci]);
    G.val.assign(G.col.size(), -1.0);
    return G;
}

// truncation: drop |w| < trunc * max|w|, keep the pmax largest, each step rescaled to
// the original row sum
void truncate_row(std::vector<std::pair<uint32_t, double>> &row, float trunc, int pmax) {
    double sum0 = 0.0, mx = 0.0;
    for (auto &e : row) {
        sum0 += e.second;
        mx = std::max(mx, std::fabs(e.second));
    }
    if (trunc > 0.0f) {
        std::vector<std::pair<uint32_t, double>> kept;
        for (auto &e : row)
            if (std::fabs(e.second) >= (double)trunc * mx)
                kept.push_back(e);
        double s1 = 0.0;
        for (auto &e : kept)
            s1 += e.second;
        if (s1 != 0.0 && kept.size() < row.size())
            for (auto &e : kept)
                e.second *= sum0 / s1;
        row.swap(kept);
    }
    if (pmax > 0 && (int)row.size() > pmax) {
        double s0 = 0.0;
        for (auto &e : row)
            s0 += e.second;
        std::stable_sort(row.begin(), row.end(),
                         [](const std::pair<uint32_t, double> &a,
                            const std::pair<uint32_t, double> &b) {
                             return std::fabs(a.second) > std::fabs(b.second);
                         });
        row.resize(pmax);
        double s1 = 0.0;
        for (auto &e : row)
            s1 += e.second;
        if (s1 != 0.0)
            for (auto &e : row)
                e.second *= s0 / s1;
    }
    std::sort(row.begin(), row.end());
}

// Multipass interpolation (Stueben): pass 1 interpolates F points with strong C
// neighbours directly (w_ij = -alpha a_ij / a_ii, alpha = sum of all off-diagonal
// entries / sum of the interpolatory ones); pass p interpolates the remaining F points
// through their strong neighbours of earlier passes: P_i = sum_k (-alpha a_ik / a_ii) P_k.
CsrHost multipass(const CsrHost &A, const std::vector<uint8_t> &S, const std::vector<int8_t> &cf,
                  float trunc, int pmax, int64_t &nc) {
    const int64_t n = A.nrows;
    std::vector<int64_t> cidx(n, -1);
    nc = 0;
    for (int64_t i = 0; i < n; ++i)
        if (cf[i] == CPT)
            cidx[i] = nc++;
    std::vector<int> pass(n, -1);
    std::vector<int64_t> frontier;
    for (int64_t i = 0; i < n; ++i)
        if (cf[i] == CPT)
            pass[i] = 0;
    // breadth-first pass numbers over strong dependencies
    for (int p = 1;; ++p) {
        frontier.clear();
        for (int64_t i = 0; i < n; ++i) {
            if (pass[i] >= 0)
                continue;
            for (uint32_t q = A.row[i]; q < A.row[i + 1]; ++q)
                if (S[q] && pass[A.col[q]] >= 0 && pass[A.col[q]] < p) {
                    frontier.push_back(i);
                    break;
                }
        }
        if (frontier.empty())
            break;
        for (int64_t i : frontier)
            pass[i] = p;
    }
    int maxpass = 0;
    for (int64_t i = 0; i < n; ++i)
        maxpass = std::max(maxpass, pass[i]);
    std::vector<std::vector<std::pair<uint32_t, double>>> P(n);
    for (int64_t i = 0; i < n; ++i)
        if (cf[i] == CPT)
            P[i] = {{(uint32_t)cidx[i], 1.0}};
    std::vector<std::vector<int64_t>> members(maxpass + 1);
    for (int64_t i = 0; i < n; ++i)
        if (pass[i] > 0)
            members[pass[i]].push_back(i);
    for (int p = 1; p <= maxpass; ++p) {
        const std::vector<int64_t> &mem = members[p];
#pragma omp parallel
        {
            std::vector<double> acc(nc, 0.0);
            std::vector<uint32_t> touched;
#pragma omp for schedule(dynamic, 256)
            for (size_t q = 0; q < mem.size(); ++q) {
                const int64_t i = mem[q];
                double aii = 0.0, sum_all = 0.0, sum_int = 0.0;
                for (uint32_t r = A.row[i]; r < A.row[i + 1]; ++r) {
                    const uint32_t k = A.col[r];
                    if (k == (uint32_t)i) {
                        aii += A.val[r];
                        continue;
                    }
                    sum_all += A.val[r];
                    if (S[r] && pass[k] >= 0 && pass[k] < p)
                        sum_int += A.val[r];
                }
                if (aii == 0.0 || sum_int == 0.0)
                    continue;
                const double alpha = sum_all / sum_int;
                touched.clear();
                for (uint32_t r = A.row[i]; r < A.row[i + 1]; ++r) {
                    const uint32_t k = A.col[r];
                    if (k == (uint32_t)i || !S[r] || pass[k] < 0 || pass[k] >= p)
                        continue;
                    const double w = -alpha * A.val[r] / aii;
                    for (auto &e : P[k]) {
                        if (acc[e.first] == 0.0)
                            touched.push_back(e.first);
                        acc[e.first] += w * e.second;
                    }
                }
                std::vector<std::pair<uint32_t, double>> row;
                for (uint32_t c : touched) {
                    if (acc[c] != 0.0)
                        row.emplace_back(c, acc[c]);
                    acc[c] = 0.0;
                }
                std::sort(row.begin(), row.end());
                row.erase(std::unique(row.begin(), row.end(),
                                      [](const std::pair<uint32_t, double> &a,
                                         const std::pair<uint32_t, double> &b) {
                                          return a.first == b.first;
                                      }),
                          row.end());
                truncate_row(row, trunc, pmax);
                P[i] = row;
            }
        }
    }
    CsrHost Pm;
    Pm.nrows = n;
    Pm.ncols = nc;
    Pm.row.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i)
        Pm.row[i + 1] = Pm.row[i] + (uint32_t)P[i].size();
    Pm.col.resize(Pm.row[n]);
    Pm.val.resize(Pm.row[n]);
    for (int64_t i = 0; i < n; ++i)
        for (size_t q = 0; q < P[i].size(); ++q) {
            Pm.col[Pm.row[i] + q] = P[i][q].first;
            Pm.val[Pm.row[i] + q] = P[i][q].second;
        }
    return Pm;
}

uint64_t splitmix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// PMIS (parallel modified independent set).  Measure w_i = |S^T_i| + u_i with u_i in
// (0, 1) from a counter-based generator (equal measures, which the generator makes
// vanishingly rare, are ordered by index), so the result does not depend on the thread
// count.  A point that no point depends on becomes F (with no strong
// dependency it gets an empty row of P, as an isolated point of the RS pass; with
// isolated_as_c it becomes C, as in the second stage of aggressive coarsening).  Each
// round, every undecided point whose measure exceeds that of all its undecided strong
// neighbours (S_i and S^T_i) becomes C, then every undecided point that strongly depends
// on one of the new C points becomes F.  Rounds see the state at their start, as the
// parallel algorithm does.
std::vector<int8_t> pmis_coarsen(const CsrHost &A, const std::vector<uint8_t> &S, uint64_t seed,
                                 bool isolated_as_c = false) {
    const int64_t n = A.nrows;
    std::vector<uint32_t> tr(n + 1, 0);
    for (int64_t i = 0; i < n; ++i)
        for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j)
            if (S[j] && (int64_t)A.col[j] != i)
                tr[A.col[j] + 1]++;
    for (int64_t i = 0; i < n; ++i)
        tr[i + 1] += tr[i];
    std::vector<uint32_t> tc(tr[n]);
    {
        std::vector<uint32_t> pos(tr.begin(), tr.end() - 1);
        for (int64_t i = 0; i < n; ++i)
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j)
                if (S[j] && (int64_t)A.col[j] != i)
                    tc[pos[A.col[j]]++] = (uint32_t)i;
    }
    std::vector<double> w(n);
    std::vector<int8_t> cf(n, UNDECIDED);
    for (int64_t i = 0; i < n; ++i) {
        const uint64_t z = splitmix64(seed * 0xD1B54A32D192ED03ull + (uint64_t)i);
        const double u = ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const int64_t deg = tr[i + 1] - tr[i];
        w[i] = (double)deg + u;
        bool dep = false;
        for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j)
            dep |= (S[j] != 0 && (int64_t)A.col[j] != i);
        if (deg == 0)
            cf[i] = (!dep && isolated_as_c) ? CPT : FPT;
    }
    std::vector<int64_t> und;
    for (int64_t i = 0; i < n; ++i)
        if (cf[i] == UNDECIDED)
            und.push_back(i);
    std::vector<uint8_t> newc(n, 0);
    while (!und.empty()) {
        std::vector<int64_t> cs;
        for (int64_t i : und) {
            // (w, index) order: equal measures cannot stall the selection
            auto beats = [&](int64_t k) { return w[k] > w[i] || (w[k] == w[i] && k > i); };
            bool best = true;
            for (uint32_t j = A.row[i]; best && j < A.row[i + 1]; ++j) {
                const int64_t k = A.col[j];
                if (S[j] && k != i && cf[k] == UNDECIDED && beats(k))
                    best = false;
            }
            for (uint32_t q = tr[i]; best && q < tr[i + 1]; ++q) {
                const int64_t k = tc[q];
                if (cf[k] == UNDECIDED && beats(k))
                    best = false;
            }
            if (best)
                cs.push_back(i);
        }
        for (int64_t i : cs)
            cf[i] = CPT;
        for (int64_t c : cs)
            for (uint32_t q = tr[c]; q < tr[c + 1]; ++q)
                if (cf[tc[q]] == UNDECIDED)
                    cf[tc[q]] = FPT;
        std::vector<int64_t> rest;
        for (int64_t i : und)
            if (cf[i] == UNDECIDED)
                rest.push_back(i);
        und.swap(rest);
    }
    return cf;
}

// Extended+i interpolation.  For an F point i with strong C neighbours C_i and strong F
// neighbours F_i, the interpolatory set is C^_i = C_i plus the strong C neighbours of
// every k in F_i.  With abar_kl = a_kl when a_kl and a_kk have opposite signs (else 0)
// and D_k = sum over l in C^_i + {i} of abar_kl:
//   w_ij = -(a_ij + sum_{k in F_i} a_ik abar_kj / D_k) / d_i,   j in C^_i
//   d_i  = a_ii + (sum of a_in over the other neighbours n, weak ones outside C^_i)
//               + sum_{k in F_i} a_ik abar_ki / D_k,
// a neighbour k with D_k = 0 lumped into d_i.  A zero-row-sum row interpolates constants
// exactly (the weights sum to 1); truncation then rescales each row to its sum.
CsrHost ext_i(const CsrHost &A, const std::vector<uint8_t> &S, const std::vector<int8_t> &cf,
              float trunc, int pmax, int64_t &nc) {
    const int64_t n = A.nrows;
    std::vector<int64_t> cidx(n, -1);
    nc = 0;
    for (int64_t i = 0; i < n; ++i)
        if (cf[i] == CPT)
            cidx[i] = nc++;
    std::vector<std::vector<std::pair<uint32_t, double>>> rows(n);
#pragma omp parallel
    {
        std::vector<int64_t> mark(n, -1);
        std::vector<double> w(n, 0.0);
#pragma omp for schedule(dynamic, 512)
        for (int64_t i = 0; i < n; ++i) {
            if (cf[i] == CPT) {
                rows[i] = {{(uint32_t)cidx[i], 1.0}};
                continue;
            }
            std::vector<uint32_t> chat;
            auto add_c = [&](uint32_t c) {
                if (mark[c] != i) {
                    mark[c] = i;
                    w[c] = 0.0;
                    chat.push_back(c);
                }
            };
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j) {
                const uint32_t k = A.col[j];
                if (k == (uint32_t)i || !S[j])
                    continue;
                if (cf[k] == CPT)
                    add_c(k);
                else
                    for (uint32_t q = A.row[k]; q < A.row[k + 1]; ++q)
                        if (S[q] && A.col[q] != k && cf[A.col[q]] == CPT)
                            add_c(A.col[q]);
            }
            if (chat.empty())
                continue; // no interpolatory point: empty row of P
            double d = 0.0;
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j) {
                const uint32_t k = A.col[j];
                if (k == (uint32_t)i)
                    d += A.val[j];
                else if (mark[k] == i)
                    w[k] += A.val[j];
                else if (!(S[j] && cf[k] != CPT))
                    d += A.val[j];
            }
            for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j) {
                const uint32_t k = A.col[j];
                if (k == (uint32_t)i || !S[j] || cf[k] == CPT)
                    continue;
                const double akk = diag_of(A, k);
                double D = 0.0, aki = 0.0;
                for (uint32_t q = A.row[k]; q < A.row[k + 1]; ++q) {
                    const uint32_t l = A.col[q];
                    if (l == k || A.val[q] * akk >= 0.0)
                        continue;
                    if (l == (uint32_t)i) {
                        D += A.val[q];
                        aki += A.val[q];
                    } else if (mark[l] == i) {
                        D += A.val[q];
                    }
                }
                if (D == 0.0) {
                    d += A.val[j];
                    continue;
                }
                const double f = A.val[j] / D;
                for (uint32_t q = A.row[k]; q < A.row[k + 1]; ++q) {
                    const uint32_t l = A.col[q];
                    if (l != k && l != (uint32_t)i && mark[l] == i && A.val[q] * akk < 0.0)
                        w[l] += f * A.val[q];
                }
                d += f * aki;
            }
            std::vector<std::pair<uint32_t, double>> row;
            for (uint32_t c : chat)
                row.emplace_back(c, d != 0.0 ? -w[c] / d : 0.0);
            truncate_row(row, trunc, pmax);
            for (auto &e : row)
                e.first = (uint32_t)cidx[e.first];
            std::sort(row.begin(), row.end());
            rows[i] = std::move(row);
        }
    }
    CsrHost P;
    P.nrows = n;
    P.ncols = nc;
    P.row.assign(n + 1, 0);
    for (int64_t i = 0; i < n; ++i)
        P.row[i + 1] = P.row[i] + (uint32_t)rows[i].size();
    P.col.resize(P.row[n]);
    P.val.resize(P.row[n]);
    for (int64_t i = 0; i < n; ++i)
        for (size_t q = 0; q < rows[i].size(); ++q) {
            P.col[P.row[i] + q] = rows[i][q].first;
            P.val[P.row[i] + q] = rows[i][q].second;
        }
    return P;
}

CsrHost interpolate_nc(const CsrHost &A, const std::vector<uint8_t> &S,
                       const std::vector<int8_t> &cf, int type, float trunc, int pmax,
                       int64_t &nc) {
    if (type == INTERP_EXT_I)
        return ext_i(A, S, cf, trunc, pmax, nc);
    if (type == INTERP_MULTIPASS)
        return multipass(A, S, cf, trunc, pmax, nc);
    return interpolation(A, S, cf, trunc, pmax, nc);
}

} // namespace

LevelKeys SetupParams::keys(int level) const {
    if (level >= 0 && level < (int)per_level.size())
        return per_level[level];
    LevelKeys k;
    k.strength = strength;
    k.max_row_sum = max_row_sum;
    k.trunc_factor = trunc_factor;
    k.pmax = pmax;
    k.coarsen_type = coarsen_type;
    k.interp_type = interp_type;
    k.agg_interp_type = agg_interp_type;
    k.agg_trunc_factor = agg_trunc_factor;
    k.agg_pmax = agg_pmax;
    k.num_paths = num_paths;
    return k;
}

bool coarsen_supported(int type) { return type == COARSEN_FALGOUT || type == COARSEN_PMIS; }
bool interp_supported(int type) { return type == INTERP_CLASSICAL || type == INTERP_EXT_I; }
bool agg_interp_supported(int type) { return type == INTERP_MULTIPASS; }

std::vector<uint8_t> strength_of(const CsrHost &A, float theta, float max_row_sum) {
    return strength(A, theta, max_row_sum);
}

std::vector<int8_t> coarsen(const CsrHost &A, const std::vector<uint8_t> &S, int type,
                            int second_pass, uint64_t seed) {
    if (type == COARSEN_PMIS)
        return pmis_coarsen(A, S, seed);
    return rs_coarsen(A, S, second_pass);
}

CsrHost interpolate(const CsrHost &A, const std::vector<uint8_t> &S,
                    const std::vector<int8_t> &cf, int type, float trunc, int pmax) {
    int64_t nc = 0;
    return interpolate_nc(A, S, cf, type, trunc, pmax, nc);
}

void estimate_eigs(const CsrHost &A, int iters, uint64_t seed, double &mn, double &mx) {
    const int64_t n = A.nrows;
    std::vector<double> isd(n), x(n), y(n);
    for (int64_t i = 0; i < n; ++i) {
        double d = diag_of(A, i);
        isd[i] = d != 0.0 ? std::sqrt(1.0 / d) : 0.0;
        uint64_t z = splitmix64(seed * 0x100000001B3ull + (uint64_t)i);
        x[i] = 2.0 * ((double)(z >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
    // dots as fixed-size chunk partials summed in chunk order: the estimate (and with it
    // every Chebyshev coefficient) is the same whatever the thread count and schedule
    constexpr int64_t kChunk = 4096;
    const int64_t nchunk = (n + kChunk - 1) / kChunk;
    std::vector<double> pxy(nchunk), pxx(nchunk), pyy(nchunk);
    double lam = 0.0;
    for (int it = 0; it < std::max(1, iters); ++it) {
#pragma omp parallel for schedule(static)
        for (int64_t c = 0; c < nchunk; ++c) {
            double cxy = 0.0, cxx = 0.0, cyy = 0.0;
            const int64_t i1 = std::min(n, (c + 1) * kChunk);
            for (int64_t i = c * kChunk; i < i1; ++i) {
                double s = 0.0;
                for (uint32_t j = A.row[i]; j < A.row[i + 1]; ++j)
                    s += A.val[j] * isd[A.col[j]] * x[A.col[j]];
                y[i] = isd[i] * s;
                cxy += x[i] * y[i];
                cxx += x[i] * x[i];
                cyy += y[i] * y[i];
            }
            pxy[c] = cxy;
            pxx[c] = cxx;
            pyy[c] = cyy;
        }
        double xy = 0.0, xx = 0.0, yy = 0.0;
        for (int64_t c = 0; c < nchunk; ++c) {
            xy += pxy[c];
            xx += pxx[c];
            yy += pyy[c];
        }
        lam = xx > 0 ? xy / xx : 0.0;
        double nrm = std::sqrt(yy);
        if (nrm == 0.0)
            break;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            x[i] = y[i] / nrm;
    }
    mx = lam;
    mn = 0.0;
}

void build(const CsrHost &A0, const SetupParams &p, Hierarchy &H) {
    H = Hierarchy();
    CsrHost A = A0;
    sort_rows_diag_first(A);
    H.A.push_back(A);
    while ((int)H.A.size() < p.max_levels && H.A.back().nrows > p.coarse_size) {
        const CsrHost &Af = H.A.back();
        const int lev = (int)H.A.size() - 1;
        const LevelKeys k = p.keys(lev);
        std::vector<uint8_t> S = strength(Af, k.strength, k.max_row_sum);
        const uint64_t seed = p.pmis_seed + 0x9E3779B97F4A7C15ull * (uint64_t)lev;
        std::vector<int8_t> cf = coarsen(Af, S, k.coarsen_type, p.second_pass, seed);
        int64_t nc = 0;
        CsrHost P;
        if (lev < p.agg_num_levels) {
            // aggressive coarsening: a second pass of the same scheme on the first-stage C
            // points with the long-range strength graph, then multipass interpolation
            std::vector<int64_t> cidx;
            CsrHost G = second_strength(Af, S, cf, k.num_paths, cidx);
            std::vector<uint8_t> S2(G.col.size(), 1);
            std::vector<int8_t> cf2 = k.coarsen_type == COARSEN_PMIS
                                          ? pmis_coarsen(G, S2, seed ^ 0x5851F42D4C957F2Dull, true)
                                          : rs_coarsen(G, S2, 0, true);
            for (int64_t i = 0; i < Af.nrows; ++i)
                if (cf[i] == CPT && cf2[cidx[i]] != CPT)
                    cf[i] = FPT;
            P = multipass(Af, S, cf, k.agg_trunc_factor, k.agg_pmax, nc);
        } else {
            P = interpolate_nc(Af, S, cf, k.interp_type, k.trunc_factor, k.pmax, nc);
        }
        if (nc == 0 || nc >= Af.nrows)
            break;
        CsrHost R = transpose(P);
        CsrHost AP = spgemm(Af, P);
        CsrHost Ac = spgemm(R, AP);
        sort_rows_diag_first(Ac);
        H.P.push_back(std::move(P));
        H.R.push_back(std::move(R));
        H.A.push_back(std::move(Ac));
    }
    H.min_eig.assign(H.A.size(), 0.0);
    H.max_eig.assign(H.A.size(), 0.0);
    if (p.eig_iters > 0)
        for (size_t l = 0; l < H.A.size(); ++l)
            estimate_eigs(H.A[l], p.eig_iters, p.eig_seed, H.min_eig[l], H.max_eig[l]);
}

} // namespace amg
} // namespace xamg
