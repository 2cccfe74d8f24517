/*
 * XAMG solve-phase oracle -- TEST INFRASTRUCTURE ONLY (see xamg_oracle.h).
 *
 * Single-threaded restatement of the reference solve phase.  Vectors carry the
 * reference's `if_zero` flag (src/primitives/vector/vector.h:88-94) because the
 * reference's short-circuits decide which arithmetic is done (e.g. an SpMV of a
 * zero-flagged vector is skipped, src/blas2/blas2.h:309-313).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   spmv / spmv-add     dense brute force, integer-exact inputs, cube eigenvectors
 *   dense matvec, GJ    numpy.linalg.inv / brute force on small SPD matrices
 *   level-1 + merged    closed forms, integer-exact dots, ||s-wt||^2 = theta - w*phi
 *   Chebyshev coeffs    shifted Chebyshev residual polynomial (numpy.polynomial)
 *   HSGS                textbook SSOR (1 block) / block-Jacobi-of-SSOR (P blocks)
 *   Jacobi              textbook weighted Jacobi
 *   Direct              numpy.linalg.solve
 *   MultiGrid           2-level cycle on 1D Poisson with explicit P; linearity
 *   BiCGStab family     A = I converges in one iteration; diagonal A with m distinct
 *                       eigenvalues converges in <= m iterations; residual histories
 *                       vs scipy-free dense restatement on tiny systems
 *   compress            round trip decode(encode(A)) == A; width thresholds
 *   cube generator      nnz = 7N - 6 edge^2; closed-form eigenvalues
 */
#include "xamg_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <type_traits>
#include <vector>
#include <omp.h>

namespace xo {

static int g_err = 0; /* set on a condition where the reference asserts */

/* Loops over rows or vector elements run on all OpenMP threads: each element is
 * computed by one thread with the same operations, so results do not depend on the
 * thread count.  Reductions keep their own fixed order (per-core partials, see dot). */
#define XO_PAR _Pragma("omp parallel for schedule(static)")

/* vector with the reference's zero flag; data layout x[i*nv+k]
 * (src/primitives/vector/vector.h:509-526) */
template <typename T>
struct Vec {
    std::vector<T> d;
    uint64_t n = 0;
    int nv = 0;
    bool zero = true;
    void alloc(uint64_t n_, int nv_) {
        n = n_;
        nv = nv_;
        d.assign((size_t)n_ * nv_, T(0));
        zero = true;
    }
    /* a work buffer kept between calls, as the reference keeps its solver buffers: its
     * memory is reused and only the zero flag is set (set_const(buf, 0.0) without force,
     * blas_wrapper.inl:30-42); nothing reads a zero-flagged vector's data */
    void reuse(uint64_t n_, int nv_) {
        if (n != n_ || nv != nv_ || d.size() != (size_t)n_ * nv_)
            alloc(n_, nv_);
        zero = true;
    }
    T *p() { return d.data(); }
    const T *p() const { return d.data(); }
    size_t sz() const { return (size_t)n * nv; }
};

template <typename T>
static Vec<T> const_vec(int nv, double v) {
    /* ConstVectorsCache a0/a1/a_1 (src/blas/blas.h:140-156) */
    Vec<T> a;
    a.alloc(1, nv);
    for (int k = 0; k < nv; ++k)
        a.d[k] = (T)v;
    a.zero = !(std::fabs(v) > 1.e-100);
    return a;
}

/* ---- partition helpers: misc::split_range (src/misc/misc.h:46-55) with
 *      XAMG_ALIGN_SIZE 64 / sizeof(float64_t) = 8; core ranges as
 *      part::get_part (src/part/part.h:115-135) ---- */
static uint64_t split_range(uint64_t size, int nblocks) {
    const uint64_t factor = 64 / sizeof(double);
    uint64_t core_size = size / (uint64_t)nblocks;
    core_size += factor - (core_size % factor);
    return core_size;
}

static std::vector<uint64_t> core_ranges(uint64_t size, int nblocks) {
    std::vector<uint64_t> off(nblocks + 1, 0);
    uint64_t cs = split_range(size, nblocks);
    for (int c = 0; c < nblocks; ++c) {
        uint64_t o = std::min(size, cs * (uint64_t)c);
        uint64_t s = std::min(cs, size - o);
        off[c] = o;
        off[c + 1] = o + s;
    }
    off[nblocks] = size;
    return off;
}

/* global emulated core count for dot partial sums */
static int g_dot_ncores = 1;

/* ---------------- level-1 BLAS (src/blas/detail/blas.inl, blas_ext.inl) ---------------- */

/* scal / ax_y coefficient access: a.nv==1 broadcasts (blas.inl:61-71) */
template <typename T>
static void ax_y(const Vec<T> &a, const Vec<T> &x, Vec<T> &y) { /* blas.inl:31-85 */
    if (a.zero || x.zero) {
        y.zero = true;
        return;
    }
    const int nv = x.nv;
    const size_t N = x.sz();
    if (a.nv == 1) {
        XO_PAR
        for (size_t i = 0; i < N; ++i)
            y.d[i] = a.d[0] * x.d[i];
    } else {
        XO_PAR
        for (size_t i = 0; i < N; i += nv)
            for (int k = 0; k < nv; ++k)
                y.d[i + k] = a.d[k] * x.d[i + k];
    }
    y.zero = false;
}

template <typename T>
static void scal(const Vec<T> &x, Vec<T> &y) { /* blas.inl:156-219 */
    if (x.zero || y.zero) {
        y.zero = true;
        return;
    }
    const int nv = y.nv;
    if (x.nv == 1) {
        if (x.n == 1) {
            XO_PAR
            for (size_t i = 0; i < y.sz(); ++i)
                y.d[i] *= x.d[0];
        } else {
            XO_PAR
            for (uint64_t i = 0; i < y.n; ++i)
                for (int k = 0; k < nv; ++k)
                    y.d[i * nv + k] *= x.d[i];
        }
    } else {
        if (x.n == 1) {
            XO_PAR
            for (size_t i = 0; i < y.sz(); i += nv)
                for (int k = 0; k < nv; ++k)
                    y.d[i + k] *= x.d[k];
        } else {
            XO_PAR
            for (size_t i = 0; i < y.sz(); ++i)
                y.d[i] *= x.d[i];
        }
    }
}

template <typename T>
static void axpby(const Vec<T> &a, const Vec<T> &x, const Vec<T> &b, Vec<T> &y) { /* blas.inl:90-151 */
    if (a.zero || x.zero) {
        scal(b, y);
        return;
    }
    if (b.zero || y.zero) {
        ax_y(a, x, y);
        return;
    }
    const int nv = y.nv;
    if (a.nv == 1) {
        XO_PAR
        for (size_t i = 0; i < y.sz(); ++i)
            y.d[i] = a.d[0] * x.d[i] + b.d[0] * y.d[i];
    } else {
        XO_PAR
        for (size_t i = 0; i < y.sz(); i += nv)
            for (int k = 0; k < nv; ++k)
                y.d[i + k] = a.d[k] * x.d[i + k] + b.d[k] * y.d[i + k];
    }
    y.zero = false;
}

template <typename T>
static void axpby_s(double a, const Vec<T> &x, double b, Vec<T> &y) { /* blas_wrapper.inl:76-99 */
    Vec<T> va = const_vec<T>(y.nv, a), vb = const_vec<T>(y.nv, b);
    axpby(va, x, vb, y);
}

template <typename T>
static void copy(const Vec<T> &x, Vec<T> &y) { /* blas.inl:224-307 */
    if (x.zero) {
        y.zero = true;
        return;
    }
    y.d = x.d;
    y.n = x.n;
    y.nv = x.nv;
    y.zero = false;
}

/* set_const<T,NV>(x, val, force) wrapper + set_const(x, vec, force)
 * (blas_wrapper.inl:30-42, blas.inl:338-385) */
template <typename T>
static void set_const(Vec<T> &x, double val, bool force = false) {
    bool vzero = !(std::fabs(val) > 1.e-100);
    if (vzero && !force) {
        x.zero = true;
        return;
    }
    std::fill(x.d.begin(), x.d.end(), (T)val);
    x.zero = false;
}

template <typename T>
static void dot(const Vec<T> &x, const Vec<T> &y, Vec<T> &res) { /* blas.inl:465-532 */
    const int nv = x.nv;
    if (x.zero || y.zero) {
        set_const(res, 0.0, true);
        return;
    }
    /* per-core sequential partial sums combined in rank order
     * (src/comm/shm_allreduce.h:112-135); one core = one sequential sum */
    std::vector<uint64_t> off = core_ranges(x.n, g_dot_ncores);
    std::vector<std::vector<T>> parts(g_dot_ncores, std::vector<T>(nv, T(0)));
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < g_dot_ncores; ++c) {
        std::vector<T> &part = parts[c];
        for (uint64_t i = off[c]; i < off[c + 1]; ++i)
            for (int k = 0; k < nv; ++k)
                part[k] += x.d[i * nv + k] * y.d[i * nv + k];
    }
    std::vector<T> acc(nv, T(0));
    for (int c = 0; c < g_dot_ncores; ++c)
        for (int k = 0; k < nv; ++k)
            acc[k] = (c == 0) ? parts[c][k] : acc[k] + parts[c][k];
    for (int k = 0; k < nv; ++k)
        res.d[k] = acc[k];
    res.zero = false;
}

template <typename T>
static void axpby_z(const Vec<T> &a, const Vec<T> &x, const Vec<T> &b, const Vec<T> &y,
                    Vec<T> &z) { /* blas_ext.inl:31-102 */
    if (a.zero || x.zero) {
        ax_y(b, y, z);
        return;
    }
    if (b.zero || y.zero) {
        ax_y(a, x, z);
        return;
    }
    const int nv = x.nv;
    if (a.nv == 1) {
        XO_PAR
        for (size_t i = 0; i < x.sz(); i += nv)
            for (int k = 0; k < nv; ++k)
                z.d[i + k] = a.d[0] * x.d[i + k] + b.d[0] * y.d[i + k];
    } else {
        XO_PAR
        for (size_t i = 0; i < x.sz(); i += nv)
            for (int k = 0; k < nv; ++k)
                z.d[i + k] = a.d[k] * x.d[i + k] + b.d[k] * y.d[i + k];
    }
    z.zero = false;
}

template <typename T>
static void axpbypcz(const Vec<T> &a, const Vec<T> &x, const Vec<T> &b, const Vec<T> &y,
                     const Vec<T> &c, Vec<T> &z) { /* blas_ext.inl:107-189 */
    if (a.zero || x.zero) {
        axpby(b, y, c, z);
        return;
    }
    if (b.zero || y.zero) {
        axpby(a, x, c, z);
        return;
    }
    if (c.zero || z.zero) {
        axpby_z(a, x, b, y, z);
        return;
    }
    const int nv = x.nv;
    if (a.nv == 1) {
        XO_PAR
        for (size_t i = 0; i < x.sz(); i += nv)
            for (int k = 0; k < nv; ++k)
                z.d[i + k] = a.d[0] * x.d[i + k] + b.d[0] * y.d[i + k] + c.d[0] * z.d[i + k];
    } else {
        XO_PAR
        for (size_t i = 0; i < x.sz(); i += nv)
            for (int k = 0; k < nv; ++k)
                z.d[i + k] = a.d[k] * x.d[i + k] + b.d[k] * y.d[i + k] + c.d[k] * z.d[i + k];
    }
    z.zero = false;
}

template <typename T>
static void xdivy_z(const Vec<T> &x, const Vec<T> &y, Vec<T> &z) { /* blas_ext.inl:194-239 */
    if (x.zero) {
        set_const(z, 0.0);
        return;
    }
    if (y.zero) {
        g_err = -2; /* "Div by 0!" assert */
        return;
    }
    XO_PAR
    for (size_t i = 0; i < x.sz(); ++i)
        z.d[i] = x.d[i] / y.d[i];
    z.zero = false;
}

/* ---------------- merged BiCGStab groups (blas_merged_bicgstab.inl) ---------------- */

template <typename T>
static void group1(Vec<T> &r0, Vec<T> &b, Vec<T> &p, Vec<T> &r, Vec<T> &rho0) { /* :33-87 */
    if (b.zero)
        set_const(b, 0.0, true);
    if (r0.zero)
        set_const(r0, 0.0, true);
    const int nv = r0.nv;
    std::vector<uint64_t> off = core_ranges(r0.n, g_dot_ncores);
    std::vector<std::vector<T>> parts(g_dot_ncores, std::vector<T>(nv, T(0)));
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < g_dot_ncores; ++c) {
        std::vector<T> &part = parts[c];
        for (uint64_t i = off[c] * nv; i < off[c + 1] * nv; i += nv)
            for (int k = 0; k < nv; ++k) {
                r0.d[i + k] = b.d[i + k] - r0.d[i + k];
                part[k] += r0.d[i + k] * r0.d[i + k];
                p.d[i + k] = r0.d[i + k];
                r.d[i + k] = r0.d[i + k];
            }
    }
    std::vector<T> acc(nv, T(0));
    for (int c = 0; c < g_dot_ncores; ++c)
        for (int k = 0; k < nv; ++k)
            acc[k] = (c == 0) ? parts[c][k] : acc[k] + parts[c][k];
    for (int k = 0; k < nv; ++k)
        rho0.d[k] = acc[k];
    r0.zero = rho0.zero = p.zero = r.zero = false;
}

template <typename T>
static void group2(Vec<T> &t, Vec<T> &s, Vec<T> &phi, Vec<T> &psi, Vec<T> &theta) { /* :89-147 */
    if (t.zero)
        set_const(t, 0.0, true);
    if (s.zero)
        set_const(s, 0.0, true);
    const int nv = s.nv;
    std::vector<uint64_t> off = core_ranges(s.n, g_dot_ncores);
    std::vector<std::vector<T>> parts(g_dot_ncores, std::vector<T>(3 * nv, T(0)));
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < g_dot_ncores; ++c) {
        T *p1 = parts[c].data(), *p2 = p1 + nv, *p3 = p2 + nv;
        for (uint64_t i = off[c] * nv; i < off[c + 1] * nv; i += nv)
            for (int k = 0; k < nv; ++k) {
                p1[k] += t.d[i + k] * s.d[i + k];
                p2[k] += t.d[i + k] * t.d[i + k];
                p3[k] += s.d[i + k] * s.d[i + k];
            }
    }
    std::vector<T> a1(nv, 0), a2(nv, 0), a3(nv, 0);
    for (int c = 0; c < g_dot_ncores; ++c) {
        const T *p1 = parts[c].data(), *p2 = p1 + nv, *p3 = p2 + nv;
        for (int k = 0; k < nv; ++k) {
            a1[k] = (c == 0) ? p1[k] : a1[k] + p1[k];
            a2[k] = (c == 0) ? p2[k] : a2[k] + p2[k];
            a3[k] = (c == 0) ? p3[k] : a3[k] + p3[k];
        }
    }
    for (int k = 0; k < nv; ++k) {
        phi.d[k] = a1[k];
        psi.d[k] = a2[k];
        theta.d[k] = a3[k];
    }
    phi.zero = psi.zero = theta.zero = false;
}

template <typename T>
static void group4(Vec<T> &s, Vec<T> &t, Vec<T> &r0, Vec<T> &r, const Vec<T> &omega,
                   Vec<T> &rho) { /* :149-194 */
    if (s.zero)
        set_const(s, 0.0, true);
    if (t.zero)
        set_const(t, 0.0, true);
    if (r0.zero)
        set_const(r0, 0.0, true);
    const int nv = s.nv;
    std::vector<uint64_t> off = core_ranges(s.n, g_dot_ncores);
    std::vector<std::vector<T>> parts(g_dot_ncores, std::vector<T>(nv, T(0)));
#pragma omp parallel for schedule(static, 1)
    for (int c = 0; c < g_dot_ncores; ++c) {
        std::vector<T> &part = parts[c];
        for (uint64_t i = off[c] * nv; i < off[c + 1] * nv; i += nv)
            for (int k = 0; k < nv; ++k) {
                r.d[i + k] = s.d[i + k] - omega.d[k] * t.d[i + k];
                part[k] += r.d[i + k] * r0.d[i + k];
            }
    }
    std::vector<T> acc(nv, 0);
    for (int c = 0; c < g_dot_ncores; ++c)
        for (int k = 0; k < nv; ++k)
            acc[k] = (c == 0) ? parts[c][k] : acc[k] + parts[c][k];
    for (int k = 0; k < nv; ++k)
        rho.d[k] = acc[k];
    r.zero = rho.zero = false;
}

template <typename T>
static void group5(Vec<T> &x, Vec<T> &p, Vec<T> &s, Vec<T> &r, Vec<T> &v, const Vec<T> &ac,
                   const Vec<T> &oc, const Vec<T> &alpha, const Vec<T> &omega,
                   const Vec<T> &beta) { /* :198-257 */
    (void)alpha;
    if (x.zero)
        set_const(x, 0.0, true);
    if (p.zero)
        set_const(p, 0.0, true);
    if (s.zero)
        set_const(s, 0.0, true);
    if (r.zero)
        set_const(r, 0.0, true);
    if (v.zero)
        set_const(v, 0.0, true);
    const int nv = x.nv;
    std::vector<T> val(nv);
    for (int k = 0; k < nv; ++k)
        val[k] = beta.d[k] * omega.d[k];
    XO_PAR
    for (size_t i = 0; i < x.sz(); i += nv)
        for (int k = 0; k < nv; ++k) {
            x.d[i + k] += ac.d[k] * p.d[i + k] + oc.d[k] * s.d[i + k];
            p.d[i + k] = r.d[i + k] + beta.d[k] * p.d[i + k] - val[k] * v.d[i + k];
        }
    x.zero = p.zero = false;
}

/* ---------------- CSR (csr_matvec.inl, csr_matrix.h) ---------------- */

template <typename T>
struct Csr {
    uint64_t nrows = 0, ncols = 0;
    std::vector<uint32_t> row, col;
    std::vector<T> val;
};

template <typename T>
static Csr<T> make_csr(const xo_csr *m) {
    /* hypre_base::parse_matrix casts the values to F (hypre_base.inl:292-301) */
    Csr<T> c;
    c.nrows = m->nrows;
    c.ncols = m->ncols;
    c.row.assign(m->row, m->row + m->nrows + 1);
    c.col.assign(m->col, m->col + m->nnz);
    c.val.resize(m->nnz);
    for (int64_t j = 0; j < m->nnz; ++j)
        c.val[j] = (T)m->val[j];
    return c;
}

template <typename T>
static void csr_ax_y(const Csr<T> &m, const T *x, T *y, int nv) { /* csr_matvec.inl:213-231 */
    XO_PAR
    for (uint64_t i = 0; i < m.nrows; ++i) {
        T *yr = y + i * nv;
        for (int k = 0; k < nv; ++k)
            yr[k] = 0.0;
        for (uint32_t j = m.row[i]; j < m.row[i + 1]; ++j) {
            const T *xr = x + (uint64_t)m.col[j] * nv;
            for (int k = 0; k < nv; ++k)
                yr[k] += xr[k] * m.val[j];
        }
    }
}

template <typename T>
static void csr_axpy(const Csr<T> &m, const T *x, T *y, int nv) { /* csr_matvec.inl:103-116 */
    XO_PAR
    for (uint64_t i = 0; i < m.nrows; ++i) {
        T *yr = y + i * nv;
        for (uint32_t j = m.row[i]; j < m.row[i + 1]; ++j) {
            const T *xr = x + (uint64_t)m.col[j] * nv;
            for (int k = 0; k < nv; ++k)
                yr[k] += xr[k] * m.val[j];
        }
    }
}

/* blas2::Ax_y / Axpy with the zero-flag short-circuits (blas2.h:284-326) */
template <typename T>
static void Ax_y(const Csr<T> &m, const Vec<T> &x, Vec<T> &y) {
    if (x.zero) {
        y.zero = true;
        return;
    }
    csr_ax_y(m, x.p(), y.p(), x.nv);
    y.zero = false;
}

template <typename T>
static void Axpy(const Csr<T> &m, const Vec<T> &x, Vec<T> &y) {
    if (x.zero)
        return;
    if (y.zero)
        set_const(y, 0.0, true); /* y is written through; a flagged y holds no data */
    csr_axpy(m, x.p(), y.p(), x.nv);
    y.zero = false;
}

template <typename T>
static std::vector<T> inv_diag(const Csr<T> &m) { /* csr_matrix.h:302-334 */
    std::vector<T> d(m.nrows, T(0));
    for (uint64_t i = 0; i < m.nrows; ++i)
        for (uint32_t j = m.row[i]; j < m.row[i + 1]; ++j)
            if (m.col[j] == (uint32_t)i)
                d[i] = 1.0 / m.val[j];
    return d;
}

template <typename T>
static std::vector<T> inv_sqrt_diag(const Csr<T> &m) { /* csr_matrix.h:336-368 */
    std::vector<T> d(m.nrows, T(0));
    for (uint64_t i = 0; i < m.nrows; ++i)
        for (uint32_t j = m.row[i]; j < m.row[i + 1]; ++j)
            if (m.col[j] == (uint32_t)i)
                d[i] = std::sqrt(1.0 / m.val[j]);
    return d;
}

/* ---------------- dense (dense_matvec.inl, inverse.inl) ---------------- */

template <typename T>
static void dense_ax_y(uint64_t n, uint64_t m, const T *D, const T *x, T *y, int nv) {
    /* dense_matvec.inl:109-117 */
    XO_PAR
    for (uint64_t i = 0; i < n; ++i) {
        for (int k = 0; k < nv; ++k)
            y[i * nv + k] = 0.0;
        for (uint64_t j = 0; j < m; ++j)
            for (int k = 0; k < nv; ++k)
                y[i * nv + k] += D[i * m + j] * x[j * nv + k];
    }
}

template <typename T>
static void gj_inverse(uint64_t n, const T *A, T *Ainv) { /* inverse.inl:31-106 */
    const uint64_t nc = 2 * n;
    std::vector<T> e(n * nc, T(0));
    for (uint64_t i = 0; i < n; ++i) {
        e[i * nc + i + n] = 1.0;
        for (uint64_t j = 0; j < n; ++j)
            e[i * nc + j] = A[i * n + j];
    }
    for (uint64_t i = 0; i < n; ++i) {
        T *ri = &e[i * nc];
        for (uint64_t k = nc - 1; k > i; --k)
            ri[k] /= ri[i];
        ri[i] = 1.0;
        for (uint64_t j = i + 1; j < n; ++j) {
            T *rj = &e[j * nc];
            for (uint64_t k = nc - 1; k > i; --k)
                rj[k] -= rj[i] * ri[k];
            rj[i] = 0.0;
        }
    }
    for (uint64_t ii = 0; ii < n; ++ii) {
        uint64_t row = n - 1 - ii;
        T *ri = &e[row * nc];
        for (uint64_t j = 0; j < row; ++j) {
            T *rj = &e[j * nc];
            for (uint64_t k = nc / 2; k < nc; ++k)
                rj[k] -= rj[row] * ri[k];
            rj[row] = 0.0;
        }
    }
    for (uint64_t i = 0; i < n; ++i)
        for (uint64_t j = 0; j < n; ++j)
            Ainv[i * n + j] = e[i * nc + n + j];
}

/* ---------------- Chebyshev coefficients (chebyshev.inl:50-121) ---------------- */

static int cheby_coeffs(double min_eig, double max_eig, int order, double fraction,
                        std::vector<double> &c) {
    double upper = max_eig * 1.1;
    double lower = (upper - min_eig) * fraction + min_eig;
    double theta = (upper + lower) / 2.;
    double delta = (upper - lower) / 2.;
    c.assign(order, 0.0);
    double den;
    switch (order) {
    case 1:
        c[0] = 1. / theta;
        return 0;
    case 2:
        den = delta * delta - 2 * theta * theta;
        c[0] = -4 * theta / den;
        c[1] = 2 / den;
        return 0;
    case 3:
        den = 3 * (delta * delta) * theta - 4 * (theta * theta * theta);
        c[0] = (3 * delta * delta - 12 * theta * theta) / den;
        c[1] = 12 * theta / den;
        c[2] = -4 / den;
        return 0;
    case 4:
        den = pow(delta, 4) - 8 * delta * delta * theta * theta + 8 * pow(theta, 4);
        c[0] = (32 * pow(theta, 3) - 16 * delta * delta * theta) / den;
        c[1] = (8 * delta * delta - 48 * theta * theta) / den;
        c[2] = 32 * theta / den;
        c[3] = -8 / den;
        return 0;
    default:
        break;
    }
    if (order < 1 || order > 16)
        return -1;
    /* Beyond the reference (it asserts for order > 4): the same polynomial family,
     * p(t) = (1 - T_k((theta-t)/delta)/T_k(theta/delta)) / t, expanded in monomials by
     * the three-term recurrence T_{m+1}(s) = 2 s T_m(s) - T_{m-1}(s), s = (theta-t)/delta. */
    std::vector<double> Tm1(order + 2, 0.0), Tm(order + 2, 0.0), Tn(order + 2, 0.0);
    Tm1[0] = 1.0;                                 /* T0 = 1 */
    Tm[0] = theta / delta, Tm[1] = -1.0 / delta;  /* T1 = s */
    for (int m = 1; m < order; ++m) {
        std::fill(Tn.begin(), Tn.end(), 0.0);
        for (int i = 0; i <= m; ++i) {
            Tn[i] += 2 * (theta / delta) * Tm[i];
            Tn[i + 1] += 2 * (-1.0 / delta) * Tm[i];
        }
        for (int i = 0; i <= m - 1; ++i)
            Tn[i] -= Tm1[i];
        Tm1 = Tm;
        Tm = Tn;
    }
    double T0 = Tm[0]; /* T_k(theta/delta) = value at t = 0 */
    for (int i = 0; i < order; ++i)
        c[i] = -Tm[i + 1] / T0;
    return 0;
}

/* ---------------- parameters ---------------- */

struct Role {
    int method = XO_NONE;
    int max_iters = 1;
    float abs_tol = 0.f, rel_tol = 0.f;
    int conv_check = 0;
    float relax = 0.99f;
    int poly_order = 2;
    float fraction = 0.3f;
};

static Role role_of(const xo_role &r) {
    Role o;
    o.method = r.method;
    o.max_iters = r.max_iters;
    o.abs_tol = r.abs_tol;
    o.rel_tol = r.rel_tol;
    o.conv_check = r.conv_check;
    o.relax = r.relax_factor;
    o.poly_order = r.poly_order;
    o.fraction = r.spectrum_fraction;
    return o;
}

/* ---------------- solver base (src/solvers/solver.h:48-256) ---------------- */

struct Stats {
    std::vector<bool> conv;
    std::vector<double> abs_res, rel_res;
    uint32_t iters = 0;
    std::vector<double> *hist = nullptr;
};

template <typename T>
struct Solver {
    Role prm;
    Stats stats;
    int nv = 0;
    virtual ~Solver() {}
    /* solve(x, b, conv) */
    virtual void solve(Vec<T> &x, Vec<T> &b, Vec<T> &iconv) = 0;
    void solve(Vec<T> &x, Vec<T> &b) {
        Vec<T> ones = const_vec<T>(x.nv, 1.0);
        solve(x, b, ones);
    }
    void reset() { /* solver_stats::reset (solver.h:54-65) */
        stats.iters = 0;
        if (prm.conv_check) {
            stats.conv.assign(nv, false);
            stats.abs_res.assign(nv, 0.0);
            stats.rel_res.assign(nv, 0.0);
        } else if ((int)stats.conv.size() != nv) {
            stats.conv.assign(nv, false);
            stats.abs_res.assign(nv, 0.0);
            stats.rel_res.assign(nv, 0.0);
        }
    }
    bool convergence_check(const Vec<T> &res, const Vec<T> &res0, const Vec<T> &conv) {
        /* solver.h:173-203 */
        float abs_tol = prm.abs_tol, rel_tol = prm.rel_tol;
        bool flag = true;
        if (stats.hist) {
            for (int k = 0; k < nv; ++k)
                stats.hist->push_back(std::sqrt((double)res.d[k]));
        }
        for (int k = 0; k < nv; ++k) {
            if (conv.d[k]) {
                T a = std::sqrt(res.d[k]);
                T r = a / std::sqrt(res0.d[k]);
                if ((a > abs_tol) && (r > rel_tol)) {
                    flag = false;
                } else if (!stats.conv[k]) {
                    stats.abs_res[k] = a;
                    stats.rel_res[k] = r;
                    stats.conv[k] = true;
                }
            }
        }
        return flag;
    }
    bool converged(const Vec<T> &res, const Vec<T> &res0, Vec<T> &conv) { /* solver.h:205-241 */
        if (prm.conv_check) {
            if (convergence_check(res, res0, conv))
                return true;
            for (int k = 0; k < nv; ++k)
                conv.d[k] *= (!stats.conv[k]);
            conv.zero = false;
        }
        return false;
    }
};

template <typename T>
struct Level;

/* ---------------- smoothers ---------------- */

template <typename T>
struct Identity : Solver<T> {
    void solve(Vec<T> &x, Vec<T> &b, Vec<T> &) override { copy(b, x); } /* solver.h:352-368 */
};

template <typename T>
struct Jacobi : Solver<T> { /* jacobi.inl:50-120 */
    const Csr<T> *A = nullptr;
    Vec<T> invd;
    Vec<T> r;
    void init(const Csr<T> *A_) {
        A = A_;
        invd.alloc(A->nrows, 1);
        invd.d = inv_diag(*A);
        invd.zero = false;
    }
    void get_residual(const Vec<T> &x, const Vec<T> &b, Vec<T> &rr, Vec<T> *res) {
        Ax_y(*A, x, rr);
        axpby_s(1.0, b, -1.0, rr);
        if (res)
            dot(rr, rr, *res);
    }
    void solve(Vec<T> &x, Vec<T> &b, Vec<T> &iconv_in) override {
        const int nv = x.nv;
        this->nv = nv;
        this->reset();
        Vec<T> iconv = iconv_in;
        Vec<T> a1 = const_vec<T>(nv, 1.0);
        Vec<T> relax;
        relax.alloc(1, nv);
        set_const(relax, (double)(T)this->prm.relax);
        r.reuse(x.n, nv);
        auto &it = this->stats.iters;
        if (this->prm.conv_check) {
            Vec<T> rho0, res;
            rho0.alloc(1, nv);
            res.alloc(1, nv);
            get_residual(x, b, r, &rho0);
            if (this->converged(rho0, rho0, iconv))
                return;
            do {
                ++it;
                scal(invd, r);
                scal(iconv, relax);
                axpby(relax, r, a1, x);
                get_residual(x, b, r, &res);
                if (this->converged(res, rho0, iconv))
                    return;
            } while ((int)it < this->prm.max_iters);
        } else {
            do {
                ++it;
                Ax_y(*A, x, r);
                merged_jacobi(x, b, r, relax);
            } while ((int)it < this->prm.max_iters);
        }
    }
    void merged_jacobi(Vec<T> &x, Vec<T> &b, Vec<T> &rr, const Vec<T> &relax) {
        /* blas_merged_jacobi.inl:33-90 */
        const int nv = x.nv;
        if (x.zero && rr.zero) {
            XO_PAR
            for (uint64_t i = 0; i < x.n; ++i)
                for (int k = 0; k < nv; ++k)
                    x.d[i * nv + k] = relax.d[k] * invd.d[i] * b.d[i * nv + k];
        } else {
            if (x.zero)
                set_const(x, 0.0, true);
            if (rr.zero)
                set_const(rr, 0.0, true);
            XO_PAR
            for (uint64_t i = 0; i < x.n; ++i)
                for (int k = 0; k < nv; ++k)
                    x.d[i * nv + k] +=
                        relax.d[k] * invd.d[i] * (b.d[i * nv + k] - rr.d[i * nv + k]);
        }
        x.zero = false;
    }
};

template <typename T>
struct HSGS : Solver<T> { /* hsgs.inl:31-115, blas2.h:185-281, csr_gs.inl:33-134 */
    const Csr<T> *A = nullptr;
    std::vector<T> invd;
    std::vector<uint64_t> blk; /* core block boundaries */
    std::vector<uint8_t> inblk;  /* per nonzero: column lies in the row's block */
    Vec<T> r, t;
    void init(const Csr<T> *A_, int nblocks) {
        A = A_;
        invd = inv_diag(*A);
        blk = core_ranges(A->nrows, std::max(1, nblocks));
        std::vector<uint32_t> owner(A->nrows);
        for (size_t c = 0; c + 1 < blk.size(); ++c)
            for (uint64_t i = blk[c]; i < blk[c + 1]; ++i)
                owner[i] = (uint32_t)c;
        inblk.assign(A->col.size(), 0);
        for (uint64_t i = 0; i < A->nrows; ++i)
            for (uint32_t j = A->row[i]; j < A->row[i + 1]; ++j)
                inblk[j] = (owner[A->col[j]] == owner[i]);
    }
    void sgs(Vec<T> &b, Vec<T> &x, const Vec<T> &relax) {
        const int nv = x.nv;
        /* t = b - (off-block part of A) x_old; off-block contributions accumulate in the
         * order of the column blocks, which for rows stored diagonal-first and then by
         * ascending column is the stored order (csr_matrix_segment.inl:107-229) */
        set_const(t, 0.0, true);
        if (!x.zero) {
            XO_PAR
            for (uint64_t i = 0; i < A->nrows; ++i)
                for (uint32_t j = A->row[i]; j < A->row[i + 1]; ++j)
                    if (!inblk[j])
                        for (int k = 0; k < nv; ++k)
                            t.d[i * nv + k] += x.d[(uint64_t)A->col[j] * nv + k] * A->val[j];
        }
        axpby_s(1.0, b, -1.0, t);
        if (x.zero)
            set_const(x, 0.0, true);
        /* blocks are independent (each reads and writes only its own rows of x) */
        const int64_t nblk = (int64_t)blk.size() - 1;
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t c = 0; c < nblk; ++c) {
            std::vector<T> sum(nv), coeff(nv);
            const uint64_t r0 = blk[c], r1 = blk[c + 1];
            for (uint64_t i = r0; i < r1; ++i) { /* forward, csr_gs.inl:77-100 */
                for (int k = 0; k < nv; ++k)
                    sum[k] = t.d[i * nv + k];
                for (uint32_t j = A->row[i]; j < A->row[i + 1]; ++j)
                    if (inblk[j])
                        for (int k = 0; k < nv; ++k)
                            sum[k] -= x.d[(uint64_t)A->col[j] * nv + k] * A->val[j];
                for (int k = 0; k < nv; ++k)
                    coeff[k] = relax.d[k] * invd[i];
                for (int k = 0; k < nv; ++k)
                    x.d[i * nv + k] += coeff[k] * sum[k];
            }
            for (uint64_t ii = r0; ii < r1; ++ii) { /* backward, csr_gs.inl:104-127 */
                uint64_t i = r1 - 1 - (ii - r0);
                for (int k = 0; k < nv; ++k)
                    sum[k] = t.d[i * nv + k];
                for (uint32_t j = A->row[i]; j < A->row[i + 1]; ++j)
                    if (inblk[j])
                        for (int k = 0; k < nv; ++k)
                            sum[k] -= x.d[(uint64_t)A->col[j] * nv + k] * A->val[j];
                for (int k = 0; k < nv; ++k)
                    coeff[k] = relax.d[k] * invd[i];
                for (int k = 0; k < nv; ++k)
                    x.d[i * nv + k] += coeff[k] * sum[k];
            }
        }
        x.zero = false;
    }
    void get_residual(const Vec<T> &x, const Vec<T> &b, Vec<T> &rr, Vec<T> &res) {
        Ax_y(*A, x, rr);
        axpby_s(1.0, b, -1.0, rr);
        dot(rr, rr, res);
    }
    void solve(Vec<T> &x, Vec<T> &b, Vec<T> &iconv_in) override {
        const int nv = x.nv;
        this->nv = nv;
        this->reset();
        Vec<T> iconv = iconv_in;
        Vec<T> relax;
        relax.alloc(1, nv);
        set_const(relax, (double)(T)this->prm.relax);
        r.reuse(x.n, nv);
        t.reuse(x.n, nv);
        auto &it = this->stats.iters;
        Vec<T> rho0, res;
        rho0.alloc(1, nv);
        res.alloc(1, nv);
        if (this->prm.conv_check) {
            get_residual(x, b, r, rho0);
            if (this->converged(rho0, rho0, iconv))
                return;
        }
        do {
            ++it;
            scal(iconv, relax);
            sgs(b, x, relax);
            if (this->prm.conv_check) {
                get_residual(x, b, r, res);
                if (this->converged(res, rho0, iconv))
                    return;
            }
        } while ((int)it < this->prm.max_iters);
    }
};

template <typename T>
struct Chebyshev : Solver<T> { /* chebyshev.inl:124-206 */
    const Csr<T> *A = nullptr;
    Vec<T> isd;
    double min_eig = 0, max_eig = 0;
    Vec<T> r, t1, t2;
    void init(const Csr<T> *A_, double mn, double mx) {
        A = A_;
        isd.alloc(A->nrows, 1);
        isd.d = inv_sqrt_diag(*A);
        isd.zero = false;
        min_eig = mn;
        max_eig = mx;
    }
    void solve(Vec<T> &x, Vec<T> &b, Vec<T> &) override {
        const int nv = x.nv;
        this->nv = nv;
        this->reset();
        const int order = this->prm.poly_order;
        Vec<T> a1 = const_vec<T>(nv, 1.0);
        r.reuse(x.n, nv);
        t1.reuse(x.n, nv);
        t2.reuse(x.n, nv);
        std::vector<double> cd;
        if (cheby_coeffs(min_eig, max_eig, order, (double)this->prm.fraction, cd)) {
            g_err = -3;
            return;
        }
        std::vector<Vec<T>> c(order);
        for (int i = 0; i < order; ++i) {
            c[i].alloc(1, nv);
            for (int k = 0; k < nv; ++k)
                c[i].d[k] = (T)cd[i];
            c[i].zero = false;
        }
        /* get_residual(x, r, rho0, conv_check=0) */
        Ax_y(*A, x, r);
        axpby_s(1.0, b, -1.0, r);
        scal(isd, r);
        ax_y(c[order - 1], r, t1);
        for (int i = order - 2; i >= 0; i--) {
            scal(isd, t1);
            Ax_y(*A, t1, t2);
            scal(isd, t2);
            axpby_z(c[i], r, a1, t2, t1);
        }
        scal(isd, t1);
        axpby(a1, t1, a1, x);
    }
};

template <typename T>
struct Direct : Solver<T> { /* direct.inl:31-56 */
    uint64_t n = 0;
    std::vector<T> inv;
    template <typename TS>
    void init(const Csr<TS> *A) {
        n = A->nrows;
        std::vector<T> dense(n * n, T(0));
        for (uint64_t i = 0; i < n; ++i)
            for (uint32_t j = A->row[i]; j < A->row[i + 1]; ++j)
                dense[i * n + A->col[j]] = (T)A->val[j];
        inv.resize(n * n);
        gj_inverse(n, dense.data(), inv.data());
    }
    void solve(Vec<T> &x, Vec<T> &b, Vec<T> &) override {
        if (b.zero) {
            x.zero = true;
            return;
        }
        dense_ax_y(n, n, inv.data(), b.p(), x.p(), b.nv);
        x.zero = false;
    }
};

/* ---------------- MultiGrid (mg.inl) ---------------- */

struct AnySolver {
    std::unique_ptr<Solver<double>> d;
    std::unique_ptr<Solver<float>> f;
};

struct LevelData {
    bool reduced = false; /* this level's A and smoothers run in fp32 */
    Csr<double> Ad, Pd, Rd;
    Csr<float> Af, Pf, Rf;
    double min_eig = 0, max_eig = 0;
    /* buffers (mg.inl:136-163): temp (T1), temp32 (T2), b_H, x_H (this level's T) */
    Vec<double> tmp_d, bH_d, xH_d;
    Vec<float> tmp_f, bH_f, xH_f;
};

struct Params {
    Role solver, precond, pre, post, coarse;
    /* per-level smoother roles: params.get("pre_smoother", l) (mg.inl:113-131) */
    std::vector<Role> pre_lev, post_lev;
    const Role &pre_of(int l) const { return l < (int)pre_lev.size() ? pre_lev[l] : pre; }
    const Role &post_of(int l) const { return l < (int)post_lev.size() ? post_lev[l] : post; }
    int mg_cycle = 0;
    int mg_prec = 0;
    int hsgs_ncores = 1, hsgs_block_rows = 0;
    int coarse_fp64 = 0;
    int true_res_period = 0;
};

template <typename T>
static std::unique_ptr<Solver<T>> make_smoother(const Role &r, const Csr<T> *A, double mn,
                                                double mx, const Params &P) {
    std::unique_ptr<Solver<T>> s;
    switch (r.method) {
    case XO_JACOBI: {
        auto j = new Jacobi<T>();
        j->init(A);
        s.reset(j);
        break;
    }
    case XO_HSGS: {
        auto h = new HSGS<T>();
        int nb = P.hsgs_ncores;
        if (P.hsgs_block_rows > 0)
            nb = (int)std::max<uint64_t>(1, (A->nrows + P.hsgs_block_rows - 1) / P.hsgs_block_rows);
        h->init(A, nb);
        s.reset(h);
        break;
    }
    case XO_CHEBYSHEV: {
        auto c = new Chebyshev<T>();
        c->init(A, mn, mx);
        s.reset(c);
        break;
    }
    case XO_DIRECT: {
        auto d = new Direct<T>();
        d->init(A);
        s.reset(d);
        break;
    }
    case XO_IDENTITY:
        s.reset(new Identity<T>());
        break;
    default:
        g_err = -4;
        return nullptr;
    }
    s->prm = r;
    return s;
}

struct MultiGrid : Solver<double> {
    Params P;
    std::vector<LevelData> L;
    std::vector<AnySolver> pre, post;
    AnySolver coarse;
    bool coarse_as_fp64_in_f32 = false;
    Vec<double> r;      /* residual buffer of solve() */
    Vec<float> xf, bf;  /* fp32 x and b of the fp32 preconditioner (mg_prec 2) */

    void setup(const xo_hier *H, const Params &p) { /* mg.inl:97-164 */
        P = p;
        const int nl = H->nlevels;
        L.resize(nl);
        for (int l = 0; l < nl; ++l) {
            LevelData &d = L[l];
            d.reduced = (p.mg_prec == 2) || (p.mg_prec == 1 && l >= 1);
            d.min_eig = H->min_eig[l];
            d.max_eig = H->max_eig[l];
            if (d.reduced)
                d.Af = make_csr<float>(&H->A[l]);
            else
                d.Ad = make_csr<double>(&H->A[l]);
            if (l + 1 < nl) {
                /* R_l, P_l carry the precision of the coarse side (hypre_base.inl:397) */
                bool red = (p.mg_prec != 0);
                if (red) {
                    d.Pf = make_csr<float>(&H->P[l]);
                    d.Rf = make_csr<float>(&H->R[l]);
                } else {
                    d.Pd = make_csr<double>(&H->P[l]);
                    d.Rd = make_csr<double>(&H->R[l]);
                }
            }
        }
        pre.resize(nl > 0 ? nl - 1 : 0);
        post.resize(nl > 0 ? nl - 1 : 0);
        for (int l = 0; l + 1 < nl; ++l) {
            if (L[l].reduced) {
                pre[l].f = make_smoother<float>(p.pre_of(l), &L[l].Af, L[l].min_eig,
                                                L[l].max_eig, p);
                post[l].f = make_smoother<float>(p.post_of(l), &L[l].Af, L[l].min_eig,
                                                 L[l].max_eig, p);
            } else {
                pre[l].d = make_smoother<double>(p.pre_of(l), &L[l].Ad, L[l].min_eig,
                                                 L[l].max_eig, p);
                post[l].d = make_smoother<double>(p.post_of(l), &L[l].Ad, L[l].min_eig,
                                                  L[l].max_eig, p);
            }
        }
        LevelData &c = L[nl - 1];
        if (c.reduced && p.mg_prec == 2 && p.coarse_fp64) {
            /* fp32 hierarchy with an fp64 coarsest solve (beyond the reference) */
            c.Ad = make_csr<double>(&H->A[nl - 1]);
            coarse.d = make_smoother<double>(p.coarse, &c.Ad, c.min_eig, c.max_eig, p);
            coarse_as_fp64_in_f32 = true;
        } else if (c.reduced) {
            coarse.f = make_smoother<float>(p.coarse, &c.Af, c.min_eig, c.max_eig, p);
        } else {
            coarse.d = make_smoother<double>(p.coarse, &c.Ad, c.min_eig, c.max_eig, p);
        }
    }

    void alloc(int nv) {
        for (size_t l = 0; l < L.size(); ++l) {
            uint64_t n = L[l].reduced ? L[l].Af.nrows : L[l].Ad.nrows;
            LevelData &d = L[l];
            if (d.tmp_d.nv != nv || d.tmp_d.n != n) {
                d.tmp_d.alloc(n, nv);
                d.tmp_f.alloc(n, nv);
                d.bH_d.alloc(n, nv);
                d.bH_f.alloc(n, nv);
                d.xH_d.alloc(n, nv);
                d.xH_f.alloc(n, nv);
            }
        }
    }

    template <typename T>
    static Vec<T> &bufH_b(LevelData &d);
    template <typename T>
    static Vec<T> &bufH_x(LevelData &d);
    template <typename T>
    static Vec<T> &buf_tmp(LevelData &d);
    template <typename T>
    static const Csr<T> &matA(LevelData &d);
    template <typename T>
    static const Csr<T> &matP(LevelData &d);
    template <typename T>
    static const Csr<T> &matR(LevelData &d);
    template <typename T>
    static Solver<T> *get(AnySolver &s);

    template <typename T>
    void coarse_solve(Vec<T> &x, Vec<T> &b, Vec<T> &iconv) {
        if (coarse_as_fp64_in_f32) {
            Vec<double> bd, xd;
            bd.alloc(b.n, b.nv);
            xd.alloc(x.n, x.nv);
            bd.zero = b.zero;
            XO_PAR
            for (size_t i = 0; i < b.sz(); ++i)
                bd.d[i] = (double)b.d[i];
            Vec<double> ic = const_vec<double>(b.nv, 1.0);
            coarse.d->solve(xd, bd, ic);
            x.zero = xd.zero;
            XO_PAR
            for (size_t i = 0; i < x.sz(); ++i)
                x.d[i] = (T)xd.d[i];
            return;
        }
        get<T>(coarse)->solve(x, b, iconv);
    }

    /* MG_cycle<T1,T2> (mg.inl:166-260) */
    template <typename T1, typename T2>
    void cycle(Vec<T1> &x, Vec<T1> &b, int level, int mg_cycle, Vec<T1> &iconv) {
        LevelData &d = L[level];
        LevelData &dH = L[level + 1];
        Vec<T1> a1 = const_vec<T1>(x.nv, 1.0), a_1 = const_vec<T1>(x.nv, -1.0);
        Vec<T1> &temp = buf_tmp<T1>(d);
        Vec<T2> &b_H = bufH_b<T2>(dH);
        Vec<T2> &x_H = bufH_x<T2>(dH);
        Vec<T2> iconv2;
        iconv2.alloc(1, x.nv);
        iconv2.zero = iconv.zero;
        for (int k = 0; k < x.nv; ++k)
            iconv2.d[k] = (T2)iconv.d[k];

        if (mg_cycle != 2)
            get<T1>(pre[level])->solve(x, b, iconv);

        Ax_y(matA<T1>(d), x, temp);
        axpby(a1, b, a_1, temp);

        if constexpr (!std::is_same<T1, T2>::value) {
            Vec<T2> &temp32 = buf_tmp<T2>(d);
            temp32.zero = temp.zero;
            if (!temp.zero)
                XO_PAR
                for (size_t i = 0; i < temp.sz(); ++i)
                    temp32.d[i] = (T2)temp.d[i];
            Ax_y(matR<T2>(d), temp32, b_H);
        } else {
            Ax_y(matR<T2>(d), temp, b_H);
        }
        set_const(x_H, 0.0);

        if (level + 1 == (int)L.size() - 1) {
            coarse_solve<T2>(x_H, b_H, iconv2);
        } else {
            cycle<T2, T2>(x_H, b_H, level + 1, mg_cycle, iconv2);
            if (mg_cycle == 1)
                cycle<T2, T2>(x_H, b_H, level + 1, 1, iconv2);
            else if (mg_cycle == 2)
                cycle<T2, T2>(x_H, b_H, level + 1, 0, iconv2);
        }

        if constexpr (!std::is_same<T1, T2>::value) {
            Vec<T2> &temp32 = buf_tmp<T2>(d);
            Ax_y(matP<T2>(d), x_H, temp32);
            temp.zero = temp32.zero;
            if (!temp32.zero)
                XO_PAR
                for (size_t i = 0; i < temp.sz(); ++i)
                    temp.d[i] = (T1)temp32.d[i];
            axpby(a1, temp, a1, x);
        } else {
            Axpy(matP<T1>(d), x_H, x);
        }

        get<T1>(post[level])->solve(x, b, iconv);
    }

    void get_residual(const Vec<double> &x, const Vec<double> &b, Vec<double> &r,
                      Vec<double> &res) {
        Ax_y(L[0].Ad, x, r);
        axpby_s(1.0, b, -1.0, r);
        dot(r, r, res);
    }

    void solve(Vec<double> &x, Vec<double> &b, Vec<double> &iconv_in) override { /* mg.inl:262-334 */
        const int nv = x.nv;
        this->nv = nv;
        this->reset();
        alloc(nv);
        Vec<double> iconv = iconv_in;
        auto &it = this->stats.iters;
        Vec<double> rho0, res;
        rho0.alloc(1, nv);
        res.alloc(1, nv);
        r.reuse(x.n, nv);
        if (this->prm.conv_check) {
            if (P.mg_prec == 2) {
                g_err = -5; /* level-0 fp64 operator is not kept in that mode */
                return;
            }
            get_residual(x, b, r, rho0);
            if (this->converged(rho0, rho0, iconv))
                return;
        }
        do {
            ++it;
            if (L.size() == 1) {
                coarse_solve<double>(x, b, iconv);
            } else if (P.mg_prec == 2) {
                /* fp32 preconditioner including level 0 (beyond the reference): b and x are
                 * converted at the preconditioner boundary */
                Vec<float> icf;
                xf.reuse(x.n, nv);
                bf.reuse(b.n, nv);
                icf.alloc(1, nv);
                for (int k = 0; k < nv; ++k)
                    icf.d[k] = (float)iconv.d[k];
                icf.zero = iconv.zero;
                bf.zero = b.zero;
                if (!b.zero)
                    XO_PAR
                    for (size_t i = 0; i < b.sz(); ++i)
                        bf.d[i] = (float)b.d[i];
                xf.zero = x.zero;
                if (!x.zero)
                    XO_PAR
                    for (size_t i = 0; i < x.sz(); ++i)
                        xf.d[i] = (float)x.d[i];
                cycle<float, float>(xf, bf, 0, P.mg_cycle, icf);
                x.zero = xf.zero;
                if (!xf.zero)
                    XO_PAR
                    for (size_t i = 0; i < x.sz(); ++i)
                        x.d[i] = (double)xf.d[i];
            } else if (L[1].reduced) {
                cycle<double, float>(x, b, 0, P.mg_cycle, iconv);
            } else {
                cycle<double, double>(x, b, 0, P.mg_cycle, iconv);
            }
            if (this->prm.conv_check) {
                get_residual(x, b, r, res);
                if (this->converged(res, rho0, iconv))
                    return;
            }
        } while ((int)it < this->prm.max_iters);
    }
};

template <>
Vec<double> &MultiGrid::bufH_b<double>(LevelData &d) { return d.bH_d; }
template <>
Vec<float> &MultiGrid::bufH_b<float>(LevelData &d) { return d.bH_f; }
template <>
Vec<double> &MultiGrid::bufH_x<double>(LevelData &d) { return d.xH_d; }
template <>
Vec<float> &MultiGrid::bufH_x<float>(LevelData &d) { return d.xH_f; }
template <>
Vec<double> &MultiGrid::buf_tmp<double>(LevelData &d) { return d.tmp_d; }
template <>
Vec<float> &MultiGrid::buf_tmp<float>(LevelData &d) { return d.tmp_f; }
template <>
const Csr<double> &MultiGrid::matA<double>(LevelData &d) { return d.Ad; }
template <>
const Csr<float> &MultiGrid::matA<float>(LevelData &d) { return d.Af; }
template <>
const Csr<double> &MultiGrid::matP<double>(LevelData &d) { return d.Pd; }
template <>
const Csr<float> &MultiGrid::matP<float>(LevelData &d) { return d.Pf; }
template <>
const Csr<double> &MultiGrid::matR<double>(LevelData &d) { return d.Rd; }
template <>
const Csr<float> &MultiGrid::matR<float>(LevelData &d) { return d.Rf; }
template <>
Solver<double> *MultiGrid::get<double>(AnySolver &s) { return s.d.get(); }
template <>
Solver<float> *MultiGrid::get<float>(AnySolver &s) { return s.f.get(); }

/* ---------------- Krylov solvers ---------------- */

struct Krylov : Solver<double> {
    const Csr<double> *A = nullptr;
    Solver<double> *precond = nullptr;
    int true_res_period = 0;
    void get_residual(const Vec<double> &x, const Vec<double> &b, Vec<double> &r,
                      Vec<double> &res) { /* solver.h:243-256 */
        Ax_y(*A, x, r);
        axpby_s(1.0, b, -1.0, r);
        dot(r, r, res);
    }
    // Periodic residual replacement (no reference counterpart): on every k-th iteration
    // the recurrence residual r, just updated, is replaced by b - A x (the same
    // arithmetic as get_residual) and rho_new = (r, r0) is taken from it.
    void replace_residual(int it, const Vec<double> &x, const Vec<double> &b, Vec<double> &r,
                          const Vec<double> &r0, Vec<double> &rho_new) {
        if (true_res_period <= 0 || it % true_res_period != 0)
            return;
        Ax_y(*A, x, r);
        axpby_s(1.0, b, -1.0, r);
        dot(r, r0, rho_new);
    }
};

struct Scalars {
    Vec<double> rho, rho0, res, scal1, scal2, scal3, scal4, alpha, beta, omega, ac, oc, theta,
        phi, psi, delta, eta, temp;
    explicit Scalars(int nv) {
        for (Vec<double> *v : {&rho, &rho0, &res, &scal1, &scal2, &scal3, &scal4, &alpha, &beta,
                               &omega, &ac, &oc, &theta, &phi, &psi, &delta, &eta, &temp})
            v->alloc(1, nv);
    }
};

struct PBiCGStab : Krylov { /* pbicgstab.inl:44-237 */
    void solve(Vec<double> &x, Vec<double> &b, Vec<double> &conv) override {
        const int nv = x.nv;
        this->nv = nv;
        reset();
        auto &it = stats.iters;
        Vec<double> iconv = conv;
        Vec<double> a1 = const_vec<double>(nv, 1.0), a_1 = const_vec<double>(nv, -1.0);
        Scalars S(nv);
        Vec<double> r0, r, p, v, s, t, p_, s_;
        for (Vec<double> *q : {&r0, &r, &p, &v, &s, &t, &p_, &s_})
            q->alloc(x.n, nv);
        get_residual(x, b, r0, S.rho0);
        copy(r0, r);
        if (converged(S.rho0, S.rho0, iconv))
            return;
        copy(S.rho0, S.rho);
        copy(r0, p);
        do {
            ++it;
            set_const(p_, 0.0);
            precond->solve(p_, p);
            Ax_y(*A, p_, v);
            dot(v, r0, S.scal1);
            xdivy_z(S.rho, S.scal1, S.alpha);
            scal(a_1, S.alpha);
            axpby_z(a1, r, S.alpha, v, s);
            scal(a_1, S.alpha);
            set_const(s_, 0.0);
            precond->solve(s_, s);
            Ax_y(*A, s_, t);
            dot(t, s, S.scal2);
            dot(t, t, S.scal3);
            dot(s, s, S.scal4);
            xdivy_z(S.scal2, S.scal3, S.omega);
            ax_y(iconv, S.alpha, S.ac);
            ax_y(iconv, S.omega, S.oc);
            axpbypcz(S.ac, p_, S.oc, s_, a1, x);
            ax_y(S.omega, S.scal2, S.res);
            axpby(a1, S.scal4, a_1, S.res);
            if (converged(S.res, S.rho0, iconv))
                return;
            scal(a_1, S.omega);
            axpby_z(a1, s, S.omega, t, r);
            scal(a_1, S.omega);
            dot(r, r0, S.scal1);
            replace_residual((int)it, x, b, r, r0, S.scal1);
            xdivy_z(S.scal1, S.rho, S.scal2);
            xdivy_z(S.alpha, S.omega, S.scal3);
            ax_y(S.scal2, S.scal3, S.beta);
            copy(S.scal1, S.rho);
            ax_y(S.omega, S.beta, S.scal1);
            scal(a_1, S.scal1);
            axpbypcz(a1, r, S.scal1, v, S.beta, p);
            if (g_err)
                return;
        } while ((int)it < prm.max_iters);
    }
};

struct MergedPBiCGStab : Krylov { /* merged_pbicgstab.inl:44-225 */
    void solve(Vec<double> &x, Vec<double> &b, Vec<double> &conv) override {
        const int nv = x.nv;
        this->nv = nv;
        reset();
        auto &it = stats.iters;
        Vec<double> iconv = conv;
        Vec<double> a1 = const_vec<double>(nv, 1.0), a_1 = const_vec<double>(nv, -1.0);
        Scalars S(nv);
        Vec<double> r0, r, p, v, s, t, p_, s_;
        for (Vec<double> *q : {&r0, &r, &p, &v, &s, &t, &p_, &s_})
            q->alloc(x.n, nv);
        Ax_y(*A, x, r0);
        group1(r0, b, p, r, S.rho0);
        if (converged(S.rho0, S.rho0, iconv))
            return;
        copy(S.rho0, S.rho);
        do {
            ++it;
            set_const(p_, 0.0);
            precond->solve(p_, p);
            Ax_y(*A, p_, v);
            dot(v, r0, S.scal1);
            xdivy_z(S.rho, S.scal1, S.alpha);
            scal(a_1, S.alpha);
            axpby_z(a1, r, S.alpha, v, s);
            scal(a_1, S.alpha);
            set_const(s_, 0.0);
            precond->solve(s_, s);
            Ax_y(*A, s_, t);
            group2(t, s, S.phi, S.psi, S.theta);
            xdivy_z(S.phi, S.psi, S.omega);
            ax_y(S.omega, S.phi, S.res);
            axpby(a1, S.theta, a_1, S.res);
            ax_y(iconv, S.alpha, S.ac);
            ax_y(iconv, S.omega, S.oc);
            axpbypcz(S.ac, p_, S.oc, s_, a1, x);
            if (converged(S.res, S.rho0, iconv))
                return;
            group4(s, t, r0, r, S.omega, S.scal1);
            replace_residual((int)it, x, b, r, r0, S.scal1);
            xdivy_z(S.scal1, S.rho, S.scal2);
            xdivy_z(S.alpha, S.omega, S.scal3);
            ax_y(S.scal2, S.scal3, S.beta);
            copy(S.scal1, S.rho);
            ax_y(S.omega, S.beta, S.scal1);
            scal(a_1, S.scal1);
            axpbypcz(a1, r, S.scal1, v, S.beta, p);
            if (g_err)
                return;
        } while ((int)it < prm.max_iters);
    }
};

struct BiCGStab : Krylov { /* bicgstab.inl:44-221 */
    void solve(Vec<double> &x, Vec<double> &b, Vec<double> &conv) override {
        const int nv = x.nv;
        this->nv = nv;
        reset();
        auto &it = stats.iters;
        Vec<double> iconv = conv;
        Vec<double> a1 = const_vec<double>(nv, 1.0), a_1 = const_vec<double>(nv, -1.0);
        Scalars S(nv);
        Vec<double> r0, r, p, v, s, t;
        for (Vec<double> *q : {&r0, &r, &p, &v, &s, &t})
            q->alloc(x.n, nv);
        get_residual(x, b, r0, S.rho0);
        copy(r0, r);
        if (converged(S.rho0, S.rho0, iconv))
            return;
        copy(S.rho0, S.rho);
        copy(r0, p);
        do {
            ++it;
            Ax_y(*A, p, v);
            dot(v, r0, S.scal1);
            xdivy_z(S.rho, S.scal1, S.alpha);
            scal(a_1, S.alpha);
            axpby_z(a1, r, S.alpha, v, s);
            scal(a_1, S.alpha);
            Ax_y(*A, s, t);
            dot(t, s, S.scal2);
            dot(t, t, S.scal3);
            dot(s, s, S.scal4);
            xdivy_z(S.scal2, S.scal3, S.omega);
            ax_y(iconv, S.alpha, S.ac);
            ax_y(iconv, S.omega, S.oc);
            axpbypcz(S.ac, p, S.oc, s, a1, x);
            ax_y(S.omega, S.scal2, S.res);
            axpby(a1, S.scal4, a_1, S.res);
            if (converged(S.res, S.rho0, iconv))
                return;
            scal(a_1, S.omega);
            axpby_z(a1, s, S.omega, t, r);
            scal(a_1, S.omega);
            dot(r, r0, S.scal1);
            replace_residual((int)it, x, b, r, r0, S.scal1);
            xdivy_z(S.scal1, S.rho, S.scal2);
            xdivy_z(S.alpha, S.omega, S.scal3);
            ax_y(S.scal2, S.scal3, S.beta);
            copy(S.scal1, S.rho);
            ax_y(S.omega, S.beta, S.scal1);
            scal(a_1, S.scal1);
            axpbypcz(a1, r, S.scal1, v, S.beta, p);
            if (g_err)
                return;
        } while ((int)it < prm.max_iters);
    }
};

struct MergedBiCGStab : Krylov { /* merged_bicgstab.inl:44-208 */
    void solve(Vec<double> &x, Vec<double> &b, Vec<double> &conv) override {
        const int nv = x.nv;
        this->nv = nv;
        reset();
        auto &it = stats.iters;
        Vec<double> iconv = conv;
        Vec<double> a1 = const_vec<double>(nv, 1.0), a_1 = const_vec<double>(nv, -1.0);
        Scalars S(nv);
        Vec<double> r0, r, p, v, s, t;
        for (Vec<double> *q : {&r0, &r, &p, &v, &s, &t})
            q->alloc(x.n, nv);
        Ax_y(*A, x, r0);
        group1(r0, b, p, r, S.rho0);
        if (converged(S.rho0, S.rho0, iconv))
            return;
        copy(S.rho0, S.rho);
        do {
            ++it;
            Ax_y(*A, p, v);
            dot(v, r0, S.scal1);
            xdivy_z(S.rho, S.scal1, S.alpha);
            scal(a_1, S.alpha);
            axpby_z(a1, r, S.alpha, v, s);
            scal(a_1, S.alpha);
            Ax_y(*A, s, t);
            group2(t, s, S.phi, S.psi, S.theta);
            xdivy_z(S.phi, S.psi, S.omega);
            ax_y(S.omega, S.phi, S.res);
            axpby(a1, S.theta, a_1, S.res);
            ax_y(iconv, S.alpha, S.ac);
            ax_y(iconv, S.omega, S.oc);
            if (converged(S.res, S.rho0, iconv)) {
                axpbypcz(S.ac, p, S.oc, s, a1, x);
                return;
            }
            group4(s, t, r0, r, S.omega, S.scal1);
            xdivy_z(S.scal1, S.rho, S.scal2);
            xdivy_z(S.alpha, S.omega, S.scal3);
            ax_y(S.scal2, S.scal3, S.beta);
            copy(S.scal1, S.rho);
            ax_y(iconv, S.alpha, S.ac);
            ax_y(iconv, S.omega, S.oc);
            group5(x, p, s, r, v, S.ac, S.oc, S.alpha, S.omega, S.beta);
            if (g_err)
                return;
        } while ((int)it < prm.max_iters);
    }
};

struct PCG : Krylov { /* pcg.inl:44-163 */
    void solve(Vec<double> &x, Vec<double> &b, Vec<double> &conv) override {
        const int nv = x.nv;
        this->nv = nv;
        reset();
        auto &it = stats.iters;
        Vec<double> iconv = conv;
        Vec<double> a1 = const_vec<double>(nv, 1.0), a_1 = const_vec<double>(nv, -1.0);
        Scalars S(nv);
        Vec<double> r, z, p, Ap;
        for (Vec<double> *q : {&r, &z, &p, &Ap})
            q->alloc(x.n, nv);
        get_residual(x, b, r, S.rho0);
        if (converged(S.rho0, S.rho0, iconv))
            return;
        set_const(z, 0.0);
        precond->solve(z, r);
        copy(z, p);
        do {
            ++it;
            Ax_y(*A, p, Ap);
            dot(r, z, S.scal1);
            dot(p, Ap, S.scal2);
            xdivy_z(S.scal1, S.scal2, S.alpha);
            ax_y(iconv, S.alpha, S.ac);
            axpby(S.ac, p, a1, x);
            scal(a_1, S.alpha);
            axpby(S.alpha, Ap, a1, r);
            set_const(z, 0.0);
            precond->solve(z, r);
            dot(r, z, S.scal3);
            xdivy_z(S.scal3, S.scal1, S.beta);
            axpby(a1, z, S.beta, p);
            dot(r, r, S.res);
            if (converged(S.res, S.rho0, iconv))
                return;
            if (g_err)
                return;
        } while ((int)it < prm.max_iters);
    }
};

/* RBiCGStab (rbicgstab.inl:44-232) and MergedRBiCGStab (merged/merged_rbicgstab.inl:44-226):
 * the same recurrences; the merged form fuses steps 12-13 (merged_rbicgstab_group1,
 * blas_merged_rbicgstab.inl:33-118, the same roundings) and 23-25
 * (merged_rbicgstab_group2, :120-196: v_ = z + beta v_ - beta omega s, a different
 * association) and updates x after the convergence check with the updated iconv. */
struct RBiCGStab : Krylov {
    bool merged = false;
    void solve(Vec<double> &x, Vec<double> &b, Vec<double> &conv) override {
        const int nv = x.nv;
        this->nv = nv;
        reset();
        auto &it = stats.iters;
        Vec<double> iconv = conv;
        Vec<double> a1 = const_vec<double>(nv, 1.0), a_1 = const_vec<double>(nv, -1.0);
        Scalars S(nv);
        Vec<double> r0, r, z, v_, v, s, t_, t;
        for (Vec<double> *q : {&r0, &r, &z, &v_, &v, &s, &t_, &t})
            q->alloc(x.n, nv);
        get_residual(x, b, r0, S.rho0);
        copy(r0, r);
        if (converged(S.rho0, S.rho0, iconv))
            return;
        copy(S.rho0, S.rho);
        set_const(z, 0.0);
        precond->solve(z, r0);
        copy(z, v_);
        do {
            ++it;
            Ax_y(*A, v_, v);                          // 6
            dot(v, r0, S.delta);                      // 7
            set_const(s, 0.0);
            precond->solve(s, v);                     // 8
            xdivy_z(S.rho, S.delta, S.alpha);         // 9
            scal(a_1, S.alpha);
            axpby_z(a1, z, S.alpha, s, t_);           // 10
            scal(a_1, S.alpha);
            Ax_y(*A, t_, t);                          // 11
            if (merged) {                             // 12-13, group1
                const size_t N = r.sz();
                for (int k = 0; k < nv; ++k)
                    S.theta.d[k] = S.phi.d[k] = S.psi.d[k] = S.eta.d[k] = 0.0;
                for (size_t i = 0; i < N; i += nv)
                    for (int k = 0; k < nv; ++k) {
                        r.d[i + k] -= S.alpha.d[k] * v.d[i + k];
                        S.theta.d[k] += t.d[i + k] * r.d[i + k];
                        S.phi.d[k] += t.d[i + k] * t.d[i + k];
                        S.psi.d[k] += t.d[i + k] * r0.d[i + k];
                        S.eta.d[k] += r.d[i + k] * r.d[i + k];
                    }
                for (Vec<double> *q : {&S.theta, &S.phi, &S.psi, &S.eta})
                    q->zero = false;
            } else {
                scal(a_1, S.alpha);
                axpby(S.alpha, v, a1, r);             // 12
                scal(a_1, S.alpha);
                dot(t, r, S.theta);                   // 13
                dot(t, t, S.phi);
                dot(t, r0, S.psi);
                dot(r, r, S.eta);
            }
            set_const(z, 0.0);
            precond->solve(z, t);                     // 14
            xdivy_z(S.theta, S.phi, S.omega);         // 15
            scal(a_1, S.omega);
            axpby(S.omega, t, a1, r);                 // 16
            scal(a_1, S.omega);
            ax_y(S.omega, S.theta, S.res);            // 17
            axpby(a1, S.eta, a_1, S.res);
            ax_y(iconv, S.alpha, S.ac);
            ax_y(iconv, S.omega, S.oc);
            if (!merged)
                axpbypcz(S.ac, v_, S.oc, t_, a1, x);  // 23
            if (converged(S.res, S.rho0, iconv)) {
                if (merged)
                    axpbypcz(S.ac, v_, S.oc, t_, a1, x);
                return;
            }
            ax_y(S.omega, S.psi, S.rho);              // 21
            scal(a_1, S.rho);
            ax_y(S.delta, S.omega, S.temp);           // 22
            xdivy_z(S.rho, S.temp, S.beta);
            if (!merged) {
                scal(a_1, S.omega);
                axpby(a1, t_, S.omega, z);            // 24
                ax_y(S.omega, S.beta, S.delta);       // 25
                scal(a_1, S.omega);
                axpbypcz(a1, z, S.delta, s, S.beta, v_);
            } else {                                  // group2 with the updated iconv
                ax_y(iconv, S.alpha, S.ac);
                ax_y(iconv, S.omega, S.oc);
                const size_t N = x.sz();
                for (size_t i = 0; i < N; i += nv)
                    for (int k = 0; k < nv; ++k) {
                        x.d[i + k] += S.ac.d[k] * v_.d[i + k] + S.oc.d[k] * t_.d[i + k];
                        z.d[i + k] = t_.d[i + k] - S.omega.d[k] * z.d[i + k];
                        v_.d[i + k] = z.d[i + k] + S.beta.d[k] * v_.d[i + k] -
                                      S.beta.d[k] * S.omega.d[k] * s.d[i + k];
                    }
                x.zero = z.zero = v_.zero = false;
            }
            if (g_err)
                return;
        } while ((int)it < prm.max_iters);
    }
};

static Params params_of(const xo_params *p) {
    Params P;
    P.solver = role_of(p->solver);
    P.precond = role_of(p->precond);
    P.pre = role_of(p->pre);
    P.post = role_of(p->post);
    P.coarse = role_of(p->coarse);
    for (int l = 0; l < std::min(std::max(0, p->nlev_roles), XO_MAX_LEVELS); ++l) {
        P.pre_lev.push_back(role_of(p->pre_lev[l]));
        P.post_lev.push_back(role_of(p->post_lev[l]));
    }
    P.mg_cycle = p->mg_cycle;
    P.mg_prec = p->mg_prec;
    P.hsgs_ncores = std::max(1, p->hsgs_ncores);
    P.hsgs_block_rows = p->hsgs_block_rows;
    P.coarse_fp64 = p->coarse_fp64;
    P.true_res_period = p->true_res_period;
    return P;
}

/* builds the preconditioner (or a whole-solver smoother) on the finest level */
static std::unique_ptr<Solver<double>> make_precond(const xo_hier *H, const Params &P,
                                                    const Role &r, const Csr<double> *A0,
                                                    std::unique_ptr<MultiGrid> &mg_out) {
    if (r.method == XO_MULTIGRID) {
        mg_out.reset(new MultiGrid());
        mg_out->setup(H, P);
        mg_out->prm = r;
        return nullptr;
    }
    return make_smoother<double>(r, A0, H->min_eig[0], H->max_eig[0], P);
}

struct SolverSet {
    Params P;
    Csr<double> A0;
    std::unique_ptr<MultiGrid> mg;
    std::unique_ptr<Solver<double>> pre;
    std::unique_ptr<Solver<double>> top;
    Solver<double> *S = nullptr;
    int dot_ncores = 1;
};

static int build_set(const xo_hier *H, const xo_params *prm, SolverSet &Z) {
    g_err = 0;
    Z.dot_ncores = std::max(1, prm->dot_ncores);
    Z.P = params_of(prm);
    Params &P = Z.P;
    Z.A0 = make_csr<double>(&H->A[0]);
    Solver<double> *precond = nullptr;
    const int m = P.solver.method;
    if (m == XO_PBICGSTAB || m == XO_MERGED_PBICGSTAB || m == XO_PCG || m == XO_RBICGSTAB ||
        m == XO_MERGED_RBICGSTAB) {
        Z.pre = make_precond(H, P, P.precond, &Z.A0, Z.mg);
        precond = Z.mg ? static_cast<Solver<double> *>(Z.mg.get()) : Z.pre.get();
        if (!precond)
            return -4;
    }
    switch (m) {
    case XO_PBICGSTAB:
        Z.top.reset(new PBiCGStab());
        break;
    case XO_MERGED_PBICGSTAB:
        Z.top.reset(new MergedPBiCGStab());
        break;
    case XO_BICGSTAB:
        Z.top.reset(new BiCGStab());
        break;
    case XO_MERGED_BICGSTAB:
        Z.top.reset(new MergedBiCGStab());
        break;
    case XO_PCG:
        Z.top.reset(new PCG());
        break;
    case XO_RBICGSTAB:
    case XO_MERGED_RBICGSTAB: {
        auto *rb = new RBiCGStab();
        rb->merged = m == XO_MERGED_RBICGSTAB;
        Z.top.reset(rb);
        break;
    }
    case XO_MULTIGRID:
        Z.mg.reset(new MultiGrid());
        Z.mg->setup(H, P);
        Z.mg->prm = P.solver;
        break;
    default:
        Z.top = make_smoother<double>(P.solver, &Z.A0, H->min_eig[0], H->max_eig[0], P);
        if (!Z.top)
            return -4;
        break;
    }
    Z.S = Z.top ? Z.top.get() : Z.mg.get();
    Z.S->prm = P.solver;
    if (auto *k = dynamic_cast<Krylov *>(Z.S)) {
        k->A = &Z.A0;
        k->precond = precond;
        // MergedBiCGStab updates x only after r (group5), so b - A x is not r's residual
        // periodic replacement is defined for the BiCGStab family of the benchmark only
        if (P.true_res_period > 0 &&
            !(P.solver.method == XO_BICGSTAB || P.solver.method == XO_PBICGSTAB ||
              P.solver.method == XO_MERGED_PBICGSTAB))
            return -5;
        k->true_res_period = P.true_res_period;
    }
    return g_err;
}

static int run_set(SolverSet &Z, int nv, const double *b, double *x, int32_t *iters,
                   int32_t *converged, double *abs_res, double *rel_res,
                   std::vector<double> *hist) {
    g_err = 0;
    g_dot_ncores = Z.dot_ncores;
    const uint64_t n = Z.A0.nrows;
    Vec<double> xv, bv;
    xv.alloc(n, nv);
    bv.alloc(n, nv);
    std::memcpy(xv.p(), x, sizeof(double) * n * nv);
    std::memcpy(bv.p(), b, sizeof(double) * n * nv);
    xv.zero = false; /* user vectors are uploaded, never zero-flagged (blas_ext.inl:243-283) */
    bv.zero = false;
    Solver<double> *S = Z.S;
    S->stats.hist = hist;
    S->solve(xv, bv);
    S->stats.hist = nullptr;
    if (g_err)
        return g_err;
    std::memcpy(x, xv.p(), sizeof(double) * n * nv);
    if (iters)
        *iters = (int32_t)S->stats.iters;
    for (int k = 0; k < nv; ++k) {
        if (converged)
            converged[k] = k < (int)S->stats.conv.size() ? (int)S->stats.conv[k] : 0;
        if (abs_res)
            abs_res[k] = k < (int)S->stats.abs_res.size() ? S->stats.abs_res[k] : 0.0;
        if (rel_res)
            rel_res[k] = k < (int)S->stats.rel_res.size() ? S->stats.rel_res[k] : 0.0;
    }
    return 0;
}

static int run_solve(const xo_hier *H, const xo_params *prm, int nv, const double *b,
                     double *x, int32_t *iters, int32_t *converged, double *abs_res,
                     double *rel_res, std::vector<double> *hist) {
    SolverSet Z;
    int rc = build_set(H, prm, Z);
    if (rc)
        return rc;
    return run_set(Z, nv, b, x, iters, converged, abs_res, rel_res, hist);
}

} // namespace xo

using namespace xo;

extern "C" {

/* OpenMP threads of the oracle's parallel loops (the reference's cores per rank) */
int xo_set_threads(int n) {
    if (n < 1)
        return -1;
    omp_set_num_threads(n);
    return 0;
}
int xo_get_threads(void) { return omp_get_max_threads(); }

void xo_spmv_d(const xo_csr *A, int nv, const double *x, double *y, int accumulate) {
    Csr<double> m = make_csr<double>(A);
    if (accumulate)
        csr_axpy(m, x, y, nv);
    else
        csr_ax_y(m, x, y, nv);
}

void xo_spmv_f(const xo_csr *A, int nv, const float *x, float *y, int accumulate) {
    Csr<float> m = make_csr<float>(A);
    if (accumulate)
        csr_axpy(m, x, y, nv);
    else
        csr_ax_y(m, x, y, nv);
}

void xo_dense_matvec_d(int64_t n, int64_t m, const double *D, int nv, const double *x,
                       double *y) {
    dense_ax_y<double>(n, m, D, x, y, nv);
}

void xo_dense_matvec_f(int64_t n, int64_t m, const float *D, int nv, const float *x, float *y) {
    dense_ax_y<float>(n, m, D, x, y, nv);
}

void xo_inverse_d(int64_t n, const double *A, double *Ainv) { gj_inverse<double>(n, A, Ainv); }
void xo_inverse_f(int64_t n, const float *A, float *Ainv) { gj_inverse<float>(n, A, Ainv); }

int xo_cheby_coeffs(double min_eig, double max_eig, int order, double fraction, double *coeffs) {
    std::vector<double> c;
    int rc = cheby_coeffs(min_eig, max_eig, order, fraction, c);
    if (rc)
        return rc;
    for (int i = 0; i < order; ++i)
        coeffs[i] = c[i];
    return 0;
}

uint64_t xo_split_range(uint64_t size, int nblocks) { return split_range(size, nblocks); }

void xo_core_ranges(uint64_t size, int nblocks, uint64_t *offsets) {
    std::vector<uint64_t> o = core_ranges(size, nblocks);
    for (int i = 0; i <= nblocks; ++i)
        offsets[i] = o[i];
}

void xo_hsgs_d(const xo_csr *A, int nv, int nblocks, float relax, int sweeps, const double *b,
               double *x, int x_is_zero) {
    Csr<double> m = make_csr<double>(A);
    HSGS<double> h;
    h.init(&m, nblocks);
    h.prm.relax = relax;
    h.prm.max_iters = sweeps;
    h.prm.conv_check = 0;
    Vec<double> xv, bv;
    xv.alloc(m.nrows, nv);
    bv.alloc(m.nrows, nv);
    std::memcpy(xv.p(), x, sizeof(double) * m.nrows * nv);
    std::memcpy(bv.p(), b, sizeof(double) * m.nrows * nv);
    xv.zero = x_is_zero != 0;
    bv.zero = false;
    h.Solver<double>::solve(xv, bv);
    if (xv.zero)
        std::fill(xv.d.begin(), xv.d.end(), 0.0);
    std::memcpy(x, xv.p(), sizeof(double) * m.nrows * nv);
}

void xo_chebyshev_d(const xo_csr *A, int nv, double min_eig, double max_eig, int order,
                    float fraction, const double *b, double *x, int x_is_zero) {
    Csr<double> m = make_csr<double>(A);
    Chebyshev<double> c;
    c.init(&m, min_eig, max_eig);
    c.prm.poly_order = order;
    c.prm.fraction = fraction;
    Vec<double> xv, bv;
    xv.alloc(m.nrows, nv);
    bv.alloc(m.nrows, nv);
    std::memcpy(xv.p(), x, sizeof(double) * m.nrows * nv);
    std::memcpy(bv.p(), b, sizeof(double) * m.nrows * nv);
    xv.zero = x_is_zero != 0;
    bv.zero = false;
    c.Solver<double>::solve(xv, bv);
    if (xv.zero)
        std::fill(xv.d.begin(), xv.d.end(), 0.0);
    std::memcpy(x, xv.p(), sizeof(double) * m.nrows * nv);
}

void xo_jacobi_d(const xo_csr *A, int nv, float relax, int sweeps, const double *b, double *x,
                 int x_is_zero) {
    Csr<double> m = make_csr<double>(A);
    Jacobi<double> j;
    j.init(&m);
    j.prm.relax = relax;
    j.prm.max_iters = sweeps;
    j.prm.conv_check = 0;
    Vec<double> xv, bv;
    xv.alloc(m.nrows, nv);
    bv.alloc(m.nrows, nv);
    std::memcpy(xv.p(), x, sizeof(double) * m.nrows * nv);
    std::memcpy(bv.p(), b, sizeof(double) * m.nrows * nv);
    xv.zero = x_is_zero != 0;
    bv.zero = false;
    j.Solver<double>::solve(xv, bv);
    if (xv.zero)
        std::fill(xv.d.begin(), xv.d.end(), 0.0);
    std::memcpy(x, xv.p(), sizeof(double) * m.nrows * nv);
}

int xo_define_int_type(uint64_t range) { /* bit_operations.h:119-129 */
    if (range <= UINT8_MAX)
        return 0;
    if (range <= UINT16_MAX)
        return 1;
    if (range <= UINT32_MAX)
        return 2;
    return 3;
}

int xo_compress(const xo_csr *A, int64_t block_ncols, uint32_t *row_out, uint32_t *col_out,
                uint32_t *row_ind, uint32_t *col_ind, int64_t *nrows_out, int64_t *ncols_out) {
    /* csr_matrix::compress (csr_matrix.inl:86-170) */
    const int64_t n = A->nrows;
    std::vector<uint32_t> row(A->row, A->row + n + 1);
    int64_t ne = 0;
    for (int64_t i = 0; i < n; ++i)
        if (row[i + 1] - row[i])
            row_ind[ne++] = (uint32_t)i;
    std::vector<uint64_t> label(block_ncols, 0);
    for (int64_t j = 0; j < A->nnz; ++j)
        label[A->col[j]] = 1;
    int64_t nc = 0;
    for (int64_t c = 0; c < block_ncols; ++c)
        if (label[c]) {
            col_ind[nc] = (uint32_t)c;
            label[c] = (uint64_t)nc;
            ++nc;
        }
    int64_t k = 0;
    for (int64_t i = 0; i < n; ++i)
        if (row[i + 1] != row[k]) {
            row[k + 1] = row[i + 1];
            ++k;
        }
    for (int64_t i = 0; i <= k; ++i)
        row_out[i] = row[i];
    for (int64_t j = 0; j < A->nnz; ++j)
        col_out[j] = (uint32_t)label[A->col[j]];
    *nrows_out = k;
    *ncols_out = nc;
    return 0;
}

int64_t xo_cube_nnz(int64_t nx, int64_t ny, int64_t nz) {
    int64_t nnz = 0;
    for (int64_t indx = 0; indx < nx * ny * nz; ++indx) {
        int64_t ii = indx % nx, jj = (indx / nx) % ny, kk = indx / (nx * ny);
        nnz += 1 + (kk > 0) + (jj > 0) + (ii > 0) + (ii < nx - 1) + (jj < ny - 1) + (kk < nz - 1);
    }
    return nnz;
}

void xo_cube(int64_t nx, int64_t ny, int64_t nz, uint32_t *row, uint32_t *col, double *val) {
    /* examples/common/generator/cube.inl:84-146, one process */
    double hx = 1.0 / nx, hy = 1.0 / ny, hz = 1.0 / nz;
    double q1 = 1.0 / hx / hx, q2 = 1.0 / hy / hy, q3 = 1.0 / hz / hz;
    uint64_t c = 0;
    row[0] = 0;
    for (int64_t indx = 0; indx < nx * ny * nz; ++indx) {
        int64_t ii = indx % nx, jj = (indx / nx) % ny, kk = indx / (nx * ny);
        uint64_t d = c;
        col[c] = (uint32_t)indx;
        val[c++] = 0.0;
        if (kk > 0) {
            val[d] += q3;
            col[c] = (uint32_t)(indx - nx * ny);
            val[c++] = -q3;
        } else if (nz > 1)
            val[d] += 2.0 * q3;
        if (jj > 0) {
            val[d] += q2;
            col[c] = (uint32_t)(indx - nx);
            val[c++] = -q2;
        } else if (ny > 1)
            val[d] += 2.0 * q2;
        if (ii > 0) {
            val[d] += q1;
            col[c] = (uint32_t)(indx - 1);
            val[c++] = -q1;
        } else if (nx > 1)
            val[d] += 2.0 * q1;
        if (ii < nx - 1) {
            val[d] += q1;
            col[c] = (uint32_t)(indx + 1);
            val[c++] = -q1;
        } else if (nx > 1)
            val[d] += 2.0 * q1;
        if (jj < ny - 1) {
            val[d] += q2;
            col[c] = (uint32_t)(indx + nx);
            val[c++] = -q2;
        } else if (ny > 1)
            val[d] += 2.0 * q2;
        if (kk < nz - 1) {
            val[d] += q3;
            col[c] = (uint32_t)(indx + nx * ny);
            val[c++] = -q3;
        } else if (nz > 1)
            val[d] += 2.0 * q3;
        row[indx + 1] = (uint32_t)c;
    }
}

int xo_solve(const xo_hier *H, const xo_params *prm, int nv, const double *b, double *x,
             int32_t *iters, int32_t *converged, double *abs_res, double *rel_res) {
    return run_solve(H, prm, nv, b, x, iters, converged, abs_res, rel_res, nullptr);
}

int xo_solve_hist(const xo_hier *H, const xo_params *prm, int nv, const double *b, double *x,
                  int32_t *iters, int32_t *converged, double *abs_res, double *rel_res,
                  double *hist, int32_t max_hist) {
    std::vector<double> h;
    int rc = run_solve(H, prm, nv, b, x, iters, converged, abs_res, rel_res, &h);
    size_t m = std::min(h.size(), (size_t)max_hist * nv);
    for (size_t i = 0; i < m; ++i)
        hist[i] = h[i];
    return rc;
}

struct xo_solver_s {
    SolverSet Z;
};

int xo_solver_create(const xo_hier *H, const xo_params *prm, xo_solver *out) {
    auto *s = new xo_solver_s();
    int rc = build_set(H, prm, s->Z);
    if (rc) {
        delete s;
        return rc;
    }
    *out = s;
    return 0;
}

int xo_solver_solve(xo_solver S, int nv, const double *b, double *x, int32_t *iters,
                    int32_t *converged, double *abs_res, double *rel_res) {
    return run_set(S->Z, nv, b, x, iters, converged, abs_res, rel_res, nullptr);
}

void xo_solver_free(xo_solver S) { delete S; }

int xo_mg_apply(const xo_hier *H, const xo_params *prm, int nv, const double *b, double *x) {
    g_err = 0;
    g_dot_ncores = std::max(1, prm->dot_ncores);
    Params P = params_of(prm);
    MultiGrid mg;
    mg.setup(H, P);
    mg.prm = P.precond;
    const uint64_t n = H->A[0].nrows;
    Vec<double> xv, bv;
    xv.alloc(n, nv);
    bv.alloc(n, nv);
    std::memcpy(bv.p(), b, sizeof(double) * n * nv);
    bv.zero = false;
    set_const(xv, 0.0);
    mg.Solver<double>::solve(xv, bv);
    if (xv.zero)
        std::fill(xv.d.begin(), xv.d.end(), 0.0);
    std::memcpy(x, xv.p(), sizeof(double) * n * nv);
    return g_err;
}

} /* extern "C" */
