// Level-1 / merged BiCGStab vector kernels with fused per-RHS dot partials, the
// single-workgroup finalize kernels that turn partials into the Krylov scalars and the
// convergence decision on the device, and the smoother / coarse-solve kernels.
//
// References: blas::axpby_z / axpbypcz / xdivy_z (src/blas/detail/blas_ext.inl:31-239),
// merged_bicgstab_group1/2/4/5 (src/blas/detail/specific/blas_merged_bicgstab.inl:33-257),
// merged_rbicgstab_group2 (src/blas/detail/specific/blas_merged_rbicgstab.inl:120-196),
// base_solver::convergence_check / converged (src/solvers/solver.h:173-241),
// merged_jacobi (src/blas/detail/specific/blas_merged_jacobi.inl:33-90),
// csr_SGS (src/blas2/detail/csr_gs.inl:33-134), dense_Ax_y
// (src/blas2/detail/dense_matvec.inl:72-121), vector::convert
// (src/primitives/vector/operations/convert.inl:31-83).
//
// Every expression keeps the reference's operand order; the file is built with
// -ffp-contract=off so a*b+c rounds twice as in the reference's x86-64 build.
#include "xh_internal.hpp"
#include "xh_krylov.hpp"

#include <algorithm>

namespace xh {
namespace {

__device__ __forceinline__ int64_t logical_block() {
    const uint32_t G = gridDim.x, p = blockIdx.x;
    return (int64_t)(p & 7u) * (int64_t)(G >> 3) + (int64_t)(p >> 3);
}

template <int NV, int ND>
__device__ __forceinline__ void block_partials(double (&p)[ND], double *__restrict__ out,
                                               int64_t lb) {
    __shared__ double red[kBlock / 64][ND][NV];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int off = NV; off < 64; off <<= 1)
            p[d] = p[d] + __shfl_xor(p[d], off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane < NV)
#pragma unroll
        for (int d = 0; d < ND; ++d)
            red[w][d][lane] = p[d];
    __syncthreads();
    if ((int)threadIdx.x < ND * NV) {
        const int d = threadIdx.x / NV, k = threadIdx.x % NV;
        double s = red[0][d][k];
#pragma unroll
        for (int q = 1; q < kBlock / 64; ++q)
            s = s + red[q][d][k];
        out[lb * (ND * NV) + d * NV + k] = s;
    }
}

enum Gate : int { GATE_NOT_DONE = 0, GATE_CONV_ALL = 1 };

template <int NV, int OP>
__global__ __launch_bounds__(kBlock) void k_vec(int64_t n, int64_t chunk, double *__restrict__ x,
                                                double *__restrict__ p, double *__restrict__ s,
                                                double *__restrict__ r,
                                                const double *__restrict__ r0,
                                                const double *__restrict__ v,
                                                const double *__restrict__ t,
                                                const double *__restrict__ p_,
                                                const double *__restrict__ s_,
                                                const SolverState *__restrict__ st, int gate,
                                                double *__restrict__ partials) {
    if (st->done)
        return;
    if (gate == GATE_CONV_ALL && !st->conv_all)
        return;
    constexpr int ND =
        (OP == V_XR_UPDATE || OP == V_R_UPDATE || OP == V_DOT1 || OP == V_COPY_DOT) ? 1
        : (OP == V_DOT2)                                                             ? 2
        : (OP == V_DOT3)                                                             ? 3
                                                                                     : 0;
    constexpr int RPB = kBlock / NV;
    const int64_t lb = logical_block();
    const int64_t start = lb * chunk;
    const int64_t end = start + chunk < n ? start + chunk : n;
    const int tr = threadIdx.x / NV, k = threadIdx.x % NV;
    double part[ND > 0 ? ND : 1];
#pragma unroll
    for (int d = 0; d < (ND > 0 ? ND : 1); ++d)
        part[d] = 0.0;
    double c1 = 0, c2 = 0, c3 = 0;
    if constexpr (OP == V_S_UPDATE) {
        c1 = st->nalpha[k];
    } else if constexpr (OP == V_XR_UPDATE) {
        c1 = st->ac[k];
        c2 = st->oc[k];
        c3 = st->omega[k];
    } else if constexpr (OP == V_X_UPDATE) {
        c1 = st->ac[k];
        c2 = st->oc[k];
    } else if constexpr (OP == V_R_UPDATE) {
        c3 = st->omega[k];
    } else if constexpr (OP == V_P_UPDATE) {
        c1 = st->pcoef[k];
        c2 = st->beta[k];
    } else if constexpr (OP == V_GROUP5) {
        c1 = st->ac[k];
        c2 = st->oc[k];
        c3 = st->beta[k] * st->omega[k];
    } else if constexpr (OP == V_CG_XR) {
        c1 = st->ac[k];
        c2 = st->nalpha[k];
    } else if constexpr (OP == V_CG_P) {
        c1 = st->beta[k];
    } else if constexpr (OP == V_RB_UPDATE) {
        c1 = st->ac[k];
        c2 = st->oc[k];
        c3 = st->nomega[k];
    } else if constexpr (OP == V_RB_UPDATE_MERGED) {
        c1 = st->ac[k];
        c2 = st->oc[k];
        c3 = st->nomega[k];
    }
    // RBiCGStab step 25 coefficients: (-omega)*beta (ax_y(omega, beta, delta) with omega
    // negated) for the plain form, beta*omega for merged_rbicgstab_group2
    const double be = (OP == V_RB_UPDATE || OP == V_RB_UPDATE_MERGED) ? st->beta[k] : 0.0;
    const double d2 = (OP == V_RB_UPDATE) ? st->nomega[k] * be : 0.0;
    const double om = (OP == V_RB_UPDATE_MERGED) ? st->omega[k] : 0.0;
    const double bo = (OP == V_RB_UPDATE_MERGED) ? be * om : 0.0;
    // unrolled so that the loads of several rows are in flight together; each thread
    // still visits its rows, and adds its dot terms, in row order
#pragma unroll 4
    for (int64_t i = start + tr; i < end; i += RPB) {
        const size_t o = (size_t)i * NV + k;
        if constexpr (OP == V_S_UPDATE) {
            // axpby_z(a1, r, -alpha, v, s): s = 1*r + (-alpha)*v
            s[o] = r[o] + c1 * v[o];
        } else if constexpr (OP == V_XR_UPDATE) {
            // axpbypcz(ac, p_, oc, s_, a1, x) then merged_bicgstab_group4
            x[o] = c1 * p_[o] + c2 * s_[o] + x[o];
            const double rv = s[o] - c3 * t[o];
            r[o] = rv;
            part[0] = part[0] + rv * r0[o];
        } else if constexpr (OP == V_X_UPDATE) {
            x[o] = c1 * p_[o] + c2 * s_[o] + x[o];
        } else if constexpr (OP == V_R_UPDATE) {
            const double rv = s[o] - c3 * t[o];
            r[o] = rv;
            part[0] = part[0] + rv * r0[o];
        } else if constexpr (OP == V_P_UPDATE) {
            // axpbypcz(a1, r, -(omega*beta), v, beta, p)
            p[o] = r[o] + c1 * v[o] + c2 * p[o];
        } else if constexpr (OP == V_GROUP5) {
            x[o] += c1 * p[o] + c2 * s[o];
            p[o] = r[o] + st->beta[k] * p[o] - c3 * v[o];
        } else if constexpr (OP == V_DOT3) {
            const double tv = t[o], sv = s[o];
            part[0] = part[0] + tv * sv;
            part[1] = part[1] + tv * tv;
            part[2] = part[2] + sv * sv;
        } else if constexpr (OP == V_DOT1) {
            part[0] = part[0] + r[o] * r0[o];
        } else if constexpr (OP == V_CG_XR) {
            // axpby(alpha_conv, p, a1, x) ; axpby(-alpha, Ap, a1, r)
            x[o] = c1 * p_[o] + x[o];
            r[o] = c2 * t[o] + r[o];
        } else if constexpr (OP == V_DOT2) {
            const double rv = r[o], zv = r0[o];
            part[0] = part[0] + rv * zv;
            part[1] = part[1] + rv * rv;
        } else if constexpr (OP == V_CG_P) {
            p[o] = v[o] + c1 * p[o]; // axpby(a1, z, beta, p)
        } else if constexpr (OP == V_COPY_DOT) {
            const double zv = v[o];
            p[o] = zv;
            part[0] = part[0] + r[o] * zv;
        } else if constexpr (OP == V_COPY) {
            p[o] = v[o];
        } else if constexpr (OP == V_RB_UPDATE) {
            r[o] = c3 * t[o] + r[o];                  // 16: axpby(-omega, t, a1, r)
            x[o] = c1 * p[o] + c2 * p_[o] + x[o];     // 23: axpbypcz(ac, v_, oc, t_, a1, x)
            const double zn = p_[o] + c3 * s[o];      // 24: axpby(a1, t_, -omega, z)
            s[o] = zn;
            p[o] = zn + d2 * v[o] + be * p[o];        // 25: axpbypcz(a1, z, delta, s, beta, v_)
        } else if constexpr (OP == V_RB_UPDATE_MERGED) {
            r[o] = c3 * t[o] + r[o];                  // 16
            x[o] += c1 * p[o] + c2 * p_[o];           // group2: x += ac v_ + oc t_
            const double zn = p_[o] - om * s[o];      //         z = t_ - omega z
            s[o] = zn;
            p[o] = zn + be * p[o] - bo * v[o];        //         v_ = z + beta v_ - beta omega s
        }
    }
    if constexpr (ND > 0)
        block_partials<NV, ND>(part, partials, lb);
}

template <int NV>
int vec_nv(VecOp op, int64_t n, const VecArgs &a, int gate, hipStream_t s) {
    const int rpb = kBlock / NV;
    const bool dots = (op == V_XR_UPDATE || op == V_R_UPDATE || op == V_DOT1 || op == V_DOT3 ||
                       op == V_DOT2 || op == V_COPY_DOT);
    int64_t need = (n + rpb - 1) / rpb;
    int64_t g = std::min<int64_t>(need, dots ? g_tune.vec_dot_groups : 8192);
    g = std::max<int64_t>(8, (g + 7) / 8 * 8);
    int64_t chunk = (n + g - 1) / g;
    chunk = (chunk + rpb - 1) / rpb * rpb;
#define XH_VEC_CASE(OPV)                                                                        \
    case OPV:                                                                                   \
        hipLaunchKernelGGL((k_vec<NV, OPV>), dim3(g), dim3(kBlock), 0, s, n, chunk,             \
                           (double *)a.x, (double *)a.p, (double *)a.s, (double *)a.r,          \
                           (const double *)a.r0, (const double *)a.v, (const double *)a.t,      \
                           (const double *)a.p_, (const double *)a.s_, a.st, gate, a.partials); \
        break;
    switch (op) {
        XH_VEC_CASE(V_S_UPDATE)
        XH_VEC_CASE(V_XR_UPDATE)
        XH_VEC_CASE(V_P_UPDATE)
        XH_VEC_CASE(V_X_UPDATE)
        XH_VEC_CASE(V_R_UPDATE)
        XH_VEC_CASE(V_GROUP5)
        XH_VEC_CASE(V_DOT3)
        XH_VEC_CASE(V_DOT1)
        XH_VEC_CASE(V_CG_XR)
        XH_VEC_CASE(V_DOT2)
        XH_VEC_CASE(V_CG_P)
        XH_VEC_CASE(V_COPY_DOT)
        XH_VEC_CASE(V_COPY)
        XH_VEC_CASE(V_RB_UPDATE)
        XH_VEC_CASE(V_RB_UPDATE_MERGED)
    default:
        return -1;
    }
#undef XH_VEC_CASE
    return dots ? (int)g : 0;
}

// ---- finalize: partial sums -> Krylov scalars and the convergence decision ----

__global__ __launch_bounds__(256) void k_finalize(int op, SolverState *__restrict__ st,
                                                  const double *__restrict__ partials,
                                                  int nparts, int nv) {
    finalize_body(op, st, partials, nparts, nv, fin_ndots(op) * nv);
}

// ---- smoother / coarse / conversion kernels ----

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_cheb_init_zero(int64_t count, int nv,
                                                           const VT *__restrict__ b,
                                                           const VT *__restrict__ isd, VT c,
                                                           VT *__restrict__ r,
                                                           VT *__restrict__ u, int last,
                                                           const int *__restrict__ done) {
    if (done != nullptr && *done)
        return;
    // get_residual with a zero-flagged x gives r = 1*b; r *= isd; t1 = c*r; t1 *= isd
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * kBlock) {
        const VT di = isd[e / nv];
        const VT rv = b[e] * di;
        VT t1 = c * rv;
        t1 = t1 * di;
        if (!last)
            r[e] = rv;
        u[e] = t1;
    }
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_jacobi(int64_t count, int nv, VT *__restrict__ x,
                                                   const VT *__restrict__ b,
                                                   const VT *__restrict__ r,
                                                   const VT *__restrict__ invd, VT relax,
                                                   int x_zero, const int *__restrict__ done) {
    if (done != nullptr && *done)
        return;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * kBlock) {
        const VT d = invd[e / nv];
        if (x_zero)
            x[e] = relax * d * b[e];
        else
            x[e] += relax * d * (b[e] - r[e]);
    }
}

// HSGS: one group of NV lanes sweeps one core block forward then backward; the block's
// x lives in LDS, each lane only touches its own RHS column so no barrier is needed.
template <typename VT, typename IT, int NV>
__global__ __launch_bounds__(64) void k_hsgs_lds(const uint32_t *__restrict__ row,
                                                 const IT *__restrict__ col,
                                                 const VT *__restrict__ val,
                                                 const uint32_t *__restrict__ pbase, int shift,
                                                 const VT *__restrict__ invd, int64_t n,
                                                 int64_t B, const VT *__restrict__ t,
                                                 VT *__restrict__ x, int x_zero, VT relax,
                                                 const int *__restrict__ done) {
    if (done != nullptr && *done)
        return;
    extern __shared__ unsigned char smem[];
    constexpr int G = 64 / NV;
    const int g = threadIdx.x / NV, k = threadIdx.x % NV;
    VT *xs = (VT *)smem + (size_t)g * B * NV;
    const int64_t blk = (int64_t)blockIdx.x * G + g;
    const int64_t lo = blk * B;
    if (lo >= n)
        return;
    const int64_t hi = lo + B < n ? lo + B : n;
    for (int64_t i = lo; i < hi; ++i)
        xs[(i - lo) * NV + k] = x_zero ? VT(0) : x[(size_t)i * NV + k];
    for (int sweep = 0; sweep < 2; ++sweep) {
        for (int64_t q = 0; q < hi - lo; ++q) {
            const int64_t i = sweep == 0 ? lo + q : hi - 1 - q;
            uint32_t base = 0;
            if constexpr (sizeof(IT) == 2)
                base = pbase[i >> shift];
            VT sum = t[(size_t)i * NV + k];
            // unrolled so that the row's column and value loads are in flight together;
            // the subtractions keep the stored order
#pragma unroll 8
            for (uint32_t j = row[i]; j < row[i + 1]; ++j) {
                const int64_t c = (int64_t)(base + (uint32_t)col[j]);
                if (c >= lo && c < hi)
                    sum = sum - xs[(c - lo) * NV + k] * val[j];
            }
            const VT coeff = relax * invd[i];
            xs[(i - lo) * NV + k] = xs[(i - lo) * NV + k] + coeff * sum;
        }
    }
    for (int64_t i = lo; i < hi; ++i)
        x[(size_t)i * NV + k] = xs[(i - lo) * NV + k];
}

// Copies count elements from global memory into LDS with the lanes lane, lane + nl, ...;
// U loads are issued before their LDS writes so that a copy costs count / (nl * U) memory
// round trips instead of one per element.  16-byte vectors where both ends are aligned.
// Loads and stores past the end re-read and re-write the last element (always in
// bounds, same value), so that no load can be moved behind a branch and the batch stays
// in flight together.
template <typename VT, int U>
__device__ inline void stage_copy(VT *__restrict__ dst, const VT *__restrict__ src,
                                  int64_t count, int lane, int nl) {
    int64_t done_el = 0;
    if ((((uintptr_t)src | (uintptr_t)dst) & 15) == 0) {
        const int64_t nvec = count * (int64_t)sizeof(VT) / 16;
        const uint4 *s4 = (const uint4 *)src;
        uint4 *d4 = (uint4 *)dst;
        for (int64_t b = lane; b < nvec; b += (int64_t)nl * U) {
            uint4 r[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t i = b + (int64_t)u * nl;
                r[u] = s4[i < nvec ? i : nvec - 1];
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int64_t i = b + (int64_t)u * nl;
                d4[i < nvec ? i : nvec - 1] = r[u];
            }
        }
        done_el = nvec * 16 / (int64_t)sizeof(VT);
    }
    for (int64_t b = done_el + lane; b < count; b += (int64_t)nl * U) {
        VT r[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = b + (int64_t)u * nl;
            r[u] = src[i < count ? i : count - 1];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = b + (int64_t)u * nl;
            dst[i < count ? i : count - 1] = r[u];
        }
    }
}

// The same sweeps on a core-block matrix (only entries inside the row's block, u32
// columns; each row padded to whole batches of KB entries, pads pointing at the block's
// end with value +0).  A group first stages its block's x, row pointers, local columns,
// values and inverse diagonal in LDS (loads issued in batches), with one more x row of
// zeros at local row nr for the pads, then sweeps from LDS.  In the sweep, t of the next R
// rows is loaded while the current R rows are swept, and the next row's row pointers,
// inverse diagonal and first KB columns and values are read from LDS while the current
// row computes; the x operands of a row are read after the previous row's x was written
// (LDS accesses of one wave are ordered).  Each lane reads and writes only its own RHS
// column of x during the sweep.  Entry order and arithmetic as k_hsgs_lds: the products
// x_c * a_ic are rounded, then subtracted from t one by one in stored order; a pad
// subtracts +0 * +0 = +0, which changes no value.  ONE: every row is one batch.
template <typename VT, int NV, int KB, bool ONE>
__global__ __launch_bounds__(64) void k_hsgs_core(const uint32_t *__restrict__ row,
                                                  const uint32_t *__restrict__ col,
                                                  const VT *__restrict__ val,
                                                  const VT *__restrict__ invd, int64_t n,
                                                  int64_t B, int64_t E, size_t gbytes,
                                                  const VT *__restrict__ t,
                                                  VT *__restrict__ x, int x_zero, VT relax,
                                                  const int *__restrict__ done) {
    if (done != nullptr && *done)
        return;
    constexpr int R = 16;
    extern __shared__ unsigned char smem[];
    constexpr int G = 64 / NV;
    const int g = threadIdx.x / NV, k = threadIdx.x % NV;
    // per group: x [B + 1][NV], values [E], inverse diagonal [B] (VT); local columns [E],
    // row pointers [B + 1] (u32)
    VT *__restrict__ xs = (VT *)(smem + (size_t)g * gbytes);
    VT *__restrict__ cv = xs + (B + 1) * NV;
    VT *__restrict__ ivs = cv + E;
    uint32_t *__restrict__ cc = (uint32_t *)(ivs + B);
    uint32_t *__restrict__ rp = cc + E;
    const int64_t lo = ((int64_t)blockIdx.x * G + g) * B;
    const bool active = lo < n;
    const int nr = active ? (int)(lo + B < n ? B : n - lo) : 0;
    const uint32_t e0 = active ? row[lo] : 0u;
    const uint32_t ne = active ? row[lo + nr] - e0 : 0u;
    if (active) {
        for (int b = k; b <= nr; b += NV * 8) {
            uint32_t r[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                r[u] = row[lo + (b + u * NV <= nr ? b + u * NV : nr)];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                rp[b + u * NV <= nr ? b + u * NV : nr] = r[u] - e0;
        }
        for (uint32_t b = k; b < ne; b += NV * 8u) {
            uint32_t c[8];
            VT v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t i = b + u * NV < ne ? b + u * NV : ne - 1;
                c[u] = col[e0 + i];
                v[u] = val[e0 + i];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t i = b + u * NV < ne ? b + u * NV : ne - 1;
                cc[i] = c[u] - (uint32_t)lo;
                cv[i] = v[u];
            }
        }
        stage_copy<VT, 8>(ivs, invd + lo, nr, k, NV);
        if (x_zero)
            for (int q = 0; q < nr; ++q)
                xs[q * NV + k] = VT(0);
        else
            stage_copy<VT, 12>(xs, x + (size_t)lo * NV, (int64_t)nr * NV, k, NV);
        xs[nr * NV + k] = VT(0);
    }
    __syncthreads();
    const VT *__restrict__ tb = t + (size_t)lo * NV + k;
    for (int sweep = 0; sweep < 2 && nr > 0; ++sweep) {
        // step q of the sweep works on local row r0 + dir * q
        const int dir = sweep == 0 ? 1 : -1, r0 = sweep == 0 ? 0 : nr - 1;
        // the same, with steps past the end clamped to the last one
        auto row_of = [&](int q) { return r0 + dir * (q < nr ? q : nr - 1); };
        VT tc[R], tn[R];
#pragma unroll
        for (int u = 0; u < R; ++u)
            tc[u] = tb[row_of(u) * NV];
        int lc = r0;
        uint32_t j0 = rp[lc], j1 = rp[lc + 1];
        VT di = ivs[lc];
        uint32_t c4[KB];
        VT v4[KB];
#pragma unroll
        for (int v = 0; v < KB; ++v) {
            c4[v] = cc[j0 + v];
            v4[v] = cv[j0 + v];
        }
        // sweeps the row lc with t value tv, then moves to row ln
        auto step = [&](const VT tv, const int ln) {
            VT sum = tv;
            VT p[KB];
#pragma unroll
            for (int v = 0; v < KB; ++v)
                p[v] = xs[c4[v] * NV + k] * v4[v];
            const VT xo = xs[lc * NV + k];
            // the next row's descriptor: independent of x
            const uint32_t nj0 = rp[ln], nj1 = rp[ln + 1];
            const VT ndi = ivs[ln];
            uint32_t nc4[KB];
            VT nv4[KB];
#pragma unroll
            for (int v = 0; v < KB; ++v) {
                nc4[v] = cc[nj0 + v];
                nv4[v] = cv[nj0 + v];
            }
#pragma unroll
            for (int v = 0; v < KB; ++v)
                sum = sum - p[v];
            // rows of more than one batch (ONE: no row has more than KB entries)
            for (uint32_t j = j0 + KB; !ONE && j < j1; j += KB) {
                VT pp[KB];
#pragma unroll
                for (int v = 0; v < KB; ++v)
                    pp[v] = xs[cc[j + v] * NV + k] * cv[j + v];
#pragma unroll
                for (int v = 0; v < KB; ++v)
                    sum = sum - pp[v];
            }
            const VT coeff = relax * di;
            xs[lc * NV + k] = xo + coeff * sum;
            lc = ln;
            j0 = nj0;
            j1 = nj1;
            di = ndi;
#pragma unroll
            for (int v = 0; v < KB; ++v) {
                c4[v] = nc4[v];
                v4[v] = nv4[v];
            }
        };
        for (int q0 = 0; q0 < nr; q0 += R) {
            // t of the next R steps, in flight while these R rows are swept
            if (q0 + 2 * R <= nr) {
#pragma unroll
                for (int u = 0; u < R; ++u)
                    tn[u] = tb[(r0 + dir * (q0 + R + u)) * NV];
            } else {
#pragma unroll
                for (int u = 0; u < R; ++u)
                    tn[u] = tb[row_of(q0 + R + u) * NV];
            }
            if (q0 + R <= nr) {
#pragma unroll
                for (int u = 0; u < R; ++u)
                    step(tc[u], u + 1 < R ? r0 + dir * (q0 + u + 1) : row_of(q0 + u + 1));
            } else {
#pragma unroll
                for (int u = 0; u < R; ++u)
                    if (q0 + u < nr)
                        step(tc[u], row_of(q0 + u + 1));
            }
#pragma unroll
            for (int u = 0; u < R; ++u)
                tc[u] = tn[u];
        }
    }
    for (int q = 0; q < nr; ++q)
        x[(size_t)(lo + q) * NV + k] = xs[q * NV + k];
}

// same sweep order on x in global memory, for blocks that do not fit in LDS
template <typename VT, typename IT, int NV>
__global__ __launch_bounds__(64) void k_hsgs_glb(const uint32_t *__restrict__ row,
                                                 const IT *__restrict__ col,
                                                 const VT *__restrict__ val,
                                                 const uint32_t *__restrict__ pbase, int shift,
                                                 const VT *__restrict__ invd, int64_t n,
                                                 int64_t B, const VT *__restrict__ t, VT *x,
                                                 int x_zero, VT relax,
                                                 const int *__restrict__ done) {
    if (done != nullptr && *done)
        return;
    constexpr int G = 64 / NV;
    const int g = threadIdx.x / NV, k = threadIdx.x % NV;
    const int64_t blk = (int64_t)blockIdx.x * G + g;
    const int64_t lo = blk * B;
    if (lo >= n)
        return;
    const int64_t hi = lo + B < n ? lo + B : n;
    if (x_zero)
        for (int64_t i = lo; i < hi; ++i)
            x[(size_t)i * NV + k] = VT(0);
    for (int sweep = 0; sweep < 2; ++sweep) {
        for (int64_t q = 0; q < hi - lo; ++q) {
            const int64_t i = sweep == 0 ? lo + q : hi - 1 - q;
            uint32_t base = 0;
            if constexpr (sizeof(IT) == 2)
                base = pbase[i >> shift];
            VT sum = t[(size_t)i * NV + k];
#pragma unroll 8
            for (uint32_t j = row[i]; j < row[i + 1]; ++j) {
                const int64_t c = (int64_t)(base + (uint32_t)col[j]);
                if (c >= lo && c < hi)
                    sum = sum - x[(size_t)c * NV + k] * val[j];
            }
            const VT coeff = relax * invd[i];
            x[(size_t)i * NV + k] = x[(size_t)i * NV + k] + coeff * sum;
        }
    }
}

template <typename VT>
__global__ __launch_bounds__(kBlock) void k_dense(int64_t n, int nv, const VT *__restrict__ D,
                                                  const VT *__restrict__ x,
                                                  VT *__restrict__ y,
                                                  const int *__restrict__ done) {
    if (done != nullptr && *done)
        return;
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= n * nv)
        return;
    const int64_t i = e / nv;
    const int k = (int)(e % nv);
    VT acc = VT(0);
    // j order as dense_Ax_y; unrolled so that many loads are in flight
#pragma unroll 16
    for (int64_t j = 0; j < n; ++j)
        acc = acc + D[i * n + j] * x[j * nv + k];
    y[e] = acc;
}

template <typename T1, typename T2>
__global__ __launch_bounds__(kBlock) void k_convert(int64_t count, const T1 *__restrict__ a,
                                                    T2 *__restrict__ b,
                                                    const int *__restrict__ done) {
    if (done != nullptr && *done)
        return;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * kBlock)
        b[e] = (T2)a[e];
}

// x_c[c][k] = x_full[col_ind[c] + off][k]: the column decode of a compressed block
// (col_ind[col[j]] + block_col_offset, csr_matrix.h:290-297) applied to the operand
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather_cols(int64_t count, int nv,
                                                        const uint32_t *__restrict__ col_ind,
                                                        int64_t off, const T *__restrict__ xf,
                                                        T *__restrict__ xc) {
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * kBlock) {
        const int64_t c = e / nv, k = e - c * nv;
        xc[e] = xf[((int64_t)col_ind[c] + off) * nv + k];
    }
}

// b[p][k] = sum over the bumps of RHS k of a exp(-|x_p - c|^2 / (2 w^2)), cells within the
// box of half-width ceil(4w) around c only (cell centres at integer + 0.5)
__global__ __launch_bounds__(kBlock) void k_bumps(int nx, int ny, int nz, int nv, int maxb,
                                                  const int *__restrict__ count,
                                                  const double *__restrict__ P,
                                                  double *__restrict__ b) {
    const int64_t n = (int64_t)nx * ny * nz * nv;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n;
         e += (int64_t)gridDim.x * kBlock) {
        const int k = (int)(e % nv);
        const int64_t p = e / nv;
        const int x = (int)(p % nx), y = (int)((p / nx) % ny), z = (int)(p / ((int64_t)nx * ny));
        double s = 0.0;
        for (int q = 0; q < count[k]; ++q) {
            const double *Q = P + ((size_t)k * maxb + q) * 5;
            const double w = Q[3];
            const int64_t R = (int64_t)ceil(4.0 * w);
            if (x < (int64_t)Q[0] - R || x > (int64_t)Q[0] + R || y < (int64_t)Q[1] - R ||
                y > (int64_t)Q[1] + R || z < (int64_t)Q[2] - R || z > (int64_t)Q[2] + R)
                continue;
            const double dx = (double)x + 0.5 - Q[0], dy = (double)y + 0.5 - Q[1],
                         dz = (double)z + 0.5 - Q[2];
            s += Q[4] * exp(-(dx * dx + dy * dy + dz * dz) * (1.0 / (2.0 * w * w)));
        }
        b[e] = s;
    }
}

// x = u + x (the axpby(a1, t1, a1, x) that ends a Chebyshev application,
// chebyshev.inl:201-202)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_add(int64_t count, const T *__restrict__ u,
                                                T *__restrict__ x, const int *__restrict__ done) {
    if (done != nullptr && *done)
        return;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * kBlock)
        x[e] = u[e] + x[e];
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill0(int64_t count, T *__restrict__ b,
                                                  const int *__restrict__ done) {
    if (done != nullptr && *done)
        return;
    for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < count;
         e += (int64_t)gridDim.x * kBlock)
        b[e] = T(0);
}

int grid_for(int64_t count) {
    int64_t g = (count + kBlock - 1) / kBlock;
    g = std::min<int64_t>(g, 16384);
    return (int)std::max<int64_t>(1, g);
}

template <typename VT, typename IT, int NV>
const char *hsgs_t(const Mat &A, const void *t, void *x, int x_zero, double relax, int64_t B,
                   const int *done, hipStream_t s) {
    constexpr int G = 64 / NV;
    const int64_t nblk = (A.nrows + B - 1) / B;
    const int grid = (int)((nblk + G - 1) / G);
    const size_t lds = (size_t)64 * B * sizeof(VT);
    if constexpr (sizeof(IT) == 4) {
        // a padded core-block matrix of this block size whose blocks fit in LDS: staged
        // sweeps.  Per group: x with its zero row, values and inverse diagonal (VT), then
        // columns and row pointers (u32)
        const size_t gb = ((size_t)(B + 1) * NV * sizeof(VT) +
                           (size_t)A.core_emax * (sizeof(VT) + 4) + (size_t)B * sizeof(VT) +
                           (size_t)(B + 1) * 4 + 15) / 16 * 16;
        if (A.core_B == B && (A.core_kb == 3 || A.core_kb == 4) && G * gb <= 64 * 1024) {
            const bool one = A.core_rowmax <= A.core_kb;
            auto kern = A.core_kb == 3 ? (one ? k_hsgs_core<VT, NV, 3, true> : k_hsgs_core<VT, NV, 3, false>)
                                       : (one ? k_hsgs_core<VT, NV, 4, true> : k_hsgs_core<VT, NV, 4, false>);
            hipLaunchKernelGGL(kern, dim3(grid), dim3(64), G * gb, s, A.row,
                               (const uint32_t *)A.col, (const VT *)A.val,
                               (const VT *)A.inv_diag, A.nrows, B, A.core_emax, gb,
                               (const VT *)t, (VT *)x, x_zero, (VT)relax, done);
            return "k_hsgs_core";
        }
    }
    if (lds <= 64 * 1024) {
        hipLaunchKernelGGL((k_hsgs_lds<VT, IT, NV>), dim3(grid), dim3(64), lds, s, A.row,
                           (const IT *)A.col, (const VT *)A.val, A.panel_base, A.panel_shift,
                           (const VT *)A.inv_diag, A.nrows, B, (const VT *)t, (VT *)x, x_zero,
                           (VT)relax, done);
        return "k_hsgs_lds";
    }
    hipLaunchKernelGGL((k_hsgs_glb<VT, IT, NV>), dim3(grid), dim3(64), 0, s, A.row,
                       (const IT *)A.col, (const VT *)A.val, A.panel_base, A.panel_shift,
                       (const VT *)A.inv_diag, A.nrows, B, (const VT *)t, (VT *)x, x_zero,
                       (VT)relax, done);
    return "k_hsgs_glb";
}

template <typename VT, typename IT>
const char *hsgs_nv(const Mat &A, int nv, const void *t, void *x, int x_zero, double relax, int64_t B,
             const int *done, hipStream_t s) {
    switch (nv) {
    case 1:
        return hsgs_t<VT, IT, 1>(A, t, x, x_zero, relax, B, done, s);
    case 2:
        return hsgs_t<VT, IT, 2>(A, t, x, x_zero, relax, B, done, s);
    case 4:
        return hsgs_t<VT, IT, 4>(A, t, x, x_zero, relax, B, done, s);
    case 8:
        return hsgs_t<VT, IT, 8>(A, t, x, x_zero, relax, B, done, s);
    case 16:
        return hsgs_t<VT, IT, 16>(A, t, x, x_zero, relax, B, done, s);
    case 32:
        return hsgs_t<VT, IT, 32>(A, t, x, x_zero, relax, B, done, s);
    case 64:
        return hsgs_t<VT, IT, 64>(A, t, x, x_zero, relax, B, done, s);
    default:
        return nullptr;
    }
}

} // namespace

namespace {

// vectors of n x nv doubles each vector op reads plus writes, and its dot count
int vec_streams(VecOp op) {
    static const int rw[] = {3, 8, 4, 5, 4, 4, 7, 2, 2, 6, 2, 3, 3, 2, 11, 11};
    return (op >= 0 && op <= V_RB_UPDATE_MERGED) ? rw[op] : 0;
}
int vec_ndots(VecOp op) {
    return (op == V_XR_UPDATE || op == V_R_UPDATE || op == V_DOT1 || op == V_COPY_DOT) ? 1
           : op == V_DOT2                                                             ? 2
           : op == V_DOT3                                                             ? 3
                                                                                      : 0;
}
const char *vec_name(VecOp op) {
    static const char *names[] = {"V_S_UPDATE", "V_XR_UPDATE", "V_P_UPDATE", "V_GROUP1",
                                  "V_X_UPDATE", "V_R_UPDATE", "V_GROUP5", "V_DOT3", "V_DOT1",
                                  "V_CG_XR", "V_DOT2", "V_CG_P", "V_COPY_DOT", "V_COPY",
                                  "V_RB_UPDATE", "V_RB_UPDATE_MERGED"};
    return (op >= 0 && op <= V_RB_UPDATE_MERGED) ? names[op] : "V_?";
}

int vec_dispatch(VecOp op, int64_t n, int nv, const VecArgs &a, hipStream_t s) {
    const int gate = a.gate;
    const VecArgs &b = a;
    switch (nv) {
    case 1:
        return vec_nv<1>(op, n, b, gate, s);
    case 2:
        return vec_nv<2>(op, n, b, gate, s);
    case 4:
        return vec_nv<4>(op, n, b, gate, s);
    case 8:
        return vec_nv<8>(op, n, b, gate, s);
    case 16:
        return vec_nv<16>(op, n, b, gate, s);
    case 32:
        return vec_nv<32>(op, n, b, gate, s);
    case 64:
        return vec_nv<64>(op, n, b, gate, s);
    default:
        return -1;
    }
}

} // namespace

int vec_op(VecOp op, int64_t n, int nv, const VecArgs &a, hipStream_t s) {
    LedgerScope led(s);
    const int np = vec_dispatch(op, n, nv, a, s);
    if (led.on) {
        led.kernel = std::string("k_vec<") + std::to_string(nv) + "," + vec_name(op) + ">";
        led.bytes = (double)vec_streams(op) * n * nv * 8.0 +
                    (np > 0 ? (double)np * vec_ndots(op) * nv * 8.0 : 0.0);
    }
    return np;
}

void finalize(FinOp op, SolverState *st, const double *partials, int nparts, int nv,
              hipStream_t s) {
    LedgerScope led(s);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(256), 0, s, (int)op, st, partials, nparts, nv);
    if (led.on) {
        led.kernel = "k_finalize";
        led.bytes = (double)nparts * fin_ndots(op) * nv * 8.0;
    }
}

void cheb_init_zero(int64_t n, int nv, int prec, const void *b, const void *isd, double c,
                    void *r, void *u_or_x, int last, const int *done, hipStream_t s) {
    const int64_t count = n * nv;
    LedgerScope led(s);
    if (led.on) {
        const double sv = prec == F64 ? 8.0 : 4.0;
        led.kernel = std::string("k_cheb_init_zero<") + prec_name(prec) + ">";
        led.bytes = sv * (double)(count * (last ? 2 : 3) + n); // b, isd ; u or x (and r)
    }
    if (prec == F64)
        hipLaunchKernelGGL(k_cheb_init_zero<double>, dim3(grid_for(count)), dim3(kBlock), 0, s,
                           count, nv, (const double *)b, (const double *)isd, c, (double *)r,
                           (double *)u_or_x, last, done);
    else
        hipLaunchKernelGGL(k_cheb_init_zero<float>, dim3(grid_for(count)), dim3(kBlock), 0, s,
                           count, nv, (const float *)b, (const float *)isd, (float)c,
                           (float *)r, (float *)u_or_x, last, done);
}

void jacobi_update(int64_t n, int nv, int prec, void *x, const void *b, const void *r,
                   const void *inv_diag, double relax, int x_zero, const int *done,
                   hipStream_t s) {
    const int64_t count = n * nv;
    LedgerScope led(s);
    if (led.on) {
        const double sv = prec == F64 ? 8.0 : 4.0;
        led.kernel = std::string("k_jacobi<") + prec_name(prec) + ">";
        // b, inverse diagonal ; x written (x and r read unless x is zero-flagged)
        led.bytes = sv * (double)(count * (x_zero ? 2 : 4) + n);
    }
    if (prec == F64)
        hipLaunchKernelGGL(k_jacobi<double>, dim3(grid_for(count)), dim3(kBlock), 0, s, count,
                           nv, (double *)x, (const double *)b, (const double *)r,
                           (const double *)inv_diag, relax, x_zero, done);
    else
        hipLaunchKernelGGL(k_jacobi<float>, dim3(grid_for(count)), dim3(kBlock), 0, s, count,
                           nv, (float *)x, (const float *)b, (const float *)r,
                           (const float *)inv_diag, (float)relax, x_zero, done);
}

int64_t hsgs_block_rows(int64_t n, int nblocks) {
    // misc::split_range<float64_t> (src/misc/misc.h:46-55)
    int64_t cs = n / nblocks;
    cs += 8 - (cs % 8);
    return cs;
}

void hsgs_sweeps(const Mat &A, int nv, const void *t, void *x, int x_zero, double relax,
                 int64_t B, const int *done, hipStream_t s) {
    LedgerScope led(s);
    const char *name;
    if (A.prec == F64)
        name = A.idx == IDX_U16 ? hsgs_nv<double, uint16_t>(A, nv, t, x, x_zero, relax, B, done, s)
                                : hsgs_nv<double, uint32_t>(A, nv, t, x, x_zero, relax, B, done, s);
    else
        name = A.idx == IDX_U16 ? hsgs_nv<float, uint16_t>(A, nv, t, x, x_zero, relax, B, done, s)
                                : hsgs_nv<float, uint32_t>(A, nv, t, x, x_zero, relax, B, done, s);
    if (led.on && name) {
        // the swept matrix (core blocks or the whole operator), its inverse diagonal, t ;
        // x written (and read unless zero-flagged)
        const double sv = A.prec == F64 ? 8.0 : 4.0, si = A.idx == IDX_U16 ? 2.0 : 4.0;
        const double V = (double)A.nrows * nv * sv;
        led.kernel = std::string(name) + "<" + prec_name(A.prec) + "," + std::to_string(nv) + ">";
        led.bytes = (double)A.nnz * (sv + si) + (double)(A.nrows + 1) * 4.0 +
                    (double)A.nrows * sv + V + (x_zero ? V : 2.0 * V);
    }
}

void dense_matvec(int64_t n, int nv, int prec, const void *D, const void *x, void *y,
                  const int *done, hipStream_t s) {
    const int64_t count = n * nv;
    const int grid = (int)((count + kBlock - 1) / kBlock);
    LedgerScope led(s);
    if (led.on) {
        const double sv = prec == F64 ? 8.0 : 4.0;
        led.kernel = std::string("k_dense<") + prec_name(prec) + ">";
        led.bytes = sv * (double)(n * n + 2 * count); // D, x, y
    }
    if (prec == F64)
        hipLaunchKernelGGL(k_dense<double>, dim3(grid), dim3(kBlock), 0, s, n, nv,
                           (const double *)D, (const double *)x, (double *)y, done);
    else
        hipLaunchKernelGGL(k_dense<float>, dim3(grid), dim3(kBlock), 0, s, n, nv,
                           (const float *)D, (const float *)x, (float *)y, done);
}

void convert(int64_t count, int from_prec, const void *src, void *dst, const int *done,
             hipStream_t s) {
    LedgerScope led(s);
    if (led.on) {
        led.kernel = from_prec == F64 ? "k_convert<double,float>" : "k_convert<float,double>";
        led.bytes = 12.0 * (double)count;
    }
    if (from_prec == F64)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(count)), dim3(kBlock), 0, s,
                           count, (const double *)src, (float *)dst, done);
    else
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(count)), dim3(kBlock), 0, s,
                           count, (const float *)src, (double *)dst, done);
}

void add_vec(int64_t count, int prec, const void *u, void *x, const int *done, hipStream_t s) {
    LedgerScope led(s);
    if (led.on) {
        led.kernel = std::string("k_add<") + prec_name(prec) + ">";
        led.bytes = (prec == F64 ? 24.0 : 12.0) * (double)count;
    }
    if (prec == F64)
        hipLaunchKernelGGL(k_add<double>, dim3(grid_for(count)), dim3(kBlock), 0, s, count,
                           (const double *)u, (double *)x, done);
    else
        hipLaunchKernelGGL(k_add<float>, dim3(grid_for(count)), dim3(kBlock), 0, s, count,
                           (const float *)u, (float *)x, done);
}

void copy_vec(int64_t count, int prec, const void *src, void *dst, const int *done,
              hipStream_t s) {
    LedgerScope led(s);
    if (led.on) {
        led.kernel = std::string("k_convert<") + prec_name(prec) + "," + prec_name(prec) + ">";
        led.bytes = (prec == F64 ? 16.0 : 8.0) * (double)count;
    }
    if (prec == F64)
        hipLaunchKernelGGL((k_convert<double, double>), dim3(grid_for(count)), dim3(kBlock), 0,
                           s, count, (const double *)src, (double *)dst, done);
    else
        hipLaunchKernelGGL((k_convert<float, float>), dim3(grid_for(count)), dim3(kBlock), 0, s,
                           count, (const float *)src, (float *)dst, done);
}

void gather_cols(int64_t nc, int nv, int prec, const uint32_t *col_ind, int64_t off,
                 const void *x_full, void *x_c, hipStream_t s) {
    const int64_t count = nc * nv;
    if (count == 0)
        return;
    if (prec == F64)
        hipLaunchKernelGGL(k_gather_cols<double>, dim3(grid_for(count)), dim3(kBlock), 0, s,
                           count, nv, col_ind, off, (const double *)x_full, (double *)x_c);
    else
        hipLaunchKernelGGL(k_gather_cols<float>, dim3(grid_for(count)), dim3(kBlock), 0, s,
                           count, nv, col_ind, off, (const float *)x_full, (float *)x_c);
}

void eval_bumps(int nx, int ny, int nz, int nv, int maxb, const int *count, const double *P,
                double *b, hipStream_t s) {
    const int64_t n = (int64_t)nx * ny * nz * nv;
    hipLaunchKernelGGL(k_bumps, dim3(grid_for(n)), dim3(kBlock), 0, s, nx, ny, nz, nv, maxb,
                       count, P, b);
}

void fill_zero(int64_t count, int prec, void *dst, const int *done, hipStream_t s) {
    LedgerScope led(s);
    if (led.on) {
        led.kernel = std::string("k_fill0<") + prec_name(prec) + ">";
        led.bytes = (prec == F64 ? 8.0 : 4.0) * (double)count;
    }
    if (prec == F64)
        hipLaunchKernelGGL(k_fill0<double>, dim3(grid_for(count)), dim3(kBlock), 0, s, count,
                           (double *)dst, done);
    else
        hipLaunchKernelGGL(k_fill0<float>, dim3(grid_for(count)), dim3(kBlock), 0, s, count,
                           (float *)dst, done);
}

} // namespace xh
