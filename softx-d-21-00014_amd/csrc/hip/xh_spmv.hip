// Multi-RHS CSR SpMV (SpMM) for gfx950 with fused epilogues.
//
// Restates backend::csr_Ax_y / csr_Axpy (src/blas2/detail/csr_matvec.inl:33-260) on
// NV-interleaved vectors x[i*NV+k]: one group of NV lanes per row, lane k owns RHS k,
// so every nonzero reads one contiguous NV-wide row of x (a coalesced 8*NV-byte load
// in fp64) and the row sum is accumulated in stored nonzero order exactly as the
// reference (acc = acc + x*val, no FMA contraction; the file is built with
// -ffp-contract=off).  Column indices are u32 or u16 panel-relative (col = base[row >>
// shift] + u16), decoded in-kernel.  The epilogues fuse the level-1 work that follows
// an SpMV in the reference solvers (residual, dot partials, Chebyshev steps,
// mixed-precision prolongation, HSGS off-block residual).
//
// Workgroups are dealt round-robin over the 8 XCDs; logical_block() renumbers them
// so that each XCD walks one contiguous eighth of the rows, keeping the +-1 and +-nx
// neighbours of x in that XCD's L2.
#include "xh_internal.hpp"

#include <algorithm>

namespace xh {
namespace {

__device__ __forceinline__ int64_t logical_block() {
    const uint32_t G = gridDim.x, p = blockIdx.x;
    return (int64_t)(p & 7u) * (int64_t)(G >> 3) + (int64_t)(p >> 3);
}

template <typename IT>
__device__ __forceinline__ uint32_t panel_base(const uint32_t *__restrict__ pb, int shift,
                                               int64_t i) {
    if constexpr (sizeof(IT) == 4)
        return 0u;
    else
        return pb[i >> shift];
}

// acc + sum_j x[col_j]*val_j over row i, stored order, 4-way load batching
template <typename VT, typename XT, typename IT, int NV>
__device__ __forceinline__ VT row_acc(const uint32_t *__restrict__ row,
                                      const IT *__restrict__ col, const VT *__restrict__ val,
                                      uint32_t base, const XT *__restrict__ x, int64_t i, int k,
                                      VT acc) {
    uint32_t j = row[i];
    const uint32_t e = row[i + 1];
    for (; j + 4 <= e; j += 4) {
        const uint32_t c0 = base + (uint32_t)col[j], c1 = base + (uint32_t)col[j + 1];
        const uint32_t c2 = base + (uint32_t)col[j + 2], c3 = base + (uint32_t)col[j + 3];
        const VT v0 = val[j], v1 = val[j + 1], v2 = val[j + 2], v3 = val[j + 3];
        const VT x0 = (VT)x[(size_t)c0 * NV + k], x1 = (VT)x[(size_t)c1 * NV + k];
        const VT x2 = (VT)x[(size_t)c2 * NV + k], x3 = (VT)x[(size_t)c3 * NV + k];
        acc = acc + x0 * v0;
        acc = acc + x1 * v1;
        acc = acc + x2 * v2;
        acc = acc + x3 * v3;
    }
    for (; j < e; ++j) {
        const uint32_t c = base + (uint32_t)col[j];
        acc = acc + (VT)x[(size_t)c * NV + k] * val[j];
    }
    return acc;
}

// sum over the lanes of a wave that own the same RHS, then over the 4 waves; one
// partial per (dot, RHS) per workgroup, in a fixed order (deterministic)
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

template <int OP>
struct NDots {
    static constexpr int value = (OP == OP_DOT1 || OP == OP_RESID0) ? 1 : (OP == OP_DOT3 ? 3 : 0);
};

template <typename VT, typename IT, int NV, int OP>
__global__ __launch_bounds__(kBlock) void k_spmv(
    const uint32_t *__restrict__ row, const IT *__restrict__ col, const VT *__restrict__ val,
    const uint32_t *__restrict__ pbase, int shift, int64_t nrows, int64_t chunk,
    const VT *__restrict__ x, void *__restrict__ yv, const VT *__restrict__ b,
    const VT *__restrict__ o1, const VT *__restrict__ isd, const VT *__restrict__ rr,
    VT *__restrict__ out2, VT *__restrict__ out3, void *__restrict__ xupdv, VT c, int last,
    int x_zero, int out_f32, double *__restrict__ partials, const int *__restrict__ done,
    int64_t blk_rows) {
    constexpr int ND = NDots<OP>::value;
    if (done != nullptr && *done)
        return;
    constexpr int RPB = kBlock / NV;
    const int64_t lb = logical_block();
    const int64_t start = lb * chunk;
    const int64_t end = start + chunk < nrows ? start + chunk : nrows;
    const int tr = threadIdx.x / NV, k = threadIdx.x % NV;
    double part[ND > 0 ? ND : 1];
#pragma unroll
    for (int d = 0; d < (ND > 0 ? ND : 1); ++d)
        part[d] = 0.0;

    for (int64_t i = start + tr; i < end; i += RPB) {
        const size_t o = (size_t)i * NV + k;
        const uint32_t base = panel_base<IT>(pbase, shift, i);
        if constexpr (OP == OP_SET) {
            VT acc = row_acc<VT, VT, IT, NV>(row, col, val, base, x, i, k, VT(0));
            ((VT *)yv)[o] = acc;
        } else if constexpr (OP == OP_ADD) {
            VT *y = (VT *)yv;
            y[o] = row_acc<VT, VT, IT, NV>(row, col, val, base, x, i, k, y[o]);
        } else if constexpr (OP == OP_RESID) {
            VT acc = row_acc<VT, VT, IT, NV>(row, col, val, base, x, i, k, VT(0));
            VT t = b[o] - acc;
            if (out_f32)
                ((float *)yv)[o] = (float)t;
            else
                ((VT *)yv)[o] = t;
        } else if constexpr (OP == OP_RESID0) {
            VT acc = row_acc<VT, VT, IT, NV>(row, col, val, base, x, i, k, VT(0));
            VT t = b[o] - acc;
            ((VT *)yv)[o] = t;
            out2[o] = t;
            out3[o] = t;
            part[0] = part[0] + (double)(t * t);
        } else if constexpr (OP == OP_DOT1) {
            VT acc = row_acc<VT, VT, IT, NV>(row, col, val, base, x, i, k, VT(0));
            ((VT *)yv)[o] = acc;
            part[0] = part[0] + (double)(acc * o1[o]);
        } else if constexpr (OP == OP_DOT3) {
            VT acc = row_acc<VT, VT, IT, NV>(row, col, val, base, x, i, k, VT(0));
            ((VT *)yv)[o] = acc;
            const VT sv = o1[o];
            part[0] = part[0] + (double)(acc * sv);
            part[1] = part[1] + (double)(acc * acc);
            part[2] = part[2] + (double)(sv * sv);
        } else if constexpr (OP == OP_CHEB0) {
            // r = b - A x ; r *= isd ; t1 = c*r ; then t1 *= isd (u) or x += t1*isd
            VT acc = row_acc<VT, VT, IT, NV>(row, col, val, base, x, i, k, VT(0));
            const VT di = isd[i];
            VT rv = b[o] - acc;
            rv = rv * di;
            VT t1 = c * rv;
            t1 = t1 * di;
            if (last) {
                VT *xu = (VT *)xupdv;
                xu[o] = x_zero ? t1 : t1 + xu[o];
            } else {
                out2[o] = rv;
                ((VT *)yv)[o] = t1;
            }
        } else if constexpr (OP == OP_CHEB) {
            // t2 = (A u)*isd ; t1 = c*r + t2 ; then u' = t1*isd or x (+)= t1*isd
            VT acc = row_acc<VT, VT, IT, NV>(row, col, val, base, x, i, k, VT(0));
            const VT di = isd[i];
            VT t2 = acc * di;
            VT t1 = c * rr[o] + t2;
            t1 = t1 * di;
            if (last) {
                VT *xu = (VT *)xupdv;
                xu[o] = x_zero ? t1 : t1 + xu[o];
            } else {
                ((VT *)yv)[o] = t1;
            }
        } else if constexpr (OP == OP_PROLONG32) {
            // VT = float: temp32 = P xH ; x64 = (double)temp32 + x64
            VT acc = row_acc<VT, VT, IT, NV>(row, col, val, base, x, i, k, VT(0));
            double *xd = (double *)yv;
            xd[o] = x_zero ? (double)acc : (double)acc + xd[o];
        } else if constexpr (OP == OP_OFFBLOCK) {
            // t = b - sum over off-block columns (old x), HSGS core blocks of blk_rows
            const int64_t blk = i / blk_rows;
            const int64_t lo = blk * blk_rows, hi = lo + blk_rows;
            VT acc = VT(0);
            for (uint32_t j = row[i]; j < row[i + 1]; ++j) {
                const int64_t cc = (int64_t)(base + (uint32_t)col[j]);
                if (cc < lo || cc >= hi)
                    acc = acc + x[(size_t)cc * NV + k] * val[j];
            }
            ((VT *)yv)[o] = b[o] - acc;
        }
    }
    if constexpr (ND > 0)
        block_partials<NV, ND>(part, partials, lb);
}

struct Launch {
    int grid;
    int64_t chunk;
};

Launch plan(int64_t nrows, int nv, int cap) {
    const int rpb = kBlock / nv;
    int64_t need = (nrows + rpb - 1) / rpb;
    int64_t g = std::min<int64_t>(need, cap);
    g = std::max<int64_t>(8, (g + 7) / 8 * 8);
    int64_t chunk = (nrows + g - 1) / g;
    chunk = (chunk + rpb - 1) / rpb * rpb;
    return {(int)g, chunk};
}

template <typename VT, typename IT, int NV, int OP>
int launch_t(const Mat &A, const void *x, void *y, const SpmvArgs &a, hipStream_t s) {
    constexpr int ND = NDots<OP>::value;
    const int cap = ND > 0 ? kMaxPartials : (a.grid_cap > 0 ? a.grid_cap : 16384);
    Launch L = plan(A.nrows, NV, cap);
    hipLaunchKernelGGL((k_spmv<VT, IT, NV, OP>), dim3(L.grid), dim3(kBlock), 0, s, A.row,
                       (const IT *)A.col, (const VT *)A.val, A.panel_base, A.panel_shift,
                       A.nrows, L.chunk, (const VT *)x, y, (const VT *)a.b, (const VT *)a.o1,
                       (const VT *)a.isd, (const VT *)a.r, (VT *)a.out2, (VT *)a.out3, a.xupd,
                       (VT)a.c, a.last, a.x_zero, a.out_f32, a.partials, a.done, a.blk_rows);
    return ND > 0 ? L.grid : 0;
}

template <typename VT, typename IT, int NV>
int by_op(const Mat &A, const void *x, void *y, SpmvOp op, const SpmvArgs &a, hipStream_t s) {
    switch (op) {
    case OP_SET:
        return launch_t<VT, IT, NV, OP_SET>(A, x, y, a, s);
    case OP_ADD:
        return launch_t<VT, IT, NV, OP_ADD>(A, x, y, a, s);
    case OP_RESID:
        return launch_t<VT, IT, NV, OP_RESID>(A, x, y, a, s);
    case OP_CHEB0:
        return launch_t<VT, IT, NV, OP_CHEB0>(A, x, y, a, s);
    case OP_CHEB:
        return launch_t<VT, IT, NV, OP_CHEB>(A, x, y, a, s);
    case OP_OFFBLOCK:
        return launch_t<VT, IT, NV, OP_OFFBLOCK>(A, x, y, a, s);
    default:
        break;
    }
    if constexpr (sizeof(VT) == 8) {
        switch (op) {
        case OP_DOT1:
            return launch_t<VT, IT, NV, OP_DOT1>(A, x, y, a, s);
        case OP_DOT3:
            return launch_t<VT, IT, NV, OP_DOT3>(A, x, y, a, s);
        case OP_RESID0:
            return launch_t<VT, IT, NV, OP_RESID0>(A, x, y, a, s);
        default:
            break;
        }
    } else {
        if (op == OP_PROLONG32)
            return launch_t<VT, IT, NV, OP_PROLONG32>(A, x, y, a, s);
    }
    return -1;
}

template <typename VT, typename IT>
int by_nv(const Mat &A, int nv, const void *x, void *y, SpmvOp op, const SpmvArgs &a,
          hipStream_t s) {
    switch (nv) {
    case 1:
        return by_op<VT, IT, 1>(A, x, y, op, a, s);
    case 2:
        return by_op<VT, IT, 2>(A, x, y, op, a, s);
    case 4:
        return by_op<VT, IT, 4>(A, x, y, op, a, s);
    case 8:
        return by_op<VT, IT, 8>(A, x, y, op, a, s);
    case 16:
        return by_op<VT, IT, 16>(A, x, y, op, a, s);
    case 32:
        return by_op<VT, IT, 32>(A, x, y, op, a, s);
    case 64:
        return by_op<VT, IT, 64>(A, x, y, op, a, s);
    default:
        return -1;
    }
}

} // namespace

int spmv(const Mat &A, int nv, const void *x, void *y, SpmvOp op, const SpmvArgs &a,
         hipStream_t s) {
    if (A.nrows == 0)
        return 0;
    if (A.prec == F64)
        return A.idx == IDX_U16 ? by_nv<double, uint16_t>(A, nv, x, y, op, a, s)
                                : by_nv<double, uint32_t>(A, nv, x, y, op, a, s);
    return A.idx == IDX_U16 ? by_nv<float, uint16_t>(A, nv, x, y, op, a, s)
                            : by_nv<float, uint32_t>(A, nv, x, y, op, a, s);
}

} // namespace xh
