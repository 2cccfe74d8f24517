// Multi-RHS CSR SpMV (SpMM) for gfx950 with fused epilogues.
//
// Restates backend::csr_Ax_y / csr_Axpy (src/blas2/detail/csr_matvec.inl:33-260) on
// NV-interleaved vectors x[i*NV+k].  A row is owned by L = NV/W lanes and each lane owns
// W consecutive right-hand sides (W = 2 in fp64, 4 in fp32: one 16-byte load per lane),
// so every nonzero reads one contiguous NV-wide row of x with 16-byte loads.  Each RHS
// is accumulated in stored nonzero order exactly as the reference
// (acc = acc + x*val, no FMA contraction: the file is built with -ffp-contract=off).
// Column indices are u32 or u16 panel-relative (col = base[row >> shift] + u16), decoded
// in-kernel.  The epilogues fuse the level-1 work that follows an SpMV in the reference
// solvers (residual, dot partials, Chebyshev steps, mixed-precision prolongation, HSGS
// off-block residual).
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

template <typename T, int W>
struct alignas(sizeof(T) * W) Pack {
    T v[W];
};

template <typename T, int W>
__device__ __forceinline__ Pack<T, W> ld(const T *__restrict__ p) {
    return *reinterpret_cast<const Pack<T, W> *>(p);
}

template <typename T, int W>
__device__ __forceinline__ void st(T *__restrict__ p, const Pack<T, W> &v) {
    *reinterpret_cast<Pack<T, W> *>(p) = v;
}

template <typename IT>
__device__ __forceinline__ uint32_t panel_base(const uint32_t *__restrict__ pb, int shift,
                                               int64_t i) {
    if constexpr (sizeof(IT) == 4)
        return 0u;
    else
        return pb[i >> shift];
}

// acc[w] + sum_j x[col_j][k0+w]*val_j over the nonzeros [j, e) of one row, stored order.
// Nonzeros go in predicated batches of 8: all column/value loads of a batch are issued
// together, then all x loads, so a row of up to 8 nonzeros costs two dependent memory
// round trips.  Only existing entries are added (no +0 terms), so each RHS gets the
// reference's sequential sum exactly.
template <typename VT, typename IT, int NV, int W>
__device__ __forceinline__ void row_acc(uint32_t j, const uint32_t e, const IT *__restrict__ col,
                                        const VT *__restrict__ val, uint32_t base,
                                        const VT *__restrict__ x, int k0, VT (&acc)[W]) {
    constexpr int BW = 8;
    constexpr int L = NV / W; // lanes sharing this row
    if constexpr (L >= 2) {
        // The L lanes of the row load its column indices and values cooperatively
        // (E entries per lane, one coalesced load each) and broadcast them across the
        // group with cross-lane shuffles, instead of every lane loading every entry.
        constexpr int E = (BW + L - 1) / L;
        constexpr int CH = L * E; // entries per chunk (>= 8)
        const int lane = threadIdx.x & 63;
        const int g0 = lane - (lane % L);
        const int gl = lane - g0;
        for (; j < e; j += CH) {
            uint32_t cl[E];
            VT vl[E];
#pragma unroll
            for (int q = 0; q < E; ++q) {
                const uint32_t jj = j + q * L + gl;
                const bool ok = jj < e;
                cl[q] = ok ? base + (uint32_t)col[jj] : 0u;
                vl[q] = ok ? val[jj] : VT(0);
            }
#pragma unroll
            for (int b0 = 0; b0 < CH; b0 += BW) {
                // uniform within the row's lane group, so every shuffle source is active
                if (j + b0 >= e)
                    break;
                uint32_t c[BW];
                VT v[BW];
#pragma unroll
                for (int q = 0; q < BW; ++q) {
                    const int t = b0 + q; // entry index within the chunk
                    c[q] = (uint32_t)__shfl((int)cl[t / L], g0 + (t % L), 64);
                    v[q] = __shfl(vl[t / L], g0 + (t % L), 64);
                }
                Pack<VT, W> xv[BW];
#pragma unroll
                for (int q = 0; q < BW; ++q)
                    if (j + b0 + q < e)
                        xv[q] = ld<VT, W>(x + (size_t)c[q] * NV + k0);
#pragma unroll
                for (int q = 0; q < BW; ++q)
                    if (j + b0 + q < e)
#pragma unroll
                        for (int w = 0; w < W; ++w)
                            acc[w] = acc[w] + xv[q].v[w] * v[q];
            }
        }
        return;
    }
    for (; j < e; j += BW) {
        uint32_t c[BW];
        VT v[BW];
#pragma unroll
        for (int q = 0; q < BW; ++q) {
            const bool ok = j + q < e;
            c[q] = ok ? base + (uint32_t)col[j + q] : 0u;
            v[q] = ok ? val[j + q] : VT(0);
        }
        Pack<VT, W> xv[BW];
#pragma unroll
        for (int q = 0; q < BW; ++q)
            if (j + q < e)
                xv[q] = ld<VT, W>(x + (size_t)c[q] * NV + k0);
#pragma unroll
        for (int q = 0; q < BW; ++q)
            if (j + q < e)
#pragma unroll
                for (int w = 0; w < W; ++w)
                    acc[w] = acc[w] + xv[q].v[w] * v[q];
    }
}

// Sum over the lanes that own the same RHS (lanes l and l + L, l + 2L, ... of a wave),
// then over the 4 waves: one partial per (dot, RHS) per workgroup, fixed order.
template <int NV, int W, int ND>
__device__ __forceinline__ void block_partials(double (&p)[ND][W], double *__restrict__ out,
                                               int64_t lb) {
    constexpr int L = NV / W;
    __shared__ double red[kBlock / 64][ND][NV];
#pragma unroll
    for (int d = 0; d < ND; ++d)
#pragma unroll
        for (int w = 0; w < W; ++w)
#pragma unroll
            for (int off = L; off < 64; off <<= 1)
                p[d][w] = p[d][w] + __shfl_xor(p[d][w], off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane < L)
#pragma unroll
        for (int d = 0; d < ND; ++d)
#pragma unroll
            for (int w = 0; w < W; ++w)
                red[wv][d][lane * W + w] = p[d][w];
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

template <typename VT, int NV>
struct Width {
    static constexpr int value = (16 / (int)sizeof(VT)) < NV ? (16 / (int)sizeof(VT)) : NV;
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
    constexpr int W = Width<VT, NV>::value;
    constexpr int L = NV / W;        // lanes per row
    constexpr int RPB = kBlock / L;  // rows per workgroup pass
    using P = Pack<VT, W>;
    if (done != nullptr && *done)
        return;
    const int64_t lb = logical_block();
    const int64_t start = lb * chunk;
    const int64_t end = start + chunk < nrows ? start + chunk : nrows;
    const int tr = threadIdx.x / L, k0 = (threadIdx.x % L) * W;
    double part[ND > 0 ? ND : 1][W];
#pragma unroll
    for (int d = 0; d < (ND > 0 ? ND : 1); ++d)
#pragma unroll
        for (int w = 0; w < W; ++w)
            part[d][w] = 0.0;

    // row pointers of the next row are loaded one iteration ahead
    int64_t i = start + tr;
    uint32_t jb = 0, je = 0;
    if (i < end) {
        jb = row[i];
        je = row[i + 1];
    }
    for (; i < end; i += RPB) {
        const size_t o = (size_t)i * NV + k0;
        const uint32_t base = panel_base<IT>(pbase, shift, i);
        uint32_t nb = 0, ne = 0;
        if (i + RPB < end) {
            nb = row[i + RPB];
            ne = row[i + RPB + 1];
        }
        VT acc[W];
        if constexpr (OP == OP_ADD) {
            const P y0 = ld<VT, W>((const VT *)yv + o);
#pragma unroll
            for (int w = 0; w < W; ++w)
                acc[w] = y0.v[w];
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w)
                acc[w] = VT(0);
        }
        if constexpr (OP == OP_OFFBLOCK) {
            // t = b - sum over off-block columns (old x), HSGS core blocks of blk_rows
            const int64_t lo = (i / blk_rows) * blk_rows, hi = lo + blk_rows;
            for (uint32_t j = jb; j < je; ++j) {
                const int64_t cc = (int64_t)(base + (uint32_t)col[j]);
                if (cc < lo || cc >= hi) {
                    const P xv = ld<VT, W>(x + (size_t)cc * NV + k0);
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        acc[w] = acc[w] + xv.v[w] * val[j];
                }
            }
        } else {
            row_acc<VT, IT, NV, W>(jb, je, col, val, base, x, k0, acc);
        }
        if constexpr (OP == OP_SET || OP == OP_ADD) {
            P yo;
#pragma unroll
            for (int w = 0; w < W; ++w)
                yo.v[w] = acc[w];
            st<VT, W>((VT *)yv + o, yo);
        } else if constexpr (OP == OP_RESID || OP == OP_OFFBLOCK) {
            const P bv = ld<VT, W>(b + o);
            if (OP == OP_RESID && out_f32) {
                Pack<float, W> t;
#pragma unroll
                for (int w = 0; w < W; ++w)
                    t.v[w] = (float)(bv.v[w] - acc[w]);
                st<float, W>((float *)yv + o, t);
            } else {
                P t;
#pragma unroll
                for (int w = 0; w < W; ++w)
                    t.v[w] = bv.v[w] - acc[w];
                st<VT, W>((VT *)yv + o, t);
            }
        } else if constexpr (OP == OP_RESID0) {
            const P bv = ld<VT, W>(b + o);
            P t;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                t.v[w] = bv.v[w] - acc[w];
                part[0][w] = part[0][w] + (double)(t.v[w] * t.v[w]);
            }
            st<VT, W>((VT *)yv + o, t);
            st<VT, W>(out2 + o, t);
            st<VT, W>(out3 + o, t);
        } else if constexpr (OP == OP_DOT1) {
            const P ov = ld<VT, W>(o1 + o);
            P yo;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                yo.v[w] = acc[w];
                part[0][w] = part[0][w] + (double)(acc[w] * ov.v[w]);
            }
            st<VT, W>((VT *)yv + o, yo);
        } else if constexpr (OP == OP_DOT3) {
            const P sv = ld<VT, W>(o1 + o);
            P yo;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                yo.v[w] = acc[w];
                part[0][w] = part[0][w] + (double)(acc[w] * sv.v[w]);
                part[1][w] = part[1][w] + (double)(acc[w] * acc[w]);
                part[2][w] = part[2][w] + (double)(sv.v[w] * sv.v[w]);
            }
            st<VT, W>((VT *)yv + o, yo);
        } else if constexpr (OP == OP_CHEB0) {
            // r = b - A x ; r *= isd ; t1 = c*r ; then t1 *= isd (u) or x (+)= t1*isd
            const VT di = isd[i];
            const P bv = ld<VT, W>(b + o);
            P rv, t1;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                rv.v[w] = bv.v[w] - acc[w];
                rv.v[w] = rv.v[w] * di;
                t1.v[w] = c * rv.v[w];
                t1.v[w] = t1.v[w] * di;
            }
            if (last) {
                VT *xu = (VT *)xupdv;
                if (!x_zero) {
                    const P xo = ld<VT, W>(xu + o);
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        t1.v[w] = t1.v[w] + xo.v[w];
                }
                st<VT, W>(xu + o, t1);
            } else {
                st<VT, W>(out2 + o, rv);
                st<VT, W>((VT *)yv + o, t1);
            }
        } else if constexpr (OP == OP_CHEB) {
            // t2 = (A u)*isd ; t1 = c*r + t2 ; then u' = t1*isd or x (+)= t1*isd
            const VT di = isd[i];
            const P r0 = ld<VT, W>(rr + o);
            P t1;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const VT t2 = acc[w] * di;
                t1.v[w] = c * r0.v[w] + t2;
                t1.v[w] = t1.v[w] * di;
            }
            if (last) {
                VT *xu = (VT *)xupdv;
                if (!x_zero) {
                    const P xo = ld<VT, W>(xu + o);
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        t1.v[w] = t1.v[w] + xo.v[w];
                }
                st<VT, W>(xu + o, t1);
            } else {
                st<VT, W>((VT *)yv + o, t1);
            }
        } else if constexpr (OP == OP_PROLONG32) {
            // VT = float: temp32 = P xH ; x64 = (double)temp32 + x64
            double *xd = (double *)yv;
            Pack<double, W> xo;
            if (!x_zero)
                xo = ld<double, W>(xd + o);
#pragma unroll
            for (int w = 0; w < W; ++w)
                xo.v[w] = x_zero ? (double)acc[w] : (double)acc[w] + xo.v[w];
            st<double, W>(xd + o, xo);
        }
        jb = nb;
        je = ne;
    }
    if constexpr (ND > 0)
        block_partials<NV, W, ND>(part, partials, lb);
}

struct Launch {
    int grid;
    int64_t chunk;
};

Launch plan(int64_t nrows, int rpb, int cap) {
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
    constexpr int W = Width<VT, NV>::value;
    constexpr int RPB = kBlock / (NV / W);
    const int cap = ND > 0 ? kMaxPartials : (a.grid_cap > 0 ? a.grid_cap : 16384);
    Launch L = plan(A.nrows, RPB, cap);
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
