// Multi-RHS CSR SpMV (SpMM) for gfx950 with fused epilogues.
//
// Restates backend::csr_Ax_y / csr_Axpy (src/blas2/detail/csr_matvec.inl:33-260) on
// NV-interleaved vectors x[i*NV+k].  A row is owned by L = NV/W lanes and each lane owns
// W consecutive right-hand sides (W = 2 in fp64, 4 in fp32: one 16-byte load per lane),
// so every nonzero reads one contiguous NV-wide row of x with 16-byte loads.  Each RHS
// is accumulated in stored nonzero order exactly as the reference
// (acc = acc + x*val, no FMA contraction: the file is built with -ffp-contract=off).
// Column indices are u32 or u16 panel-relative (col = base[row >> shift] + u16), decoded
// in-kernel; a compressed block (the reference's compact-index format) has u8/u16/u32
// compact columns and writes row row_ind[i] of y.  The epilogues fuse the level-1 work that follows an SpMV in the reference
// solvers (residual, dot partials, Chebyshev steps, mixed-precision prolongation, HSGS
// off-block residual).
//
// Workgroups are dealt round-robin over the 8 XCDs; logical_block() renumbers them
// so that each XCD walks one contiguous eighth of the tile schedule (Sched): in plain
// order one eighth of the rows, in folded order one column strip (or a few), so that
// the +-1, +-nx and +-fold neighbours of x are reused from that XCD's L2.
#include "xh_internal.hpp"

#include <algorithm>
#include <type_traits>

namespace xh {

Tuning g_tune;

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

// Results are streamed out with nontemporal stores, so they do not evict the x rows
// that later rows of the launch will read again from L2.
template <typename T, int W>
__device__ __forceinline__ void st(T *__restrict__ p, const Pack<T, W> &v) {
    typedef T vec __attribute__((ext_vector_type(W)));
    vec t;
#pragma unroll
    for (int w = 0; w < W; ++w)
        t[w] = v.v[w];
    __builtin_nontemporal_store(t, reinterpret_cast<vec *>(p));
}

template <typename IT>
__device__ __forceinline__ uint32_t panel_base(const uint32_t *__restrict__ pb, int shift,
                                               int64_t i) {
    if constexpr (sizeof(IT) == 4)
        return 0u;
    else
        return pb[i >> shift];
}

// The nonzeros of a row go in chunks.  With L >= 2 lanes per row, the first LL =
// min(L, 16) lanes of the row's group load a chunk's column indices and values
// cooperatively (E entries per lane, one coalesced load each) and broadcast them across
// the group (bcast); with one lane per row the lane loads 8 entries itself.  The first
// chunk of the next row is loaded one pass ahead (together with its row pointers and
// panel base), so a row of up to CH nonzeros costs one exposed memory round trip: its x
// loads.  Only existing entries are added (no +0 terms), so each RHS gets the
// reference's sequential sum exactly.
// BX: x loads in flight per batch (8; 16 for the few-row, long-row launches, whose
// workgroups leave the GPU mostly idle and whose time is the chain of batches of a row).
template <typename VT, int NV, int W, int BX = 8>
struct Chunk {
    static constexpr int BW = BX;
    static constexpr int L = NV / W;                    // lanes sharing the row
    static constexpr int LL = L < 16 ? L : 16;          // of which load the chunk
    static constexpr int E = L >= 2 ? (BW + LL - 1) / LL : BW; // entries per loading lane
    static constexpr int CH = L >= 2 ? LL * E : BW;     // entries per chunk: 8 or 16
    uint32_t c[E]; // raw column index (panel-relative for u16)
    VT v[E];
    VT d[E];       // SCALE loads: the scale of the entry's column (sd[base + c])
};

// SCALE: also the column's scale sd[base + c] of each loaded entry (a load that depends on
// the column's), broadcast with the column and value
template <typename VT, typename IT, int NV, int W, bool SCALE = false, int BX = 8>
__device__ __forceinline__ void load_chunk(uint32_t j, uint32_t e, const IT *__restrict__ col,
                                           const VT *__restrict__ val,
                                           Chunk<VT, NV, W, BX> &ch, uint32_t base = 0,
                                           const VT *__restrict__ sd = nullptr) {
    // Entries past the row end repeat the row's last entry (its x row exists), so the
    // x loads of a batch need no select; their products are never added.
    using C = Chunk<VT, NV, W, BX>;
    const int gl = C::L >= 2 ? (int)(threadIdx.x & 63) % C::L : 0;
    const bool live = j < e && gl < C::LL;
#pragma unroll
    for (int q = 0; q < C::E; ++q) {
        const uint32_t jj = j + (C::L >= 2 ? (uint32_t)(q * C::LL + gl) : (uint32_t)q);
        const uint32_t jc = jj < e ? jj : e - 1;
        ch.c[q] = live ? (uint32_t)col[jc] : 0u;
        ch.v[q] = live ? val[jc] : VT(0);
    }
    if constexpr (SCALE) {
#pragma unroll
        for (int q = 0; q < C::E; ++q)
            ch.d[q] = live ? sd[base + ch.c[q]] : VT(0);
    }
}

// acc[w] + sum x[base + c][k0 + w] * v over the entries [j, min(j + CH, e)) of `ch`.
// The x loads of a batch are issued together, one per existing entry (exec-masked per
// slot): every load instruction holds an L1 request slot whether it hits or misses, so a
// slot past the row end issues none (1-entry rows: 303 -> 217 us at 128^3, nv = 32);
// the accumulation of a slot past the row end is skipped the same way.
// SKIP: entries whose column lies in [slo, slo + sn) are skipped the same way (neither
// loaded nor added), which keeps the order of the others.  SCALE: each gathered value
// x[c] enters as (cz*(x[c]*d))*d with d the column's scale from the chunk (the zero-x
// first Chebyshev step of OP_CHEBZ, rounded as that step rounds it).
template <typename VT, int NV, int W, bool SKIP = false, bool SCALE = false, int BX = 8>
__device__ __forceinline__ void consume_chunk(uint32_t j, uint32_t e,
                                              const Chunk<VT, NV, W, BX> &ch,
                                              uint32_t base, const VT *__restrict__ x, int k0,
                                              VT (&acc)[W], uint32_t slo = 0, uint32_t sn = 0,
                                              VT cz = VT(0)) {
    using C = Chunk<VT, NV, W, BX>;
    constexpr int BW = C::BW, L = C::L;
    auto kept = [&](uint32_t cabs) { return !SKIP || cabs - slo >= sn; };
    auto scale = [&](Pack<VT, W> &xv, VT dq) __attribute__((always_inline)) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
            VT t = xv.v[w] * dq;
            t = cz * t;
            xv.v[w] = t * dq;
        }
    };
    if constexpr (L >= 2) {
        bool more = true; // uniform within the row's lane group
        static_for<0, C::CH / BW>([&](auto bc) {
            constexpr int b0 = decltype(bc)::value * BW;
            // every broadcast source lane of the group is active: the row (and so `more`)
            // is the same for all L lanes
            more = more && (j + b0 < e);
            if (!more)
                return;
            uint32_t c[BW];
            static_for<0, BW>([&](auto qc) {
                // entry t of the chunk sits in lane t % LL of the group, slot t / LL
                constexpr int t = b0 + decltype(qc)::value;
                c[t - b0] = (uint32_t)bcast<L, t % C::LL>((int)ch.c[t / C::LL]);
            });
            Pack<VT, W> xv[BW];
#pragma unroll
            for (int q = 0; q < BW; ++q)
                if (j + b0 + q < e && kept(base + c[q]))
                    xv[q] = ld<VT, W>(x + (size_t)(base + c[q]) * NV + k0);
            static_for<0, BW>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int t = b0 + q;
                if (j + t < e && kept(base + c[q])) {
                    const VT vq = bcast<L, t % C::LL>(ch.v[t / C::LL]);
                    if constexpr (SCALE)
                        scale(xv[q], bcast<L, t % C::LL>(ch.d[t / C::LL]));
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        acc[w] = acc[w] + xv[q].v[w] * vq;
                }
            });
        });
    } else {
        if (j >= e)
            return;
        Pack<VT, W> xv[BW];
#pragma unroll
        for (int q = 0; q < BW; ++q)
            xv[q] = ld<VT, W>(x + (size_t)(base + ch.c[q]) * NV + k0);
#pragma unroll
        for (int q = 0; q < BW; ++q)
            if (j + q < e && kept(base + ch.c[q])) {
                if constexpr (SCALE)
                    scale(xv[q], ch.d[q]);
#pragma unroll
                for (int w = 0; w < W; ++w)
                    acc[w] = acc[w] + xv[q].v[w] * ch.v[q];
            }
    }
}

// the whole row [j, e), its first chunk already in `first` (SKIP: without the columns
// in [slo, slo + sn))
template <typename VT, typename IT, int NV, int W, bool SKIP = false, bool SCALE = false,
          int BX = 8>
__device__ __forceinline__ void row_acc(uint32_t j, const uint32_t e,
                                        const Chunk<VT, NV, W, BX> &first,
                                        const IT *__restrict__ col, const VT *__restrict__ val,
                                        uint32_t base, const VT *__restrict__ x, int k0,
                                        VT (&acc)[W], uint32_t slo = 0, uint32_t sn = 0,
                                        const VT *__restrict__ sd = nullptr, VT cz = VT(0)) {
    using C = Chunk<VT, NV, W, BX>;
    // a longer row loads its next chunk before the x loads of the current one, so the
    // chunk's latency overlaps them
    C cur = first;
    while (true) {
        const bool more = j + C::CH < e; // the same for all lanes of the row
        C nxt;
        if (more)
            load_chunk<VT, IT, NV, W, SCALE, BX>(j + C::CH, e, col, val, nxt, base, sd);
        consume_chunk<VT, NV, W, SKIP, SCALE, BX>(j, e, cur, base, x, k0, acc, slo, sn, cz);
        if (!more)
            break;
        j += C::CH;
        cur = nxt;
    }
}

// Sum over the lanes that own the same RHS (lanes l and l + L, l + 2L, ... of a wave),
// then over the 4 waves: one partial per (dot, RHS) per workgroup, fixed order.
template <int NV, int W, int ND>
__device__ __forceinline__ void block_partials_in(double (&p)[ND][W], double *__restrict__ out,
                                                  int64_t lb,
                                                  double (&red)[kBlock / 64][ND][NV]) {
    constexpr int L = NV / W;
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

template <int NV, int W, int ND>
__device__ __forceinline__ void block_partials(double (&p)[ND][W], double *__restrict__ out,
                                               int64_t lb) {
    __shared__ double red[kBlock / 64][ND][NV];
    block_partials_in<NV, W, ND>(p, out, lb, red);
}

template <int OP>
struct NDots {
    static constexpr int value = (OP == OP_DOT1 || OP == OP_RESID0) ? 1
                                 : OP == OP_DOT3                     ? 3
                                 : OP == OP_RDOT4                    ? 4
                                                                     : 0;
};

template <typename VT, int NV>
struct Width {
    static constexpr int value = (16 / (int)sizeof(VT)) < NV ? (16 / (int)sizeof(VT)) : NV;
};

// Operands of the fused epilogue of one row, loaded by epi_load (right before the
// epilogue in k_spmv, a few passes ahead in k_spmv_grid).
template <typename VT, int W>
struct EpiIn {
    Pack<VT, W> a, b, c;
    Pack<double, W> xd;
    VT di;
    VT al[W];
};

template <typename VT, int NV, int W, int OP>
__device__ __forceinline__ void epi_load(int64_t i, size_t o, size_t oy, int k0,
                                         EpiIn<VT, W> &in, const void *__restrict__ yv,
                                         const VT *__restrict__ b, const VT *__restrict__ o1,
                                         const VT *__restrict__ isd, const VT *__restrict__ rr,
                                         const VT *__restrict__ out2,
                                         const void *__restrict__ xupdv, int last, int x_zero) {
    if constexpr (OP == OP_ADD) {
        in.a = ld<VT, W>((const VT *)yv + oy);
    } else if constexpr (OP == OP_RESID || OP == OP_OFFBLOCK || OP == OP_RESID0) {
        in.a = ld<VT, W>(b + o);
    } else if constexpr (OP == OP_DOT1 || OP == OP_DOT3) {
        in.a = ld<VT, W>(o1 + o);
    } else if constexpr (OP == OP_RDOT4) {
        in.a = ld<VT, W>(b + o);
        in.b = ld<VT, W>(o1 + o);
        in.c = ld<VT, W>(out2 + o);
#pragma unroll
        for (int w = 0; w < W; ++w)
            in.al[w] = rr[k0 + w];
    } else if constexpr (OP == OP_CHEB0 || OP == OP_CHEB || OP == OP_CHEBZ || OP == OP_CHEBR) {
        in.di = isd[i];
        if constexpr (OP == OP_CHEB0 || OP == OP_CHEBZ || OP == OP_CHEBR)
            in.a = ld<VT, W>(b + o);
        else
            in.a = ld<VT, W>(rr + o);
        if (last && !x_zero) {
            in.b = ld<VT, W>((const VT *)xupdv + o);
        } else {
#pragma unroll
            for (int w = 0; w < W; ++w)
                in.b.v[w] = VT(0);
        }
    } else if constexpr (OP == OP_PROLONG32) {
        if (!x_zero)
            in.xd = ld<double, W>((const double *)yv + o);
    }
}

// The fused epilogue of one row: acc holds the row's W sums (y = A x, or y + A x for
// OP_ADD); o is the row's offset in the nv-interleaved vectors, oy that of the output row;
// `in` holds the operands epi_load read for the row.
template <typename VT, int NV, int W, int OP, int NP>
__device__ __forceinline__ void row_epilogue(
    size_t o, size_t oy, const VT (&acc)[W], double (&part)[NP][W],
    const EpiIn<VT, W> &in, void *__restrict__ yv, VT *__restrict__ out2,
    VT *__restrict__ out3, void *__restrict__ xupdv, VT c, int last, int x_zero, int out_f32) {
    using P = Pack<VT, W>;
    if constexpr (OP == OP_SET || OP == OP_ADD) {
        P yo;
#pragma unroll
        for (int w = 0; w < W; ++w)
            yo.v[w] = acc[w];
        st<VT, W>((VT *)yv + oy, yo);
    } else if constexpr (OP == OP_RESID || OP == OP_OFFBLOCK) {
        const P &bv = in.a;
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
        const P &bv = in.a;
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
        const P &ov = in.a;
        P yo;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            yo.v[w] = acc[w];
            part[0][w] = part[0][w] + (double)(acc[w] * ov.v[w]);
        }
        st<VT, W>((VT *)yv + o, yo);
    } else if constexpr (OP == OP_DOT3) {
        const P &sv = in.a;
        P yo;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            yo.v[w] = acc[w];
            part[0][w] = part[0][w] + (double)(acc[w] * sv.v[w]);
            part[1][w] = part[1][w] + (double)(acc[w] * acc[w]);
            part[2][w] = part[2][w] + (double)(sv.v[w] * sv.v[w]);
        }
        st<VT, W>((VT *)yv + o, yo);
    } else if constexpr (OP == OP_RDOT4) {
        // RBiCGStab 12-13 (merged_rbicgstab_group1): r^ = r - alpha v, then
        // theta = (t, r^), phi = (t, t), psi = (t, r0), eta = (r^, r^)
        const P &vv = in.a, &r0v = in.b, &rv = in.c;
        P yo, rn;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            yo.v[w] = acc[w];
            rn.v[w] = rv.v[w] - in.al[w] * vv.v[w];
            part[0][w] = part[0][w] + (double)(acc[w] * rn.v[w]);
            part[1][w] = part[1][w] + (double)(acc[w] * acc[w]);
            part[2][w] = part[2][w] + (double)(acc[w] * r0v.v[w]);
            part[3][w] = part[3][w] + (double)(rn.v[w] * rn.v[w]);
        }
        st<VT, W>((VT *)yv + o, yo);
        st<VT, W>(out2 + o, rn);
    } else if constexpr (OP == OP_CHEB0) {
        // r = b - A x ; r *= isd ; t1 = c*r ; then t1 *= isd (u) or x (+)= t1*isd
        const VT di = in.di;
        const P &bv = in.a;
        P rv, t1;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            rv.v[w] = bv.v[w] - acc[w];
            rv.v[w] = rv.v[w] * di;
            t1.v[w] = c * rv.v[w];
            t1.v[w] = t1.v[w] * di;
        }
        // last: x (+)= t1 ; else r and u; one store through the selected pointer
        if (last && !x_zero) {
#pragma unroll
            for (int w = 0; w < W; ++w)
                t1.v[w] = t1.v[w] + in.b.v[w];
        }
        if (!last)
            st<VT, W>(out2 + o, rv);
        st<VT, W>((last ? (VT *)xupdv : (VT *)yv) + o, t1);
    } else if constexpr (OP == OP_CHEB || OP == OP_CHEBR) {
        // t2 = (A u)*isd ; t1 = c*r + t2 ; then u' = t1*isd or x (+)= t1*isd
        // (OP_CHEBR: r = b*isd here)
        const VT di = in.di;
        P r0 = in.a;
        if constexpr (OP == OP_CHEBR)
#pragma unroll
            for (int w = 0; w < W; ++w)
                r0.v[w] = r0.v[w] * di;
        P t1;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const VT t2 = acc[w] * di;
            t1.v[w] = c * r0.v[w] + t2;
            t1.v[w] = t1.v[w] * di;
        }
        if (last && !x_zero) {
#pragma unroll
            for (int w = 0; w < W; ++w)
                t1.v[w] = t1.v[w] + in.b.v[w];
        }
        st<VT, W>((last ? (VT *)xupdv : (VT *)yv) + o, t1);
    } else if constexpr (OP == OP_CHEBZ) {
        // r = b*isd (the zero-x first step's residual, stored for the later steps when
        // out2 is given and this step is not the last) ; t2 = (A u)*isd ; t1 = c*r + t2 ;
        // u' or x = t1*isd
        const VT di = in.di;
        P rv, t1;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            rv.v[w] = in.a.v[w] * di;
            const VT t2 = acc[w] * di;
            t1.v[w] = c * rv.v[w] + t2;
            t1.v[w] = t1.v[w] * di;
        }
        if (!last && out2 != nullptr)
            st<VT, W>(out2 + o, rv);
        st<VT, W>((last ? (VT *)xupdv : (VT *)yv) + o, t1);
    } else if constexpr (OP == OP_PROLONG32) {
        // VT = float: temp32 = P xH ; x64 = (double)temp32 + x64
        double *xd = (double *)yv;
        Pack<double, W> xo;
#pragma unroll
        for (int w = 0; w < W; ++w)
            xo.v[w] = x_zero ? (double)acc[w] : (double)acc[w] + in.xd.v[w];
        st<double, W>(xd + o, xo);
    }
}

template <typename VT, typename IT, int NV, int OP, int BX = 8>
__global__ __launch_bounds__(kBlock) void k_spmv(
    const uint32_t *__restrict__ row, const IT *__restrict__ col, const VT *__restrict__ val,
    const uint32_t *__restrict__ pbase, int shift, const Sched sc,
    const VT *__restrict__ x, void *__restrict__ yv, const VT *__restrict__ b,
    const VT *__restrict__ o1, const VT *__restrict__ isd, const VT *__restrict__ rr,
    VT *__restrict__ out2, VT *__restrict__ out3, void *__restrict__ xupdv, VT c, int last,
    int x_zero, int out_f32, double *__restrict__ partials, const int *__restrict__ done,
    int64_t blk_rows, const uint32_t *__restrict__ rind, VT c2) {
    constexpr int ND = NDots<OP>::value;
    constexpr int W = Width<VT, NV>::value;
    constexpr int L = NV / W;        // lanes per row
    constexpr int RPB = kBlock / L;  // rows per workgroup pass
    using P = Pack<VT, W>;
    if (done != nullptr && *done)
        return;
    const int64_t lb = logical_block();
    const int tr = threadIdx.x / L, k0 = (threadIdx.x % L) * W;
    double part[ND > 0 ? ND : 1][W];
#pragma unroll
    for (int d = 0; d < (ND > 0 ? ND : 1); ++d)
#pragma unroll
        for (int w = 0; w < W; ++w)
            part[d][w] = 0.0;

    for (int kt = 0; kt < sc.tpw; ++kt) {
    // launches with dot partials run few workgroups of many tiles each, which walk the
    // tiles of their XCD together; the others take consecutive tiles
    const int64_t k = ND > 0 ? tile_of(sc, lb, kt, gridDim.x) : lb * sc.tpw + kt;
    if (k >= sc.ntiles)
        break; // later kt give larger k
    int64_t start, end;
    tile_rows(sc, k, start, end);
    // Pipeline: the row pointers and panel base of the row two passes ahead and the
    // first chunk of the next row are loaded while this row's x loads are in flight.
    int64_t i = start + tr;
    uint32_t jb = 0, je = 0, bs = 0, nb = 0, ne = 0, nbs = 0;
    if (i < end) {
        jb = row[i];
        je = row[i + 1];
        bs = panel_base<IT>(pbase, shift, i);
    }
    if (i + RPB < end) {
        nb = row[i + RPB];
        ne = row[i + RPB + 1];
        nbs = panel_base<IT>(pbase, shift, i + RPB);
    }
    constexpr bool SC = OP == OP_CHEBZ; // chunks carry their columns' isd
    Chunk<VT, NV, W, BX> ch;
    load_chunk<VT, IT, NV, W, SC, BX>(jb, je, col, val, ch, bs, isd);
    // Drain the prologue loads once per tile: the row loop's head then merges only its
    // own back edge, where the last y store may stay in flight.
    __builtin_amdgcn_s_waitcnt(0);
    for (; i < end; i += RPB) {
        const size_t o = (size_t)i * NV + k0;
        const uint32_t base = bs;
        Chunk<VT, NV, W, BX> nch;
        load_chunk<VT, IT, NV, W, SC, BX>(nb, ne, col, val, nch, nbs, isd);
        uint32_t nnb = 0, nne = 0, nnbs = 0;
        if (i + 2 * RPB < end) {
            nnb = row[i + 2 * RPB];
            nne = row[i + 2 * RPB + 1];
            nnbs = panel_base<IT>(pbase, shift, i + 2 * RPB);
        }
        // output row: row_ind[i] for a compressed block (SET / ADD only)
        const size_t oy = (OP == OP_SET || OP == OP_ADD) && rind != nullptr
                              ? (size_t)rind[i] * NV + k0
                              : o;
        // the epilogue's operands, issued before the row's x loads so that both are in
        // flight together (nv = 8: 146 -> 140 us per finest Chebyshev step), except in
        // launches with dot partials, whose fewer workgroups of many tiles lost by it
        // (OP_DOT1 425 -> 442 us)
        EpiIn<VT, W> in;
        if constexpr (ND == 0)
            epi_load<VT, NV, W, OP>(i, o, oy, k0, in, yv, b, o1, isd, rr, out2, xupdv, last,
                                    x_zero);
        VT acc[W];
        if constexpr (OP == OP_ADD) {
            const P y0 = ld<VT, W>((const VT *)yv + oy);
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
            const uint32_t lo = (uint32_t)((i / blk_rows) * blk_rows);
            row_acc<VT, IT, NV, W, true, false, BX>(jb, je, ch, col, val, base, x, k0, acc, lo,
                                         (uint32_t)blk_rows);
        } else if constexpr (OP == OP_CHEBZ) {
            // the operand is b; each gathered entry enters as the zero-x first step's u
            row_acc<VT, IT, NV, W, false, true, BX>(jb, je, ch, col, val, base, x, k0, acc, 0u, 0u,
                                                isd, c2);
        } else {
            row_acc<VT, IT, NV, W, false, false, BX>(jb, je, ch, col, val, base, x, k0, acc);
        }
        if constexpr (ND > 0)
            epi_load<VT, NV, W, OP>(i, o, oy, k0, in, yv, b, o1, isd, rr, out2, xupdv, last,
                                    x_zero);
        row_epilogue<VT, NV, W, OP>(o, oy, acc, part, in, yv, out2, out3, xupdv, c, last,
                                    x_zero, out_f32);
        jb = nb;
        je = ne;
        bs = nbs;
        nb = nnb;
        ne = nne;
        nbs = nnbs;
        ch = nch;
    }
    }
    if constexpr (ND > 0)
        block_partials<NV, W, ND>(part, partials, lb);
}


// ---- structured-grid SpMV ----
// Rows that are the points of a gx x gy x gz grid (row = (z*gy + y)*gx + c) with columns
// that are grid neighbours (|dz|, |dy|, |dx| <= 1: 7- and 27-point operators).  A
// workgroup owns a tile of ty x tz grid lines (one row of each per pass, ty*tz = RPB) and
// sweeps it along x.  LDS holds the x rows of the tile's lines plus one halo line on
// each side, (ty+2) x (tz+2) lines, at four consecutive positions (a ring: c-1, c, c+1
// read by pass c, c+2 written during it), so an x row is read from L2 about
// (ty+2)(tz+2)/(ty*tz) times per launch instead of once per nonzero that names it.
// The view applies when every entry is a grid neighbour (Mat::gcode < 27), so every x
// read comes from the window.  The products and their order are those of k_spmv: the
// results are identical.

template <typename VT, int NV, int W>
struct GChunk {
    static constexpr int BW = 8;
    static constexpr int L = NV / W;
    static constexpr int LL = L < 16 ? L : 16;
    // a 27-point row fits one chunk when the row has 2 or more lanes
    static constexpr int E = L >= 2 ? (32 + LL - 1) / LL : BW;
    static constexpr int CH = L >= 2 ? LL * E : BW;
    uint32_t cd[E];
    VT v[E];
};

template <typename VT, int NV, int W>
__device__ __forceinline__ void load_gchunk(uint32_t j, uint32_t e, const VT *__restrict__ val,
                                            const uint8_t *__restrict__ gcode,
                                            GChunk<VT, NV, W> &ch) {
    // every lane loads, at an index clamped into the row (or to entry 0 for an empty
    // range), so no load sits behind a branch; slots past the row end are never used
    using C = GChunk<VT, NV, W>;
    const int gl = C::L >= 2 ? (int)(threadIdx.x & 63) % C::L : 0;
#pragma unroll
    for (int q = 0; q < C::E; ++q) {
        const uint32_t jj = j + (C::L >= 2 ? (uint32_t)(q * C::LL + gl) : (uint32_t)q);
        const uint32_t jc = jj < e ? jj : (e > j ? e - 1 : 0u);
        ch.cd[q] = (uint32_t)gcode[jc];
        ch.v[q] = val[jc];
    }
}

// an LDS (address space 3) object: accesses compile to ds_* instructions
template <typename T>
using lds_t = __attribute__((address_space(3))) T;

template <typename VT, int W>
__device__ __forceinline__ Pack<VT, W> lds_row(const lds_t<Pack<VT, W>> *src) {
    Pack<VT, W> r;
    if constexpr (W == 1) {
        r.v[0] = *(const lds_t<VT> *)src;
    } else {
        typedef VT vec __attribute__((ext_vector_type(W)));
        const vec t = *(const lds_t<vec> *)src;
#pragma unroll
        for (int w = 0; w < W; ++w)
            r.v[w] = t[w];
    }
    return r;
}


// acc[w] + sum over the entries [j, min(j + CH, e)) of `ch`, x from the LDS window.  The
// window reads go in batches of 4, unconditionally (a slot past the row end holds the
// row's last code, a valid window address), so a batch is in flight together; only the
// accumulation of a slot past the row end is skipped.
template <typename VT, int NV, int W, int NLY, int NL, int L>
__device__ __forceinline__ void consume_gchunk(uint32_t j, uint32_t e, const GChunk<VT, NV, W> &ch,
                                               int pos, int lbase,
                                               const lds_t<Pack<VT, W>> *win, int lane,
                                               VT (&acc)[W]) {
    using C = GChunk<VT, NV, W>;
    constexpr int BW = 4; // 8 in flight costs occupancy and measured slower
    auto addr = [&](uint32_t cd) __attribute__((always_inline)) {
        const int a = (int)(cd / 9u) - 1, b = (int)((cd / 3u) % 3u) - 1, d = (int)(cd % 3u) - 1;
        return win + ((((pos + d) & 3) * NL + lbase + a * NLY + b) * L + lane);
    };
    bool more = true; // uniform over the row's lanes
    static_for<0, C::CH / BW>([&](auto bc) {
        constexpr int b0 = decltype(bc)::value * BW;
        more = more && (j + b0 < e);
        if (!more)
            return;
        VT vq[BW];
        Pack<VT, W> xv[BW];
        static_for<0, BW>([&](auto qc) {
            constexpr int t = b0 + decltype(qc)::value;
            uint32_t cd;
            if constexpr (C::L >= 2) {
                cd = (uint32_t)bcast<C::L, t % C::LL>((int)ch.cd[t / C::LL]);
                vq[t - b0] = bcast<C::L, t % C::LL>(ch.v[t / C::LL]);
            } else {
                cd = ch.cd[t];
                vq[t - b0] = ch.v[t];
            }
            xv[t - b0] = lds_row<VT, W>(addr(cd));
        });
#pragma unroll
        for (int q = 0; q < BW; ++q)
            if (j + b0 + q < e)
#pragma unroll
                for (int w = 0; w < W; ++w)
                    acc[w] = acc[w] + xv[q].v[w] * vq[q];
    });
}

template <typename VT, int NV, int OP>
__global__ __launch_bounds__(kBlock) void k_spmv_grid(
    const uint32_t *__restrict__ row, const VT *__restrict__ val,
    const uint8_t *__restrict__ gcode, const GridSched g, const VT *__restrict__ x, void *__restrict__ yv, const VT *__restrict__ b,
    const VT *__restrict__ o1, const VT *__restrict__ isd, const VT *__restrict__ rr,
    VT *__restrict__ out2, VT *__restrict__ out3, void *__restrict__ xupdv, VT c, int last,
    int x_zero, int out_f32, double *__restrict__ partials, const int *__restrict__ done) {
    constexpr int ND = NDots<OP>::value;
    constexpr int W = Width<VT, NV>::value;
    constexpr int L = NV / W;
    constexpr int RPB = kBlock / L;
    constexpr int TZ = grid_tz(RPB), TY = RPB / TZ;
    constexpr int NLY = TY + 2, NL = NLY * (TZ + 2);
    constexpr int UNITS = NL * L; // 16-byte pieces of one window position
    constexpr int R = (UNITS + kBlock - 1) / kBlock;
    using P = Pack<VT, W>;
    using C = GChunk<VT, NV, W>;
    // the window doubles as the buffer of the final dot reduction
    union Shared {
        P win[4][NL][L];
        double red[kBlock / 64][ND > 0 ? ND : 1][NV];
    };
    __shared__ Shared sh;
    const lds_t<P> *wr = (const lds_t<P> *)&sh.win[0][0][0];
    if (done != nullptr && *done)
        return;
    const int64_t lb = logical_block();
    const int grp = threadIdx.x / L, ty = grp % TY, tz = grp / TY;
    const int lane = threadIdx.x % L, k0 = lane * W;
    const int lbase = (tz + 1) * NLY + (ty + 1);
    double part[ND > 0 ? ND : 1][W];
#pragma unroll
    for (int d = 0; d < (ND > 0 ? ND : 1); ++d)
#pragma unroll
        for (int w = 0; w < W; ++w)
            part[d][w] = 0.0;
    // One pass (one row of each tile line) uses what was loaded three passes earlier: the
    // x rows of window position c + 2, the chunk of row c and the epilogue operands of
    // row c, held in three register sets that rotate with the position.  A set also
    // carries the row pointers of its next row (c + 3), loaded one round ahead of that
    // row's chunk.
    struct Set {
        P s[R];
        C ch;
        EpiIn<VT, W> in;
        uint32_t jb, je, nb, ne;
    };
    for (int it = 0; it < g.ipw; ++it) {
        const int k = (int)lb * g.ipw + it;
        if (k >= g.nitems)
            break; // uniform over the workgroup
        int yb, zb, c0, c1;
        grid_item(g, k, yb, zb, c0, c1);
        const int wy = yb * TY - 1, wz = zb * TZ - 1; // window origin (halo line)
        const int yg = yb * TY + ty, zg = zb * TZ + tz;
        const bool valid = yg < g.ny && zg < g.nz;
        const int64_t rb = valid ? ((int64_t)zg * g.ny + yg) * g.nx : 0;
        // position p of every window line: piece u = threadIdx.x + q*kBlock is lane u % L
        // of line u / L; lines and positions outside the grid are never read
        // every piece is loaded (at a row clamped into the grid, so no load sits behind a
        // branch); pieces of lines or positions outside the grid are never read
        auto stage_load = [&](int p, P (&sr)[R]) __attribute__((always_inline)) {
            const int pc = p < 0 ? 0 : (p >= g.nx ? g.nx - 1 : p);
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int u = (int)threadIdx.x + q * kBlock;
                const int ln = u < UNITS ? u / L : NL - 1, lp = u % L;
                int yy = wy + ln % NLY, zz = wz + ln / NLY;
                yy = yy < 0 ? 0 : (yy >= g.ny ? g.ny - 1 : yy);
                zz = zz < 0 ? 0 : (zz >= g.nz ? g.nz - 1 : zz);
                sr[q] = ld<VT, W>(x + ((size_t)((int64_t)zz * g.ny + yy) * g.nx + pc) * NV +
                                  lp * W);
            }
        };
        auto stage_store = [&](int p, const P (&sr)[R]) __attribute__((always_inline)) {
#pragma unroll
            for (int q = 0; q < R; ++q) {
                const int u = (int)threadIdx.x + q * kBlock;
                if (u < UNITS)
                    sh.win[p & 3][u / L][u % L] = sr[q];
            }
        };
        // loads for row cc (chunk, epilogue operands), window position cc + 2 and the row
        // pointers of row cc + 3; S.nb / S.ne hold the row pointers of row cc
        // every load is issued (indices clamped into the line), so none sits behind a
        // branch and the compiler keeps the three sets in flight
        auto issue = [&](int cc, Set &S) __attribute__((always_inline)) {
            const bool rv = valid && cc < c1;
            S.jb = rv ? S.nb : 0u;
            S.je = rv ? S.ne : 0u;
            load_gchunk<VT, NV, W>(S.jb, S.je, val, gcode, S.ch);
            S.nb = row[rb + (cc + 3 < g.nx ? cc + 3 : g.nx)];
            S.ne = row[rb + (cc + 4 < g.nx ? cc + 4 : g.nx)];
            const int64_t i = rb + (cc < g.nx ? cc : g.nx - 1);
            const size_t o = (size_t)i * NV + k0;
            epi_load<VT, NV, W, OP>(i, o, o, k0, S.in, yv, b, o1, isd, rr, out2, xupdv, last,
                                    x_zero);
            stage_load(cc + 2, S.s);
        };
        // pass cc: rows of position cc (threads of rows outside the grid have an empty row
        // and skip the epilogue), then the set is refilled for pass cc + 3
        auto pass = [&](int cc, Set &S) __attribute__((always_inline)) {
            const int64_t i = rb + cc;
            const size_t o = (size_t)i * NV + k0;
            const uint32_t jb = S.jb, je = S.je;
            VT acc[W];
#pragma unroll
            for (int w = 0; w < W; ++w)
                acc[w] = OP == OP_ADD ? S.in.a.v[w] : VT(0);
            consume_gchunk<VT, NV, W, NLY, NL, L>(jb, je, S.ch, cc, lbase, wr, lane, acc);
            if (je - jb > (uint32_t)C::CH)
                for (uint32_t j = jb + C::CH; j < je; j += C::CH) {
                    C ech;
                    load_gchunk<VT, NV, W>(j, je, val, gcode, ech);
                    consume_gchunk<VT, NV, W, NLY, NL, L>(j, je, ech, cc, lbase, wr, lane,
                                                               acc);
                }
            if (valid)
                row_epilogue<VT, NV, W, OP>(o, o, acc, part, S.in, yv, out2, out3, xupdv,
                                            c, last, x_zero, out_f32);
            if (cc + 2 <= c1)
                stage_store(cc + 2, S.s);
            issue(cc + 3, S);
            __syncthreads();
        };
        {
            P t[R];
            stage_load(c0 - 1, t);
            stage_store(c0 - 1, t);
            stage_load(c0, t);
            stage_store(c0, t);
            stage_load(c0 + 1, t);
            stage_store(c0 + 1, t);
        }
        __syncthreads();
        Set S0, S1, S2;
        S0.nb = row[rb + c0];
        S0.ne = row[rb + (c0 + 1 < g.nx ? c0 + 1 : g.nx)];
        S1.nb = S0.ne;
        S1.ne = row[rb + (c0 + 2 < g.nx ? c0 + 2 : g.nx)];
        S2.nb = S1.ne;
        S2.ne = row[rb + (c0 + 3 < g.nx ? c0 + 3 : g.nx)];
        issue(c0, S0);
        issue(c0 + 1, S1);
        issue(c0 + 2, S2);
        int cb = c0;
        for (; cb + 3 <= c1; cb += 3) {
            pass(cb, S0);
            pass(cb + 1, S1);
            pass(cb + 2, S2);
        }
        if (cb < c1)
            pass(cb, S0);
        if (cb + 1 < c1)
            pass(cb + 1, S1);
    }
    if constexpr (ND > 0)
        block_partials_in<NV, W, ND>(part, partials, lb, sh.red);
}

// ---- 7-point stencil SpMV: register window along grid lines ----
// Matrices whose rows are the points of an nx x ny x nz grid in natural order and whose
// row i holds exactly the diagonal and its in-grid neighbours -z, -y, -x, +x, +y, +z in
// that order (diagonal first, then ascending columns: the reference generators' order)
// carry a second value layout set at upload, st7[i][8] = the row's values in the slots
// {C, -z, -y, -x, +x, +y, +z, unused} (0 where the neighbour lies outside the grid), and
// need neither column indices nor row pointers.  A wave owns GY x GZ grid lines (one row
// of each at a time, L lanes per row) and sweeps them along x: x(c-1) and x(c) stay in
// registers from the previous steps and only x(c+1) is loaded; a neighbour line inside
// the wave's block is read from the lanes that own it (cross-lane permute) and only the
// lines on the block's border are loaded.  For nv = 32 in fp64 (2 x 2 lines) a step of 4
// rows issues 3 x-row loads where the row-parallel kernel issues 7.  The next step's x
// rows, values and epilogue operands are loaded while the current step computes.  Each
// present entry is added in stored order with the same products as the row-parallel
// kernel: the results are identical.

// lines per wave in y and z for G = 64 / L row groups
template <int G>
struct StBlock {
    static constexpr int GY = G >= 32 ? 8 : G >= 8 ? 4 : G >= 2 ? 2 : 1;
    static constexpr int GZ = G / GY;
};

template <typename VT, int W>
__device__ __forceinline__ Pack<VT, W> shfl_pack(const Pack<VT, W> &v, int src) {
    Pack<VT, W> r;
#pragma unroll
    for (int w = 0; w < W; ++w)
        r.v[w] = __shfl(v.v[w], src, 64);
    return r;
}

template <typename VT, int NV, int OP>
__global__ __launch_bounds__(kBlock) void k_spmv_st7(
    const VT *__restrict__ st7, const StSched g, const VT *__restrict__ x,
    void *__restrict__ yv, const VT *__restrict__ b, const VT *__restrict__ o1,
    const VT *__restrict__ isd, const VT *__restrict__ rr, VT *__restrict__ out2,
    VT *__restrict__ out3, void *__restrict__ xupdv, VT c, int last, int x_zero, int out_f32,
    double *__restrict__ partials, const int *__restrict__ done) {
    constexpr int ND = NDots<OP>::value;
    constexpr int W = Width<VT, NV>::value;
    constexpr int L = NV / W;      // lanes per row
    constexpr int G = 64 / L;      // rows (lines) per wave
    constexpr int GY = StBlock<G>::GY, GZ = StBlock<G>::GZ;
    constexpr int LV = L < 8 ? L : 8;   // lanes of a row that load its 8 value slots
    constexpr int EV = 8 / LV;          // slots per loading lane
    using P = Pack<VT, W>;
    if (done != nullptr && *done)
        return;
    const int64_t lb = logical_block();
    const int lane = threadIdx.x & 63;
    const int grp = lane / L, a = grp % GY, bz = grp / GY;
    const int gl = lane % L, k0 = gl * W;
    const int64_t nwave = (int64_t)gridDim.x * (kBlock / 64);
    const int64_t wid = lb * (kBlock / 64) + (threadIdx.x >> 6);
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    double part[ND > 0 ? ND : 1][W];
#pragma unroll
    for (int d = 0; d < (ND > 0 ? ND : 1); ++d)
#pragma unroll
        for (int w = 0; w < W; ++w)
            part[d][w] = 0.0;
    P zero;
#pragma unroll
    for (int w = 0; w < W; ++w)
        zero.v[w] = VT(0);

    for (int64_t item = wid; item < g.nitems; item += nwave) {
        const int yb = (int)(item % g.nyb);
        const int64_t t = item / g.nyb;
        const int zb = (int)(t % g.nzb), seg = (int)(t / g.nzb);
        const int y = yb * GY + a, z = zb * GZ + bz;
        const int c0 = seg * g.seglen;
        const int c1 = c0 + g.seglen < nx ? c0 + g.seglen : nx;
        const bool rowok = y < ny && z < nz;
        const int64_t line = rowok ? ((int64_t)z * ny + y) * nx : 0;
        // the border line each row loads in y and z (-1: none), and for a single z line
        // both z neighbours
        const bool ylo = GY == 1 || a == 0, yhi = GY == 1 || a == GY - 1;
        const bool zlo = GZ == 1 || bz == 0, zhi = GZ == 1 || bz == GZ - 1;
        const bool has_s = y > 0, has_n = y + 1 < ny, has_d = z > 0, has_u = z + 1 < nz;
        // line offsets (rows) of the loaded border neighbours
        const int64_t ldy_s = -(int64_t)nx, ldy_n = nx;
        const int64_t ldz_d = -(int64_t)nx * ny, ldz_u = (int64_t)nx * ny;
        const bool ld_s = rowok && ylo && has_s, ld_n = rowok && yhi && has_n;
        const bool ld_d = rowok && zlo && has_d, ld_u = rowok && zhi && has_u;
        auto xat = [&](int64_t r) __attribute__((always_inline)) {
            return ld<VT, W>(x + (size_t)r * NV + k0);
        };
        // registers of step cc: border neighbours (y: S/N, z: D/U), values, epilogue
        // With two or more lines per direction a row loads one border line in y (south
        // for the first line of the block, north for the last; none in between) and one
        // in z, held in one register set each.
        struct Step {
            P by0, by1, bz0, bz1; // y: S, N (by1 only for GY == 1); z: D, U (bz1: GZ == 1)
            VT v[EV];
        };
        auto load_step = [&](int cc, Step &S) __attribute__((always_inline)) {
            const int64_t i = line + cc;
            if constexpr (GY == 1) {
                S.by0 = ld_s ? xat(i + ldy_s) : zero;
                S.by1 = ld_n ? xat(i + ldy_n) : zero;
            } else {
                S.by0 = ld_s ? xat(i + ldy_s) : (ld_n ? xat(i + ldy_n) : zero);
            }
            if constexpr (GZ == 1) {
                S.bz0 = ld_d ? xat(i + ldz_d) : zero;
                S.bz1 = ld_u ? xat(i + ldz_u) : zero;
            } else {
                S.bz0 = ld_d ? xat(i + ldz_d) : (ld_u ? xat(i + ldz_u) : zero);
            }
#pragma unroll
            for (int q = 0; q < EV; ++q)
                S.v[q] = (rowok && gl < LV) ? st7[(size_t)i * 8 + q * LV + gl] : VT(0);
        };
        P xw = zero, xc = zero, xe = zero;
        if (rowok) {
            if (c0 > 0)
                xw = xat(line + c0 - 1);
            xc = xat(line + c0);
            if (c0 + 1 < nx)
                xe = xat(line + c0 + 1);
        }
        Step cur, nxt;
        load_step(c0, cur);
        for (int cc = c0; cc < c1; ++cc) {
            // next step: x(cc + 2), border rows and values of cc + 1
            const bool more = cc + 1 < c1;
            P xn = zero;
            if (rowok && more && cc + 2 < nx)
                xn = xat(line + cc + 2);
            if (more)
                load_step(cc + 1, nxt);
            // neighbours inside the wave's block from the lanes that own those lines
            // epilogue operands of this row (issued before the cross-lane moves and the sums)
            EpiIn<VT, W> in;
            if (rowok) {
                const size_t o = (size_t)(line + cc) * NV + k0;
                epi_load<VT, NV, W, OP>(line + cc, o, o, k0, in, yv, b, o1, isd, rr, out2, xupdv,
                                        last, x_zero);
            }
            P ys, yn, zd, zu;
            if constexpr (GY == 1) {
                ys = cur.by0;
                yn = cur.by1;
            } else if constexpr (GY == 2) {
                const P sib = shfl_pack<VT, W>(xc, lane ^ L);
                ys = a == 0 ? cur.by0 : sib;
                yn = a == 0 ? sib : cur.by0;
            } else {
                const P sl = shfl_pack<VT, W>(xc, lane >= L ? lane - L : lane);
                const P sh = shfl_pack<VT, W>(xc, lane + L < 64 ? lane + L : lane);
                ys = a == 0 ? cur.by0 : sl;
                yn = a == GY - 1 ? cur.by0 : sh;
            }
            if constexpr (GZ == 1) {
                zd = cur.bz0;
                zu = cur.bz1;
            } else if constexpr (GZ == 2) {
                const P sib = shfl_pack<VT, W>(xc, lane ^ (GY * L));
                zd = bz == 0 ? cur.bz0 : sib;
                zu = bz == 0 ? sib : cur.bz0;
            } else {
                const P sl = shfl_pack<VT, W>(xc, lane >= GY * L ? lane - GY * L : lane);
                const P sh = shfl_pack<VT, W>(xc, lane + GY * L < 64 ? lane + GY * L : lane);
                zd = bz == 0 ? cur.bz0 : sl;
                zu = bz == GZ - 1 ? cur.bz0 : sh;
            }
            VT vs[7];
            static_for<0, 7>([&](auto sc) {
                constexpr int s = decltype(sc)::value;
                if constexpr (L >= 2)
                    vs[s] = bcast<L, s % LV>(cur.v[s / LV]);
                else
                    vs[s] = cur.v[s];
            });
            if (rowok) {
                VT acc[W];
#pragma unroll
                for (int w = 0; w < W; ++w)
                    acc[w] = OP == OP_ADD ? in.a.v[w] : VT(0);
                auto add = [&](const P &xv, VT v) __attribute__((always_inline)) {
#pragma unroll
                    for (int w = 0; w < W; ++w)
                        acc[w] = acc[w] + xv.v[w] * v;
                };
                add(xc, vs[0]);
                if (has_d)
                    add(zd, vs[1]);
                if (has_s)
                    add(ys, vs[2]);
                if (cc > 0)
                    add(xw, vs[3]);
                if (cc + 1 < nx)
                    add(xe, vs[4]);
                if (has_n)
                    add(yn, vs[5]);
                if (has_u)
                    add(zu, vs[6]);
                const size_t o = (size_t)(line + cc) * NV + k0;
                row_epilogue<VT, NV, W, OP>(o, o, acc, part, in, yv, out2, out3, xupdv, c,
                                            last, x_zero, out_f32);
            }
            xw = xc;
            xc = xe;
            xe = xn;
            if (more)
                cur = nxt;
        }
    }
    if constexpr (ND > 0)
        block_partials<NV, W, ND>(part, partials, lb);
}
} // namespace

// Tile schedule and grid of one launch: `rpb` rows per workgroup pass, at most `cap`
// workgroups, `row_bytes` bytes of x and y per row (nv * sizeof(value)).
Launch spmv_plan(int64_t n, int64_t nnz, int64_t F, int rpb, int cap, int64_t row_bytes) {
    Launch L;
    Sched &s = L.sc;
    s.nrows = n;
    if (g_tune.fold_enable && F > 0 && F < n && 2 * F * row_bytes > g_tune.fold_min_bytes) {
        // strips: multiple of 8 (one XCD each, or a few), narrow enough that the rows
        // touched between the first and last use of an x row (two strip rows of x, y
        // and matrix) fit the budget
        const int64_t per_row = 2 * row_bytes + 12 * std::max<int64_t>(1, nnz / std::max<int64_t>(1, n));
        int64_t ns = 8;
        while (2 * ((F + ns - 1) / ns) * per_row > g_tune.strip_budget && (F + ns - 1) / ns > 2 * rpb)
            ns *= 2;
        s.fold = F;
        s.strip = (F + ns - 1) / ns;
        s.na = (n + F - 1) / F;
        int64_t t = std::min<int64_t>((int64_t)g_tune.tile_passes * rpb, s.strip);
        t = std::max<int64_t>(rpb, t / rpb * rpb);
        s.tile = (int32_t)t;
        s.tps = (int32_t)((s.strip + t - 1) / t);
        s.ntiles = ((F + s.strip - 1) / s.strip) * s.na * s.tps;
    } else {
        // at least min_passes passes per workgroup, so the next row's loads overlap
        int64_t need = (n + rpb - 1) / rpb;
        need = std::max<int64_t>(1, need / std::max(1, g_tune.min_passes));
        int64_t g = std::min<int64_t>(need, cap);
        g = std::max<int64_t>(8, (g + 7) / 8 * 8);
        int64_t chunk = (n + g - 1) / g;
        chunk = (chunk + rpb - 1) / rpb * rpb;
        s.fold = n;
        s.strip = n;
        s.na = 1;
        s.tile = (int32_t)chunk;
        s.tps = (int32_t)((n + chunk - 1) / chunk);
        s.ntiles = s.tps;
    }
    s.tpw = (int32_t)((s.ntiles + cap - 1) / cap);
    int64_t g = (s.ntiles + s.tpw - 1) / s.tpw;
    L.grid = (int)std::max<int64_t>(8, (g + 7) / 8 * 8);
    return L;
}


GridLaunch grid_plan(int nx, int ny, int nz, int rpb, int cap) {
    GridLaunch G;
    GridSched &g = G.g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.tz = grid_tz(rpb);
    g.ty = rpb / g.tz;
    g.nyb = (ny + g.ty - 1) / g.ty;
    g.nzb = (nz + g.tz - 1) / g.tz;
    const int64_t tiles = (int64_t)g.nyb * g.nzb;
    int64_t nseg = std::max<int64_t>(1, (g_tune.grid_items + tiles - 1) / tiles);
    nseg = std::min<int64_t>(nseg, std::max(1, nx / std::max(1, g_tune.grid_min_seg)));
    g.seglen = (int)((nx + nseg - 1) / nseg);
    g.nseg = (nx + g.seglen - 1) / g.seglen;
    g.nitems = (int)(tiles * g.nseg);
    g.ipw = (g.nitems + cap - 1) / cap;
    const int64_t grid = (g.nitems + g.ipw - 1) / g.ipw;
    G.grid = (int)std::max<int64_t>(8, (grid + 7) / 8 * 8);
    return G;
}
StLaunch st7_plan(int nx, int ny, int nz, int gy, int gz, int cap) {
    StLaunch S;
    StSched &g = S.g;
    g.nx = nx;
    g.ny = ny;
    g.nz = nz;
    g.nyb = (ny + gy - 1) / gy;
    g.nzb = (nz + gz - 1) / gz;
    const int64_t blocks = (int64_t)g.nyb * g.nzb;
    // resident waves of one launch (at most one per item, at most `cap` workgroups)
    const int64_t waves = std::min<int64_t>(std::max(8, g_tune.st7_waves),
                                            (int64_t)cap * (kBlock / 64));
    // segments per line: the count (up to 8, segments of at least st7_min_seg rows) whose
    // items fill the waves' rounds best, the fewest on a tie
    const int64_t maxseg =
        std::max<int64_t>(1, std::min<int64_t>(8, nx / std::max(1, g_tune.st7_min_seg)));
    int64_t nseg = 1;
    double best = -1.0;
    for (int64_t q = 1; q <= maxseg; ++q) {
        const int64_t items = blocks * q;
        const int64_t w = std::min(items, waves);
        const int64_t rounds = (items + w - 1) / w;
        const double eff = (double)items / (double)(rounds * w) * (double)w / (double)waves;
        if (eff > best + 1e-9) {
            best = eff;
            nseg = q;
        }
    }
    g.seglen = (int)((nx + nseg - 1) / nseg);
    g.nseg = (nx + g.seglen - 1) / g.seglen;
    g.nitems = blocks * g.nseg;
    const int64_t nw = std::min<int64_t>(g.nitems, waves);
    int64_t grid = (nw + (kBlock / 64) - 1) / (kBlock / 64);
    grid = std::min<int64_t>(grid, cap);
    S.grid = (int)std::max<int64_t>(8, (grid + 7) / 8 * 8);
    if (S.grid > cap)
        S.grid = cap / 8 * 8;
    return S;
}

// Few rows of many entries (the coarse levels of an aggressively coarsened hierarchy: a
// launch of at most 512 workgroups, rows averaging 32 entries or more) run 16 x loads per
// batch instead of 8: such a launch leaves most of the GPU idle, and its time is the chain
// of a row's batches (about a microsecond each), not bandwidth.  With ~1000 workgroups
// the wider kernel's registers (104 against 78 VGPRs) already cost more than it saves.
bool spmv_wide_batches(const Mat &A, int rpb) {
    return g_tune.wide_batches && A.row_ind == nullptr && A.nnz >= 32 * A.nrows &&
           (A.nrows + rpb - 1) / rpb <= 512;
}

namespace {

template <typename VT, typename IT, int NV, int OP>
int launch_t(const Mat &A, const void *x, void *y, const SpmvArgs &a, hipStream_t s) {
    constexpr int ND = NDots<OP>::value;
    constexpr int W = Width<VT, NV>::value;
    constexpr int RPB = kBlock / (NV / W);
    const int cap = ND > 0 ? g_tune.dot_groups : (a.grid_cap > 0 ? a.grid_cap : 16384);
    if constexpr (sizeof(IT) != 1 && OP != OP_OFFBLOCK && OP != OP_PROLONG32 &&
                  OP != OP_CHEBZ && OP != OP_CHEBR && NV >= 8) {
        if (A.st7 != nullptr && g_tune.st7_enable && A.row_ind == nullptr) {
            constexpr int G = 64 / (NV / W);
            const StLaunch S =
                st7_plan(A.gx, A.gy, A.gz, StBlock<G>::GY, StBlock<G>::GZ, cap);
            hipLaunchKernelGGL((k_spmv_st7<VT, NV, OP>), dim3(S.grid), dim3(kBlock), 0, s,
                               (const VT *)A.st7, S.g, (const VT *)x, y, (const VT *)a.b,
                               (const VT *)a.o1, (const VT *)a.isd, (const VT *)a.r,
                               (VT *)a.out2, (VT *)a.out3, a.xupd, (VT)a.c, a.last, a.x_zero,
                               a.out_f32, a.partials, a.done);
            return ND > 0 ? S.grid : 0;
        }
    }
    // the off-block residual of hybrid GS stays on the row path: decoding each entry's
    // column to skip the row's own block made the grid sweep slower (466 against 369 us on
    // the 96^3 27-point operator)
    if constexpr (sizeof(IT) != 1 && OP != OP_OFFBLOCK && OP != OP_PROLONG32 &&
                  OP != OP_CHEBZ && OP != OP_CHEBR) {
        if (A.gcode != nullptr && grid_on(A) && A.row_ind == nullptr) {
            // the grid sweep's dot launches take their own partial count: more work items
            // pay there (27-point, 96^3: 409 -> 336 us with 2048), unlike the row path
            const GridLaunch G = grid_plan(A.gx, A.gy, A.gz, RPB,
                                           ND > 0 ? g_tune.grid_dot_groups : cap);
            hipLaunchKernelGGL((k_spmv_grid<VT, NV, OP>), dim3(G.grid), dim3(kBlock), 0, s,
                               A.row, (const VT *)A.val, A.gcode, G.g,
                               (const VT *)x, y,
                               (const VT *)a.b, (const VT *)a.o1, (const VT *)a.isd,
                               (const VT *)a.r, (VT *)a.out2, (VT *)a.out3, a.xupd, (VT)a.c,
                               a.last, a.x_zero, a.out_f32, a.partials, a.done);
            return ND > 0 ? G.grid : 0;
        }
    }
    const Launch L = spmv_plan(A.nrows, A.nnz, A.fold, RPB, cap, (int64_t)NV * (int64_t)sizeof(VT));
    if constexpr (sizeof(IT) != 1 && (NV == 8 || NV == 32) &&
                  (OP == OP_SET || OP == OP_ADD || OP == OP_RESID || OP == OP_CHEB0 ||
                   OP == OP_CHEB || OP == OP_CHEBZ || OP == OP_CHEBR)) {
        if (spmv_wide_batches(A, RPB)) {
            hipLaunchKernelGGL((k_spmv<VT, IT, NV, OP, 16>), dim3(L.grid), dim3(kBlock), 0, s,
                               A.row, (const IT *)A.col, (const VT *)A.val, A.panel_base,
                               A.panel_shift, L.sc, (const VT *)x, y, (const VT *)a.b,
                               (const VT *)a.o1, (const VT *)a.isd, (const VT *)a.r,
                               (VT *)a.out2, (VT *)a.out3, a.xupd, (VT)a.c, a.last, a.x_zero,
                               a.out_f32, a.partials, a.done, a.blk_rows, A.row_ind,
                               (VT)a.c2);
            return ND > 0 ? L.grid : 0;
        }
    }
    hipLaunchKernelGGL((k_spmv<VT, IT, NV, OP>), dim3(L.grid), dim3(kBlock), 0, s, A.row,
                       (const IT *)A.col, (const VT *)A.val, A.panel_base, A.panel_shift,
                       L.sc, (const VT *)x, y, (const VT *)a.b, (const VT *)a.o1,
                       (const VT *)a.isd, (const VT *)a.r, (VT *)a.out2, (VT *)a.out3, a.xupd,
                       (VT)a.c, a.last, a.x_zero, a.out_f32, a.partials, a.done, a.blk_rows,
                       A.row_ind, (VT)a.c2);
    return ND > 0 ? L.grid : 0;
}

template <typename VT, typename IT, int NV>
int by_op(const Mat &A, const void *x, void *y, SpmvOp op, const SpmvArgs &a, hipStream_t s) {
    if (A.row_ind != nullptr && op != OP_SET && op != OP_ADD)
        return -1; // a compressed block only takes y = A x and y += A x
    if constexpr (sizeof(IT) == 1) {
        // u8 columns only occur in compressed blocks
        if (op == OP_SET)
            return launch_t<VT, IT, NV, OP_SET>(A, x, y, a, s);
        if (op == OP_ADD)
            return launch_t<VT, IT, NV, OP_ADD>(A, x, y, a, s);
        return -1;
    }
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
    case OP_CHEBZ:
        return launch_t<VT, IT, NV, OP_CHEBZ>(A, x, y, a, s);
    case OP_CHEBR:
        return launch_t<VT, IT, NV, OP_CHEBR>(A, x, y, a, s);
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
        case OP_RDOT4:
            return launch_t<VT, IT, NV, OP_RDOT4>(A, x, y, a, s);
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

namespace {

const char *op_name(SpmvOp op) {
    static const char *names[] = {"OP_SET",  "OP_ADD",       "OP_RESID",    "OP_DOT1",
                                  "OP_DOT3", "OP_CHEB0",     "OP_CHEB",     "OP_PROLONG32",
                                  "OP_OFFBLOCK", "OP_RESID0", "OP_RDOT4", "OP_CHEBZ",
                                  "OP_CHEBR"};
    return (op >= 0 && op <= OP_CHEBR) ? names[op] : "OP_?";
}

int op_ndots(SpmvOp op) {
    return (op == OP_DOT1 || op == OP_RESID0) ? 1 : op == OP_DOT3 ? 3 : op == OP_RDOT4 ? 4 : 0;
}

// the launch_t conditions for the 7-point stencil and the structured-grid kernels
bool takes_st7(const Mat &A, SpmvOp op, int nv) {
    return A.idx != IDX_U8 && op != OP_OFFBLOCK && op != OP_PROLONG32 && op != OP_CHEBZ &&
           op != OP_CHEBR && nv >= 8 &&
           A.st7 != nullptr && g_tune.st7_enable && A.row_ind == nullptr;
}
bool takes_grid(const Mat &A, SpmvOp op) {
    return A.idx != IDX_U8 && op != OP_OFFBLOCK && op != OP_PROLONG32 && op != OP_CHEBZ &&
           op != OP_CHEBR && A.gcode != nullptr && grid_on(A) && A.row_ind == nullptr;
}

// Algorithmic HBM bytes of one SpMV launch: the matrix streams once (values, column
// indices or grid codes, row pointers, panel bases, row map), x once (ncols rows), and
// the epilogue's vector operands and results once each; dot partials are added by the
// caller from the launch's partial count.
double spmv_bytes(const Mat &A, int nv, SpmvOp op, const SpmvArgs &a, bool grid, bool st7) {
    const double sv = A.prec == F64 ? 8.0 : 4.0;
    const double si = grid ? 1.0 : A.idx == IDX_U32 ? 4.0 : A.idx == IDX_U16 ? 2.0 : 1.0;
    const double n = (double)A.nrows, V = n * nv * sv;
    double m = (double)A.nnz * (sv + si) + (n + 1.0) * 4.0;
    if (st7)
        m = n * 8.0 * sv; // the 8 value slots of every row, nothing else
    if (A.idx == IDX_U16 && !grid && !st7)
        m += (double)(((A.nrows - 1) >> A.panel_shift) + 1) * 4.0;
    if (A.row_ind != nullptr)
        m += n * 4.0;
    const double xb = (double)A.ncols * nv * sv;
    double vec = 0.0;
    switch (op) {
    case OP_SET:
        vec = V; // y
        break;
    case OP_ADD:
        vec = 2.0 * V; // y read and written
        break;
    case OP_RESID:
        vec = V + n * nv * (a.out_f32 ? 4.0 : sv); // b ; y
        break;
    case OP_OFFBLOCK:
    case OP_DOT1:
    case OP_DOT3:
        vec = 2.0 * V; // b or o1 ; y
        break;
    case OP_RESID0:
        vec = 4.0 * V; // b ; y, out2, out3
        break;
    case OP_RDOT4:
        vec = 5.0 * V; // v, r0, r ; t, r^
        break;
    case OP_CHEB0:
        // b, isd ; last: x (the operand) written, else r and u written
        vec = V + n * sv + (a.last ? V : 2.0 * V);
        break;
    case OP_CHEB:
        // r, isd ; last: x written (and read unless zero-flagged), else u' written
        vec = V + n * sv + (a.last ? (a.x_zero ? V : 2.0 * V) : V);
        break;
    case OP_CHEBZ:
        // b is the operand (counted with x) ; isd ; last: x written, else u' (and r when
        // stored) written
        vec = n * sv + (a.last ? V : (a.out2 ? 2.0 * V : V));
        break;
    case OP_CHEBR:
        // b, isd ; last: x written, else u' written
        vec = V + n * sv + V;
        break;
    case OP_PROLONG32:
        vec = n * nv * 8.0 * (a.x_zero ? 1.0 : 2.0); // fp64 x written (and read)
        break;
    default:
        break;
    }
    return m + xb + vec;
}

int spmv_dispatch(const Mat &A, int nv, const void *x, void *y, SpmvOp op, const SpmvArgs &a,
                  hipStream_t s) {
    if (A.prec == F64) {
        if (A.idx == IDX_U8)
            return by_nv<double, uint8_t>(A, nv, x, y, op, a, s);
        return A.idx == IDX_U16 ? by_nv<double, uint16_t>(A, nv, x, y, op, a, s)
                                : by_nv<double, uint32_t>(A, nv, x, y, op, a, s);
    }
    if (A.idx == IDX_U8)
        return by_nv<float, uint8_t>(A, nv, x, y, op, a, s);
    return A.idx == IDX_U16 ? by_nv<float, uint16_t>(A, nv, x, y, op, a, s)
                            : by_nv<float, uint32_t>(A, nv, x, y, op, a, s);
}

} // namespace

const char *prec_name(int prec) { return prec == F64 ? "double" : "float"; }
const char *idx_name(int idx) {
    return idx == IDX_U32 ? "u32" : idx == IDX_U16 ? "u16" : "u8";
}

// whether op on A runs on the row-parallel kernel (neither the stencil nor the grid path)
bool spmv_row_parallel(const Mat &A, SpmvOp op, int nv) {
    return !takes_st7(A, op, nv) && !takes_grid(A, op);
}

int spmv(const Mat &A, int nv, const void *x, void *y, SpmvOp op, const SpmvArgs &a,
         hipStream_t s) {
    if (A.nrows == 0)
        return 0;
    LedgerScope led(s);
    const int np = spmv_dispatch(A, nv, x, y, op, a, s);
    if (led.on) {
        const bool st7 = takes_st7(A, op, nv);
        const bool grid = !st7 && takes_grid(A, op);
        const bool wide = !st7 && !grid && (nv == 8 || nv == 32) && A.idx != IDX_U8 &&
                          (op == OP_SET || op == OP_ADD || op == OP_RESID || op == OP_CHEB0 ||
                           op == OP_CHEB || op == OP_CHEBZ || op == OP_CHEBR) &&
                          spmv_wide_batches(A, kBlock / (nv / std::min(nv, A.prec == F64 ? 2 : 4)));
        const std::string tail = std::to_string(nv) + "," + op_name(op) + (wide ? ",16>" : ">");
        led.kernel = st7    ? std::string("k_spmv_st7<") + prec_name(A.prec) + "," + tail
                     : grid ? std::string("k_spmv_grid<") + prec_name(A.prec) + "," + tail
                            : std::string("k_spmv<") + prec_name(A.prec) + "," +
                                  idx_name(A.idx) + "," + tail;
        led.bytes = spmv_bytes(A, nv, op, a, grid, st7) +
                    (np > 0 ? (double)np * op_ndots(op) * nv * 8.0 : 0.0);
    }
    return np;
}

} // namespace xh
