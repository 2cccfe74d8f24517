// Batched independent systems: many CSR systems of different sizes and RHS counts solved
// together, each iteration one sequence of launches over all systems (captured as one
// hipGraph by the runtime).  A launch walks a list of work items (system, row range);
// a fixed grid of workgroups takes items blockIdx.x, blockIdx.x + gridDim.x, ..., so
// uneven systems share one launch.  Inside an item a lane owns (row, RHS) of that
// system (nv <= 16 lanes per row, 256/nv rows per pass).
//
// Each system follows the single-system arithmetic exactly: SpMV row sums in stored
// order without FMA contraction (csr_Ax_y, src/blas2/detail/csr_matvec.inl:149-260), the
// Chebyshev steps of chebyshev.inl:179-196 and the BiCGStab updates of
// src/solvers/detail/pbicgstab.inl:141-230; dot partials go one block per item and are
// reduced per system, in item order, by finalize_body.
#include "xh_internal.hpp"
#include "xh_krylov.hpp"

#include <algorithm>

namespace xh {
namespace {

// per-item reduction of nd dots over the rows of the item, fixed order: lane tid holds W
// consecutive RHS of one row, so red[d][tid*W + w] is row tid / (nv/W), RHS
// (tid % (nv/W))*W + w, i.e. red[d][q*nv + k] for pass row q and RHS k
template <int W>
__device__ __forceinline__ void item_partials(const double (&part)[3][W], int nd, int nv,
                                              double *__restrict__ out) {
    __shared__ double red[3][kBlock * 2];
    const int tid = threadIdx.x;
    for (int d = 0; d < nd; ++d)
#pragma unroll
        for (int w = 0; w < W; ++w)
            red[d][tid * W + w] = part[d][w];
    __syncthreads();
    if (tid < nd * nv) {
        const int d = tid / nv, k = tid % nv, rpp = kBlock * W / nv;
        double s = red[d][k];
        for (int q = 1; q < rpp; ++q)
            s = s + red[d][q * nv + k];
        out[d * nv + k] = s;
    }
    __syncthreads();
}

// One work item of one batched operation with W right-hand sides per lane (L = nv/W
// lanes per row, kBlock/L rows per pass).  Each RHS gets the arithmetic of the
// single-system kernels; W changes only which lane holds it and so the grouping of the
// dot partials.
template <int OP, int W, int L>
__device__ __forceinline__ void batch_item(const BatchDev &d, const BatchItem &w,
                                           const BatchVecs &v, int ci, int last,
                                           double *__restrict__ partials) {
    constexpr int ND = (OP == B_RESID0 || OP == B_DOT1 || OP == B_XR_UPDATE) ? 1
                       : (OP == B_DOT3)                                      ? 3
                                                                             : 0;
    const int tid = threadIdx.x;
    const SolverState *st = d.st + w.sys;
    constexpr int lanes = L, rpp = kBlock / L;
    const int nv = w.nv;
    const int tr = tid / lanes, k0 = (tid % lanes) * W;
    const int64_t gb = d.rbase[w.sys];
    const size_t vo = (size_t)d.voff[w.sys];
    double part[3][W];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int u = 0; u < W; ++u)
            part[q][u] = 0.0;
    double c1[W], c2[W], c3[W];
#pragma unroll
    for (int u = 0; u < W; ++u) {
        const int k = k0 + u;
        c1[u] = c2[u] = c3[u] = 0.0;
        if constexpr (OP == B_S_UPDATE) {
            c1[u] = st->nalpha[k];
        } else if constexpr (OP == B_XR_UPDATE) {
            c1[u] = st->ac[k];
            c2[u] = st->oc[k];
            c3[u] = st->omega[k];
        } else if constexpr (OP == B_P_UPDATE) {
            c1[u] = st->pcoef[k];
            c2[u] = st->beta[k];
        } else if constexpr (OP == B_CHEB0Z || OP == B_CHEB || OP == B_CHEBZ) {
            c1[u] = d.cheb[(size_t)w.sys * kChebMax + ci];
        }
    }
    // B_CHEBZ: the zero-x first step's coefficient, applied to every gathered entry
    const double cz = OP == B_CHEBZ ? d.cheb[(size_t)w.sys * kChebMax + ci + 1] : 0.0;
    if (tr < rpp)
        for (int64_t i = w.r0 + tr; i < w.r1; i += rpp) {
            const size_t ob = vo + (size_t)i * nv + k0;
            const int64_t g = gb + i;
            if constexpr (OP == B_RESID0 || OP == B_DOT1 || OP == B_DOT3 || OP == B_CHEB ||
                          OP == B_CHEBZ) {
                const double *xs = v.in + vo;
                double acc[W];
#pragma unroll
                for (int u = 0; u < W; ++u)
                    acc[u] = 0.0;
                const uint32_t j1 = d.row[g + 1];
                // x row of entry column c: W consecutive RHS (16-B aligned for W = 2:
                // system vectors start on 256 B, k0 and nv even)
                auto ldx = [&](uint32_t c, double (&xq)[W]) __attribute__((always_inline)) {
                    const double *xp = xs + (size_t)c * nv + k0;
                    if constexpr (W == 2) {
                        const double2 t2 = *(const double2 *)xp;
                        xq[0] = t2.x;
                        xq[1] = t2.y;
                    } else {
                        xq[0] = xp[0];
                    }
                };
                // B_CHEBZ: a gathered entry enters as the zero-x first step's value
                // (cz*(b*isd_c))*isd_c, rounded as that step rounds it
                auto zstep = [&](double (&xq)[W], double dc) __attribute__((always_inline)) {
                    if constexpr (OP == B_CHEBZ)
#pragma unroll
                        for (int u = 0; u < W; ++u) {
                            double t = xq[u] * dc;
                            t = cz * t;
                            xq[u] = t * dc;
                        }
                };
                if constexpr (L == 1) {
                    // entries in batches of EB: columns and values of the batch, then its x
                    // rows, all in flight together; slots past the row end re-read the last
                    // entry (in bounds) and are not added, so the sum keeps the stored order
                    constexpr int EB = 4;
                    for (uint32_t j = d.row[g]; j < j1; j += EB) {
                        uint32_t c[EB];
                        double a[EB], xv[EB][W], dz[EB];
#pragma unroll
                        for (int q = 0; q < EB; ++q) {
                            const uint32_t jj = j + q < j1 ? j + q : j1 - 1;
                            c[q] = d.col[jj];
                            a[q] = d.val[jj];
                        }
#pragma unroll
                        for (int q = 0; q < EB; ++q) {
                            ldx(c[q], xv[q]);
                            if constexpr (OP == B_CHEBZ)
                                dz[q] = d.isd[gb + c[q]];
                        }
#pragma unroll
                        for (int q = 0; q < EB; ++q)
                            if (j + q < j1) {
                                zstep(xv[q], dz[q]);
#pragma unroll
                                for (int u = 0; u < W; ++u)
                                    acc[u] = acc[u] + xv[q][u] * a[q];
                            }
                    }
                } else {
                    // the L lanes of the row load L consecutive entries, one each, and
                    // broadcast column and value across the row's lanes (DPP); then the L x
                    // rows are loaded together and added in stored order (slots past the
                    // row end re-read the last entry and are not added).  The lanes of a
                    // row are active together, as the broadcast needs.
                    const uint32_t lr = (uint32_t)(tid % L);
                    for (uint32_t j = d.row[g]; j < j1; j += L) {
                        const uint32_t jm = j + lr < j1 ? j + lr : j1 - 1;
                        const uint32_t cm = d.col[jm];
                        const double am = d.val[jm];
                        // B_CHEBZ: the entry's column scale, loaded by the same lane
                        const double dm = OP == B_CHEBZ ? d.isd[gb + cm] : 0.0;
                        uint32_t c[L];
                        double a[L], xv[L][W], dz[L];
                        static_for<0, L>([&](auto qc) {
                            constexpr int q = decltype(qc)::value;
                            c[q] = (uint32_t)bcast<L, q>((int)cm);
                            a[q] = bcast<L, q>(am);
                            if constexpr (OP == B_CHEBZ)
                                dz[q] = bcast<L, q>(dm);
                        });
#pragma unroll
                        for (int q = 0; q < L; ++q)
                            ldx(c[q], xv[q]);
#pragma unroll
                        for (int q = 0; q < L; ++q)
                            if (j + q < j1) {
                                zstep(xv[q], dz[q]);
#pragma unroll
                                for (int u = 0; u < W; ++u)
                                    acc[u] = acc[u] + xv[q][u] * a[q];
                            }
                    }
                }
#pragma unroll
                for (int u = 0; u < W; ++u) {
                    const size_t o = ob + u;
                    if constexpr (OP == B_RESID0) {
                        // r0 = b - A x ; p = r = r0 ; (r0, r0)
                        const double t = v.b[o] - acc[u];
                        part[0][u] = part[0][u] + t * t;
                        v.out[o] = t;
                        v.out2[o] = t;
                        v.out3[o] = t;
                    } else if constexpr (OP == B_DOT1) {
                        v.out[o] = acc[u];
                        part[0][u] = part[0][u] + acc[u] * v.o1[o];
                    } else if constexpr (OP == B_DOT3) {
                        const double sv = v.o1[o];
                        v.out[o] = acc[u];
                        part[0][u] = part[0][u] + acc[u] * sv;
                        part[1][u] = part[1][u] + acc[u] * acc[u];
                        part[2][u] = part[2][u] + sv * sv;
                    } else if constexpr (OP == B_CHEB) {
                        // r = b*isd ; t1 = c*r + (A u)*isd ; u' = t1*isd or x = t1*isd
                        const double di = d.isd[g];
                        const double rv = v.o1[o] * di;
                        const double t2 = acc[u] * di;
                        double t1 = c1[u] * rv + t2;
                        t1 = t1 * di;
                        if (last)
                            v.out2[o] = t1; // x is zero-flagged: x = t1
                        else
                            v.out[o] = t1;
                    } else { // B_CHEBZ: r = b*isd ; t1 = c*r + (A u)*isd ; u' or x = t1*isd
                        const double di = d.isd[g];
                        const double rv = v.in[o] * di;
                        const double t2 = acc[u] * di;
                        double t1 = c1[u] * rv + t2;
                        t1 = t1 * di;
                        if (last)
                            v.out2[o] = t1;
                        else
                            v.out[o] = t1;
                    }
                }
            } else {
#pragma unroll
                for (int u = 0; u < W; ++u) {
                    const size_t o = ob + u;
                    if constexpr (OP == B_CHEB0Z) {
                        // residual of a zero-flagged x is b: r = b*isd ; t1 = (c*r)*isd
                        const double di = d.isd[g];
                        const double rv = v.b[o] * di;
                        double t1 = c1[u] * rv;
                        t1 = t1 * di;
                        if (last)
                            v.out2[o] = t1;
                        else
                            v.out[o] = t1;
                    } else if constexpr (OP == B_COPY) {
                        v.out[o] = v.b[o];
                    } else if constexpr (OP == B_S_UPDATE) {
                        v.out[o] = v.b[o] + c1[u] * v.o1[o]; // s = r + (-alpha) v
                    } else if constexpr (OP == B_XR_UPDATE) {
                        // x = ac p_ + oc s_ + x ; r = s - omega t ; (r, r0)
                        v.x[o] = c1[u] * v.p_[o] + c2[u] * v.s_[o] + v.x[o];
                        const double rv = v.s[o] - c3[u] * v.t[o];
                        v.out[o] = rv;
                        part[0][u] = part[0][u] + rv * v.o1[o];
                    } else if constexpr (OP == B_P_UPDATE) {
                        // p = r + (-(omega beta)) v + beta p
                        v.out[o] = v.b[o] + c1[u] * v.o1[o] + c2[u] * v.out[o];
                    }
                }
            }
        }
    if constexpr (ND > 0)
        item_partials<W>(part, ND, nv, partials + (size_t)w.slot * kBatchPart);
}

// one batched operation over the work items wg, wg + nwg, ... (one item per workgroup at
// a time); an item's `wpl` is its RHS per lane
template <int OP>
__device__ __forceinline__ void batch_phase(const BatchDev &d,
                                            const BatchItem *__restrict__ items, int nitems,
                                            const BatchVecs &v, int ci, int last,
                                            double *__restrict__ partials, int wg, int nwg) {
    for (int it = wg; it < nitems; it += nwg) {
        const BatchItem w = items[it];
        if (d.st[w.sys].done)
            continue; // uniform over the workgroup: one item per workgroup at a time
        // lanes per row L = nv / (RHS per lane), uniform over the item
        switch (w.nv / w.wpl) {
        case 1:
            if (w.wpl == 2)
                batch_item<OP, 2, 1>(d, w, v, ci, last, partials);
            else
                batch_item<OP, 1, 1>(d, w, v, ci, last, partials);
            break;
        case 2:
            if (w.wpl == 2)
                batch_item<OP, 2, 2>(d, w, v, ci, last, partials);
            else
                batch_item<OP, 1, 2>(d, w, v, ci, last, partials);
            break;
        case 4:
            if (w.wpl == 2)
                batch_item<OP, 2, 4>(d, w, v, ci, last, partials);
            else
                batch_item<OP, 1, 4>(d, w, v, ci, last, partials);
            break;
        case 8:
            if (w.wpl == 2)
                batch_item<OP, 2, 8>(d, w, v, ci, last, partials);
            else
                batch_item<OP, 1, 8>(d, w, v, ci, last, partials);
            break;
        default:
            batch_item<OP, 1, 16>(d, w, v, ci, last, partials);
            break;
        }
    }
}

template <int OP>
__global__ __launch_bounds__(kBlock) void k_batch(BatchDev d, const BatchItem *__restrict__ items,
                                                  int nitems, const int *__restrict__ nact,
                                                  BatchVecs v, int ci, int last,
                                                  double *__restrict__ partials) {
    const int n = nact != nullptr ? min(*nact, nitems) : nitems;
    batch_phase<OP>(d, items, n, v, ci, last, partials, blockIdx.x, gridDim.x);
}

// Work list of the systems not yet done: workgroup s copies system s's items (when it is
// not done) to the position after the items of the active systems before it; workgroup 0
// writes the total.  Every workgroup reads the same done flags (set by earlier launches).
__global__ __launch_bounds__(256) void k_batch_compact(const BatchItem *__restrict__ items,
                                                       const int *__restrict__ item_start,
                                                       const int *__restrict__ item_count,
                                                       const SolverState *__restrict__ st,
                                                       int nsys, BatchItem *__restrict__ out,
                                                       int *__restrict__ nact) {
    const int s = blockIdx.x;
    __shared__ int s_off;
    if (threadIdx.x == 0) {
        int off = 0;
        for (int q = 0; q < s; ++q)
            off += st[q].done ? 0 : item_count[q];
        s_off = off;
        if (s == 0) {
            int tot = 0;
            for (int q = 0; q < nsys; ++q)
                tot += st[q].done ? 0 : item_count[q];
            *nact = tot;
        }
    }
    __syncthreads();
    if (st[s].done)
        return;
    const int b = item_start[s], c = item_count[s], off = s_off;
    for (int q = threadIdx.x; q < c; q += blockDim.x)
        out[off + q] = items[b + q];
}

__global__ __launch_bounds__(256) void k_batch_finalize(int op, SolverState *__restrict__ st,
                                                        const double *__restrict__ partials,
                                                        const int *__restrict__ item_start,
                                                        const int *__restrict__ item_count,
                                                        const int *__restrict__ nvs) {
    const int s = blockIdx.x;
    finalize_body(op, st + s, partials + (size_t)item_start[s] * kBatchPart, item_count[s],
                  nvs[s], kBatchPart);
}

// flag[0] = 1 when every system is done; flag[1] = number of systems not done
__global__ __launch_bounds__(256) void k_batch_all_done(const SolverState *__restrict__ st,
                                                        int nsys, int *__restrict__ flag) {
    __shared__ int cnt;
    if (threadIdx.x == 0)
        cnt = 0;
    __syncthreads();
    int c = 0;
    for (int s = threadIdx.x; s < nsys; s += blockDim.x)
        c += st[s].done ? 0 : 1;
    atomicAdd(&cnt, c);
    __syncthreads();
    if (threadIdx.x == 0) {
        flag[0] = cnt == 0;
        flag[1] = cnt;
    }
}

__global__ void k_stamp(unsigned long long *t) { *t = wall_clock64(); }

// Grid-wide barrier of the single-launch solve, with a bounded wait.  Barrier number g
// (0, 1, 2, ... in every workgroup) is complete when the arrival counter reaches
// (g + 1) * nwg: the counter is never reset, so no waiter can miss a release.  A
// workgroup that waits longer than `limit` wall-clock ticks, or sees that another one gave
// up, sets the error flag and returns false; its caller then ends the kernel, and so every
// workgroup ends within one more limit.
__device__ bool grid_wait(BatchBarrier *gb, unsigned long long target,
                          unsigned long long limit) {
    __shared__ int s_ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence(); // this workgroup's writes before the barrier, visible device-wide
        atomicAdd(&gb->arrived, 1ull);
        const unsigned long long t0 = wall_clock64();
        int ok = 0;
        while (true) {
            if (__hip_atomic_load(&gb->arrived, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) >=
                target) {
                ok = 1;
                break;
            }
            if (__hip_atomic_load(&gb->err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                break;
            if (wall_clock64() - t0 > limit) {
                atomicExch(&gb->err, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        __threadfence(); // later loads see the other workgroups' writes
        s_ok = ok;
    }
    __syncthreads();
    return s_ok != 0;
}

// The whole batched solve in one cooperative launch: the launch sequence of one iteration
// (preconditioner, SpMVs with dot partials, per-system finalize, vector updates) runs as
// phases of this kernel separated by grid-wide barriers, and the iterations loop on the
// device until every system is done or max_iters is reached, so the host issues one
// launch per solve and never polls.  Each phase is the same device code over the same
// work items and partial slots as the separate launches, so the results are identical
// to them.  Every workgroup takes every barrier: the phase sequence and the exit test
// (all `done` flags, read after a barrier) are the same for all of them, and the loop is
// bounded by max_iters; a barrier that times out ends the kernel (P.bar->err).
__global__ __launch_bounds__(kBlock) void k_batch_solve(BatchSolve P) {
    const int wg = blockIdx.x, nwg = gridDim.x;
    const BatchDev &d = P.d;
    unsigned long long gen = 0;
    bool ok = true;
    auto sync = [&]() __attribute__((always_inline)) {
        if (ok)
            ok = grid_wait(P.bar, (++gen) * (unsigned long long)nwg, P.wait_ticks);
        return ok;
    };
    auto fin = [&](FinOp op) __attribute__((always_inline)) {
        for (int s = wg; s < P.nsys; s += nwg) {
            finalize_body(op, d.st + s, P.partials + (size_t)P.item_start[s] * kBatchPart,
                          P.item_count[s], P.nvs[s], kBatchPart);
            __syncthreads();
        }
        return sync();
    };
    // x = M^-1 b from a zero x (batch_precond)
    auto precond = [&](const double *b, double *x) __attribute__((always_inline)) -> bool {
        if (P.order > 0) {
            const int k = P.order;
            BatchVecs v;
            v.b = b;
            v.out = k == 1 ? x : P.cu1;
            v.out2 = x;
            int first = k - 2;
            if (k >= 2) {
                // the zero-x first step and the next one in one phase
                BatchVecs z;
                z.in = b;
                z.out = P.cu1;
                z.out2 = x;
                batch_phase<B_CHEBZ>(d, P.items, P.nitems, z, k - 2, k == 2, nullptr, wg, nwg);
                first = k - 3;
            } else {
                batch_phase<B_CHEB0Z>(d, P.items, P.nitems, v, k - 1, k == 1, nullptr, wg, nwg);
            }
            if (!sync())
                return false;
            double *uc = P.cu1, *un = P.cu2;
            for (int i = first; i >= 0; --i) {
                BatchVecs w;
                w.in = uc;
                w.o1 = b;
                w.out = un;
                w.out2 = x;
                batch_phase<B_CHEB>(d, P.items, P.nitems, w, i, i == 0, nullptr, wg, nwg);
                if (!sync())
                    return false;
                double *t = uc;
                uc = un;
                un = t;
            }
            return true;
        }
        BatchVecs v;
        v.b = b;
        v.out = x;
        batch_phase<B_COPY>(d, P.items, P.nitems, v, 0, 0, nullptr, wg, nwg);
        return sync();
    };
    // r0 = b - A x ; p = r = r0 ; rho0
    {
        BatchVecs v;
        v.in = P.x;
        v.b = P.b;
        v.out = P.r0;
        v.out2 = P.p;
        v.out3 = P.r;
        batch_phase<B_RESID0>(d, P.items, P.nitems, v, 0, 0, P.partials, wg, nwg);
        if (!sync() || !fin(FIN_INIT))
            return;
    }
    __shared__ int s_all;
    double *pp = P.pre ? P.p_ : P.p, *ss = P.pre ? P.s_ : P.s;
    for (int it = 0; it <= P.max_iters; ++it) {
        if (threadIdx.x == 0) {
            int all = 1;
            for (int s = 0; s < P.nsys; ++s)
                all = all && d.st[s].done;
            s_all = all;
        }
        __syncthreads();
        if (s_all)
            break; // the same decision in every workgroup: the flags were read after a barrier
        __syncthreads();
        if (P.pre && !precond(P.p, P.p_))
            return;
        BatchVecs v1;
        v1.in = pp;
        v1.o1 = P.r0;
        v1.out = P.v;
        batch_phase<B_DOT1>(d, P.items, P.nitems, v1, 0, 0, P.partials, wg, nwg);
        if (!sync() || !fin(FIN_ALPHA))
            return;
        BatchVecs vs;
        vs.b = P.r;
        vs.o1 = P.v;
        vs.out = P.s;
        batch_phase<B_S_UPDATE>(d, P.items, P.nitems, vs, 0, 0, nullptr, wg, nwg);
        if (!sync())
            return;
        if (P.pre && !precond(P.s, P.s_))
            return;
        BatchVecs v3;
        v3.in = ss;
        v3.o1 = P.s;
        v3.out = P.t;
        batch_phase<B_DOT3>(d, P.items, P.nitems, v3, 0, 0, P.partials, wg, nwg);
        if (!sync() || !fin(FIN_OMEGA))
            return;
        BatchVecs vx;
        vx.x = P.x;
        vx.p_ = pp;
        vx.s_ = ss;
        vx.s = P.s;
        vx.t = P.t;
        vx.o1 = P.r0;
        vx.out = P.r;
        batch_phase<B_XR_UPDATE>(d, P.items, P.nitems, vx, 0, 0, P.partials, wg, nwg);
        if (!sync() || !fin(FIN_BETA))
            return;
        BatchVecs vp;
        vp.b = P.r;
        vp.o1 = P.v;
        vp.out = P.p;
        batch_phase<B_P_UPDATE>(d, P.items, P.nitems, vp, 0, 0, nullptr, wg, nwg);
        if (!sync())
            return;
    }
}

} // namespace

int batch_solve_blocks() {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_solve, kBlock, 0) !=
            hipSuccess ||
        per_cu < 1)
        return 0;
    per_cu = std::min(per_cu, 2); // well inside residency: 2 workgroups (8 waves) per CU
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 0;
    return per_cu * cus;
}

hipError_t batch_solve_launch(const BatchSolve &P, int grid, hipStream_t s) {
    void *args[] = {(void *)&P};
    return hipLaunchCooperativeKernel((const void *)k_batch_solve, dim3(grid), dim3(kBlock),
                                      args, 0, s);
}

void batch_launch(BatchOp op, const BatchDev &d, const BatchItem *items, int nitems, int grid,
                  const BatchVecs &v, int ci, int last, double *partials, hipStream_t s,
                  const int *nact) {
    if (nitems == 0)
        return;
    const dim3 g(std::max(1, std::min(grid, nitems))), b(kBlock);
#define XH_BATCH_CASE(OPV)                                                                       \
    case OPV:                                                                                    \
        hipLaunchKernelGGL(k_batch<OPV>, g, b, 0, s, d, items, nitems, nact, v, ci, last,       \
                           partials);                                                            \
        break;
    switch (op) {
        XH_BATCH_CASE(B_RESID0)
        XH_BATCH_CASE(B_DOT1)
        XH_BATCH_CASE(B_DOT3)
        XH_BATCH_CASE(B_CHEB0Z)
        XH_BATCH_CASE(B_CHEB)
        XH_BATCH_CASE(B_CHEBZ)
        XH_BATCH_CASE(B_COPY)
        XH_BATCH_CASE(B_S_UPDATE)
        XH_BATCH_CASE(B_XR_UPDATE)
        XH_BATCH_CASE(B_P_UPDATE)
    }
#undef XH_BATCH_CASE
}

void batch_finalize(FinOp op, SolverState *st, const double *partials, const int *item_start,
                    const int *item_count, const int *nvs, int nsys, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_finalize, dim3(nsys), dim3(256), 0, s, (int)op, st, partials,
                       item_start, item_count, nvs);
}

void batch_compact(const BatchItem *items, const int *item_start, const int *item_count,
                   const SolverState *st, int nsys, BatchItem *items_act, int *nact,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_batch_compact, dim3(nsys), dim3(256), 0, s, items, item_start,
                       item_count, st, nsys, items_act, nact);
}

void batch_all_done(const SolverState *st, int nsys, int *flag, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_all_done, dim3(1), dim3(256), 0, s, st, nsys, flag);
}

void stamp_clock(unsigned long long *t, hipStream_t s) {
    hipLaunchKernelGGL(k_stamp, dim3(1), dim3(1), 0, s, t);
}

} // namespace xh
