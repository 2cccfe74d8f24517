// Device-side Krylov scalar logic shared by the single-system and the batched solvers:
// the reference's convergence test and the BiCGStab scalar recurrences, evaluated by one
// workgroup from deterministic per-workgroup dot partials.
//
// References: base_solver::convergence_check / converged (src/solvers/solver.h:173-241),
// xdivy_z / ax_y / axpby on [1][NV] vectors (src/blas/detail/blas_ext.inl:194-239,
// src/blas/detail/blas.inl:31-151), the scalar steps of PBiCGStab / MergedPBiCGStab
// (src/solvers/detail/pbicgstab.inl:141-230, merged/merged_pbicgstab.inl:129-221), PCG
// (src/solvers/detail/pcg.inl:129-160) and RBiCGStab / MergedRBiCGStab
// (src/solvers/detail/rbicgstab.inl:133-216, merged/merged_rbicgstab.inl:133-216).
#pragma once

#include "xh_internal.hpp"

namespace xh {


__device__ inline void conv_check(SolverState *st, const double *res, const double *res0, int k,
                           int nv, int *s_notconv) {
    // base_solver::convergence_check (solver.h:173-203), tolerances are float32
    if (k < nv && st->iconv[k] != 0.0) {
        const double a = sqrt(res[k]);
        const double rr = a / sqrt(res0[k]);
        if ((a > (double)st->abs_tol) && (rr > (double)st->rel_tol)) {
            atomicOr(s_notconv, 1);
        } else if (!st->conv[k]) {
            st->abs_res[k] = a;
            st->rel_res[k] = rr;
            st->conv[k] = 1;
        }
    }
}

// One workgroup of 256 threads turns the dot partials of one solve into the Krylov scalars
// and the convergence decision.  partials: nparts blocks, block g at partials[g * stride],
// holding nd*nv values (dot d, RHS k at d*nv + k), summed in block order.
__device__ inline void finalize_body(int op, SolverState *__restrict__ st,
                                     const double *__restrict__ partials, int nparts, int nv,
                                     int stride) {
    if (st->done)
        return;
    const int nd = fin_ndots(op);
    const int C = nd * nv;
    const int S = 256 / C; // >= 1 since nd*nv <= 4*64 = 256
    __shared__ double red[256];
    __shared__ double tot[4 * kMaxNV];
    __shared__ int s_notconv;
    const int tid = threadIdx.x;
    if (tid == 0)
        s_notconv = 0;
    if (tid < C * S) {
        const int c = tid % C, sub = tid / C;
        double acc = 0.0;
        // unrolled so that many partial loads are in flight; the additions keep their order
#pragma unroll 16
        for (int g = sub; g < nparts; g += S)
            acc = acc + partials[(size_t)g * stride + c];
        red[sub * C + c] = acc;
    }
    __syncthreads();
    if (tid < C) {
        double acc = red[tid];
        for (int sub = 1; sub < S; ++sub)
            acc = acc + red[sub * C + tid];
        tot[tid] = acc;
    }
    __syncthreads();
    const int k = tid;
    const bool act = k < nv;
    if (op == FIN_INIT) {
        if (act) {
            st->rho0[k] = tot[k];
            if (st->hist && st->hist_len > 0)
                st->hist[k] = sqrt(tot[k]);
        }
        __syncthreads();
        if (st->conv_check)
            conv_check(st, st->rho0, st->rho0, k, nv, &s_notconv);
        __syncthreads();
        if (!st->conv_check) {
            if (act)
                st->rho[k] = st->rho0[k];
            return;
        }
        if (!s_notconv) {
            if (tid == 0) {
                st->conv_all = 1;
                st->done = 1;
                st->t_done = wall_clock64();
            }
            return;
        }
        if (act) {
            st->iconv[k] = st->iconv[k] * (double)(!st->conv[k]);
            st->rho[k] = st->rho0[k];
        }
        return;
    }
    if (op == FIN_ALPHA) {
        if (tid == 0)
            st->iters += 1;
        if (act) {
            st->scal1[k] = tot[k];
            const double al = st->rho[k] / tot[k]; // xdivy_z(rho, scal1, alpha)
            st->alpha[k] = al;
            st->nalpha[k] = al * -1.0;
        }
        return;
    }
    if (op == FIN_OMEGA || op == FIN_OMEGA_MERGED) {
        if (act) {
            const double ph = tot[k], ps = tot[nv + k], th = tot[2 * nv + k];
            st->phi[k] = ph;
            st->psi[k] = ps;
            st->theta[k] = th;
            const double om = ph / ps;      // xdivy_z(phi, psi, omega)
            st->omega[k] = om;
            st->nomega[k] = om * -1.0;
            const double wr = om * ph;      // ax_y(omega, phi, res)
            st->res[k] = th + -1.0 * wr;    // axpby(a1, theta, a_1, res)
            st->ac[k] = st->iconv[k] * st->alpha[k];
            st->oc[k] = st->iconv[k] * om;
            const int it = st->iters;
            if (st->hist && it < st->hist_len)
                st->hist[(size_t)it * nv + k] = sqrt(st->res[k]);
        }
        __syncthreads();
        if (!st->conv_check)
            return;
        conv_check(st, st->res, st->rho0, k, nv, &s_notconv);
        __syncthreads();
        if (!s_notconv) {
            if (tid == 0)
                st->conv_all = 1;
            return;
        }
        if (act) {
            st->iconv[k] = st->iconv[k] * (double)(!st->conv[k]);
            if (op == FIN_OMEGA_MERGED) {
                st->ac[k] = st->iconv[k] * st->alpha[k];
                st->oc[k] = st->iconv[k] * st->omega[k];
            }
        }
        return;
    }
    if (op == FIN_CG_RHO) {
        if (act)
            st->rho[k] = tot[k]; // dot(r, z, scal1) of the first iteration
        return;
    }
    if (op == FIN_CG_ALPHA) {
        if (tid == 0)
            st->iters += 1;
        if (act) {
            const double al = st->rho[k] / tot[k]; // xdivy_z(scal1, scal2, alpha)
            st->alpha[k] = al;
            st->ac[k] = st->iconv[k] * al;         // ax_y(iconv, alpha, alpha_conv)
            st->nalpha[k] = al * -1.0;             // scal(a_1, alpha)
        }
        return;
    }
    if (op == FIN_CG_BETA) {
        if (act) {
            const double s3 = tot[k];
            st->beta[k] = s3 / st->rho[k]; // xdivy_z(scal3, scal1, beta)
            st->rho[k] = s3;               // next iteration's dot(r, z, scal1)
            st->res[k] = tot[nv + k];      // dot(r, r, res)
            const int it = st->iters;
            if (st->hist && it < st->hist_len)
                st->hist[(size_t)it * nv + k] = sqrt(st->res[k]);
        }
        __syncthreads();
        if (st->conv_check) {
            conv_check(st, st->res, st->rho0, k, nv, &s_notconv);
            __syncthreads();
            if (!s_notconv) {
                if (tid == 0) {
                    st->conv_all = 1;
                    st->done = 1;
                    st->t_done = wall_clock64();
                }
                return;
            }
            if (act)
                st->iconv[k] = st->iconv[k] * (double)(!st->conv[k]);
        }
        if (tid == 0 && st->iters >= st->max_iters) {
            st->done = 1;
            st->t_done = wall_clock64();
        }
        return;
    }
    if (op == FIN_RB_OMEGA || op == FIN_RB_OMEGA_MERGED) {
        if (act) {
            const double th = tot[k], ph = tot[nv + k], ps = tot[2 * nv + k],
                         et = tot[3 * nv + k];
            st->theta[k] = th;
            st->phi[k] = ph;
            st->psi[k] = ps;
            const double om = th / ph;   // xdivy_z(theta, phi, omega)
            st->omega[k] = om;
            st->nomega[k] = om * -1.0;
            const double wr = om * th;   // ax_y(omega, theta, res)
            st->res[k] = et + -1.0 * wr; // axpby(a1, eta, a_1, res)
            st->ac[k] = st->iconv[k] * st->alpha[k];
            st->oc[k] = st->iconv[k] * om;
            const int it = st->iters;
            if (st->hist && it < st->hist_len)
                st->hist[(size_t)it * nv + k] = sqrt(st->res[k]);
            // steps 21-22, used only when the solve goes on
            double rh = om * ps;         // ax_y(omega, psi, rho)
            rh = rh * -1.0;              // scal(a_1, rho)
            const double tp = st->scal1[k] * om; // ax_y(delta, omega, temp)
            st->beta[k] = rh / tp;       // xdivy_z(rho, temp, beta)
            st->rho[k] = rh;
        }
        __syncthreads();
        if (!st->conv_check)
            return;
        conv_check(st, st->res, st->rho0, k, nv, &s_notconv);
        __syncthreads();
        if (!s_notconv) {
            if (tid == 0)
                st->conv_all = 1;
            return;
        }
        if (act) {
            st->iconv[k] = st->iconv[k] * (double)(!st->conv[k]);
            if (op == FIN_RB_OMEGA_MERGED) {
                st->ac[k] = st->iconv[k] * st->alpha[k];
                st->oc[k] = st->iconv[k] * st->omega[k];
            }
        }
        return;
    }
    if (op == FIN_END) {
        if (tid == 0 && (st->conv_all || st->iters >= st->max_iters)) {
            st->done = 1;
            st->t_done = wall_clock64();
        }
        return;
    }
    if (op == FIN_BETA) {
        if (st->conv_all) {
            if (tid == 0) {
                st->done = 1;
                st->t_done = wall_clock64();
            }
            return;
        }
        if (act) {
            const double rn = tot[k];
            const double s2 = rn / st->rho[k];          // xdivy_z(scal1, rho, scal2)
            const double s3 = st->alpha[k] / st->omega[k]; // xdivy_z(alpha, omega, scal3)
            const double be = s2 * s3;                  // ax_y(scal2, scal3, beta)
            st->beta[k] = be;
            st->rho[k] = rn;                            // copy(scal1, rho)
            const double ob = st->omega[k] * be;        // ax_y(omega, beta, scal1)
            st->pcoef[k] = ob * -1.0;                   // scal(a_1, scal1)
        }
        if (tid == 0 && st->iters >= st->max_iters) {
            st->done = 1;
            st->t_done = wall_clock64();
        }
        return;
    }
}

} // namespace xh
