// Internal interface between the HIP kernel files and the runtime / C-ABI layer.
// Everything here is gfx950 (CDNA4) code: wave64, 256-thread workgroups, workgroups
// mapped XCD-contiguously (8 XCDs, blockIdx dealt round-robin across them).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <type_traits>

namespace xh {

constexpr int kBlock = 256;
constexpr int kMaxNV = 64;
constexpr int kMaxPartials = 4096; // workgroups that write dot partials, at most

enum Prec : int { F64 = 0, F32 = 1 };
// u16: panel-relative column (col = base + u16); u8: compact column of a compressed block
enum Idx : int { IDX_U32 = 0, IDX_U16 = 1, IDX_U8 = 2 };

// A device CSR matrix (rows stored diagonal first, ascending columns after it).
struct Mat {
    int64_t nrows = 0, ncols = 0, nnz = 0;
    int prec = F64;
    int idx = IDX_U32;
    int panel_shift = 0;       // rows per column panel = 1 << panel_shift (IDX_U16)
    uint32_t *row = nullptr;   // nrows+1
    void *col = nullptr;       // nnz of uint32_t or uint16_t
    uint32_t *panel_base = nullptr; // one base column per panel (IDX_U16)
    void *val = nullptr;       // nnz of double or float
    void *inv_diag = nullptr;      // nrows, prec of the matrix
    void *inv_sqrt_diag = nullptr; // nrows, prec of the matrix
    int64_t fold = 0;          // dominant long-range column offset (0: none), see Sched
    // compressed block (the reference's compress(): empty rows and columns dropped):
    // row i of the block is row row_ind[i] of y, column c is column col_ind[c] + col_offset
    // of the full operand; x of the SpMV is the compact vector of ncols rows
    uint32_t *row_ind = nullptr; // nrows, or null (row i is row i)
    uint32_t *col_ind = nullptr; // ncols, or null
    int64_t col_offset = 0;
    int64_t y_rows = 0;          // rows of y (block rows before compression)
    // structured-grid view (0: none): row i = (z*gy + y)*gx + c; gcode[j] names the
    // neighbour (a, b, d) = (dz, dy, dx) of entry j as 9(a+1) + 3(b+1) + (d+1), or 255
    // for a column that is not the grid neighbour row + a*gx*gy + b*gx + d
    int32_t gx = 0, gy = 0, gz = 0;
    uint8_t *gcode = nullptr;
    // complete 7-point stencil on the grid view (every row: the diagonal and all in-grid
    // neighbours in the order -z, -y, -x, +x, +y, +z): the values again as st7[i][8],
    // slots {C, -z, -y, -x, +x, +y, +z, unused}, 0 where a neighbour is outside the grid
    void *st7 = nullptr;
    // core-block matrix of hybrid GS (entries inside each row's block of core_B rows only):
    // core_B > 0, and core_emax = the largest entry count of one block.  Each row's entry
    // list is padded to a multiple of core_kb (>= 1 batch) with pad entries of column =
    // the block's end and value +0; nnz counts the real entries only; core_rowmax = the
    // longest padded row.
    int64_t core_B = 0, core_emax = 0, core_rowmax = 0;
    int core_kb = 0;
    size_t bytes = 0;            // device bytes owned
};

// Row schedule of an SpMV launch.  The rows are viewed as a (na x fold) array (row
// r = a*fold + b) cut into column strips of `strip` rows and tiles of `tile` contiguous
// rows; tiles are numbered strip-major, then by a, then by position in the strip:
//   tile k -> strip k / (na*tps), a = (k / tps) % na, t = k % tps,
//   rows [a*fold + strip*s + t*tile, ... + tile) clipped to the strip and to nrows.
// A matrix whose rows couple mostly to rows +-fold away (a structured-grid operator,
// fold = nx*ny) then reuses each x row within one strip a few tiles after its first
// use, inside one XCD's L2, instead of fold rows later.  fold = strip = nrows, na = 1
// is plain contiguous order.  The m = grid/8 workgroups of XCD x share the tiles
// [x*m*tpw, (x+1)*m*tpw): workgroup lb takes [lb*tpw, (lb+1)*tpw), or, in launches that
// write dot partials, every m-th tile of the range (tile_of).
struct Sched {
    int64_t nrows = 0, fold = 0, strip = 0, na = 1, ntiles = 0;
    int32_t tile = 0, tps = 0, tpw = 1;
};

// rows [b, e) of tile k (empty when the tile lies past its strip or the matrix)
__host__ __device__ inline void tile_rows(const Sched &s, int64_t k, int64_t &b, int64_t &e) {
    const int64_t per_strip = s.na * s.tps;
    const int64_t st = k / per_strip;
    const int64_t rem = k - st * per_strip;
    const int64_t a = rem / s.tps;
    const int64_t t = rem - a * s.tps;
    const int64_t s0 = st * s.strip;
    const int64_t s1 = s0 + s.strip < s.fold ? s0 + s.strip : s.fold;
    const int64_t lo = s0 + t * s.tile;
    const int64_t hi = lo + s.tile < s1 ? lo + s.tile : s1;
    b = a * s.fold + lo;
    e = a * s.fold + hi;
    if (b > s.nrows)
        b = s.nrows;
    if (e > s.nrows)
        e = s.nrows;
    if (e < b)
        e = b;
}

// tile kt (0 <= kt < tpw) of logical workgroup lb in a grid of G workgroups (a multiple
// of 8): workgroup q = lb % m of XCD x = lb / m (m = G/8) takes tiles q, q + m, q + 2m, ...
// of its XCD's range, so the tiles in flight together on an XCD are consecutive and their
// x rows are shared through its L2 (with each workgroup on tpw consecutive tiles instead,
// the resident workgroups of an XCD sat a plane apart)
__host__ __device__ inline int64_t tile_of(const Sched &s, int64_t lb, int kt, int64_t G) {
    const int64_t m = G / 8, x = lb / m, q = lb - x * m;
    return (x * s.tpw + kt) * m + q;
}

struct Launch {
    int grid = 0;
    Sched sc;
};

// Work items of the structured-grid SpMV.  The gy x gz grid lines are cut into tiles of
// ty x tz lines (ty*tz = rows per workgroup pass) and each line into nseg segments of
// seglen rows; item k = (seg, zb, yb), yb fastest.  Workgroup w takes items
// [w*ipw, (w+1)*ipw) and sweeps each along x.
struct GridSched {
    int32_t nx = 0, ny = 0, nz = 0, ty = 0, tz = 0, nyb = 0, nzb = 0, nseg = 1, seglen = 0;
    int32_t nitems = 0, ipw = 1;
};
struct GridLaunch {
    int grid = 0;
    GridSched g;
};
// tile lines per workgroup pass of rpb rows: tz = 2^floor(log2(rpb)/2), ty = rpb / tz
__host__ __device__ constexpr int grid_tz(int rpb) {
    return rpb >= 256 ? 16 : rpb >= 64 ? 8 : rpb >= 16 ? 4 : rpb >= 4 ? 2 : 1;
}
GridLaunch grid_plan(int nx, int ny, int nz, int rpb, int cap);
bool csr_grid(int64_t n, const uint32_t *row, const uint32_t *col, int32_t &nx, int32_t &ny,
              int32_t &nz);
int64_t grid_codes(int64_t n, const uint32_t *row, const uint32_t *col, int32_t nx, int32_t ny,
                   int32_t nz, uint8_t *code);
// item k -> tile (yb, zb) and x range [c0, c1)
__host__ __device__ inline void grid_item(const GridSched &g, int k, int &yb, int &zb, int &c0,
                                          int &c1) {
    yb = k % g.nyb;
    const int t = k / g.nyb;
    zb = t % g.nzb;
    const int seg = t / g.nzb;
    c0 = seg * g.seglen;
    c1 = c0 + g.seglen < g.nx ? c0 + g.seglen : g.nx;
}

// Work items of the 7-point stencil SpMV: blocks of GY x GZ grid lines (one wave each),
// each line cut into nseg segments of seglen rows; item k = (seg, zb, yb), yb fastest.
// Wave w takes items w, w + nwaves, w + 2 nwaves, ...
struct StSched {
    int32_t nx = 0, ny = 0, nz = 0, nyb = 0, nzb = 0, seglen = 0, nseg = 1;
    int64_t nitems = 0;
};
struct StLaunch {
    int grid = 0;
    StSched g;
};
StLaunch st7_plan(int nx, int ny, int nz, int gy, int gz, int cap);

// Tile schedule and grid of one SpMV launch over n rows with nnz nonzeros and fold F
// (0: none): rpb rows per workgroup pass, at most cap workgroups, row_bytes bytes of x
// per row.
Launch spmv_plan(int64_t n, int64_t nnz, int64_t F, int rpb, int cap, int64_t row_bytes);

// dominant long-range column offset of a square CSR matrix: the most frequent largest
// positive (col - row) over the rows, if it is the largest one of at least a quarter
// of the rows and at least `min_fold`; else 0
int64_t csr_fold(int64_t nrows, int64_t ncols, const uint32_t *row, const uint32_t *col,
                 int64_t min_fold);

// Launch-shape knobs (xh_set_option)
// rows of at least this many entries on average take the structured-grid SpMV by default
constexpr int64_t kGridMinRowNnz = 12;

struct Tuning {
    int64_t strip_budget = 8 << 20; // bytes of x+y+A touched between reuses of an x row
    // tile rows = tile_passes * rows per workgroup pass.  Short tiles keep the rows an
    // XCD's resident workgroups hold in flight (~224 of them) within what its L2 serves:
    // the 64^3 coarse level of the 128^3 cube (26 entries per row) takes 115 us per
    // Chebyshev step with 8 passes, 99 with 2; the finest level is flat from 2 to 8
    int tile_passes = 2;
    int min_passes = 1;             // row-order schedule: passes per workgroup at least
    int dot_groups = 1024;          // workgroups of a launch that writes dot partials
    // the same for the fused vector updates with dot partials (k_vec): a pure stream,
    // which more workgroups serve better (V_XR_UPDATE at 128^3, nv = 32: 957 us with 1024,
    // 833 with 2048), unlike the SpMMs (465 against 503-521 us)
    int vec_dot_groups = 2048;
    int fold_enable = 1;            // 0: always contiguous order
    int64_t fold_min_bytes = 2 << 20; // fold only when 2*fold*nv*sizeof(value) exceeds this
    // structured-grid matrices run the LDS-window line sweep (k_spmv_grid): 0 never, 1
    // always, 2 (default) when rows average at least kGridMinRowNnz entries.  The 7-point
    // cube at 128^3 only matches the row-parallel kernel at nv = 32 (slower at other nv
    // and in fp32); the 27-point operator at 96^3 gains 10-20 %
    int grid_enable = 2;
    int grid_dot_groups = 2048;     // work items of a grid-path launch that writes partials
    // 1: a Chebyshev smoother from a zero x runs its first two steps as one SpMV
    // (OP_CHEBZ) where OP_CHEB would run row-parallel
    int chebz_enable = 1;
    // 1: hybrid GS levels keep the entries outside each row's core block as a matrix of
    // their own (built at setup) and run the off-block residual as an OP_RESID over it
    // (grid path for 27-point operators); 0: OP_OFFBLOCK over A, skipping in-block columns
    int hsgs_off_csr = 1;
    // 1: few-row, long-row SpMVs (nv 8 and 32) issue 16 x loads per batch (k_spmv<..., 16>)
    int wide_batches = 1;
    int grid_items = 2048;          // target work items (tile x segment) per launch
    int grid_min_seg = 32;          // shortest x segment of a work item
    // 1: matrices with a complete 7-point stencil view run the register-window line
    // sweep (k_spmv_st7) for nv >= 8
    int st7_enable = 0;
    int st7_waves = 3072;           // waves of one launch: 3 per SIMD of 256 CUs
    int st7_min_seg = 32;           // shortest x segment of a work item
};
extern Tuning g_tune;

// whether the structured-grid SpMV takes a matrix that has a grid view (spmv_grid_enable)
inline bool grid_on(const Mat &A) {
    return g_tune.grid_enable == 1 ||
           (g_tune.grid_enable == 2 && A.nnz >= kGridMinRowNnz * A.nrows);
}

// ---- launch accounting (xh_solver_profile) ----
// While a profile solve runs, every launch helper of the solve path brackets its kernel
// with two HIP events on its stream and records the kernel, the hierarchy level the
// runtime is at (-1: the outer Krylov method) and the launch's algorithmic HBM bytes:
// each operand the launch needs read once and each result written once (matrix values,
// column indices, row pointers and panel bases; vectors; partial sums).
bool ledger_active();
void ledger_open(hipStream_t s);
void ledger_close(hipStream_t s, const std::string &kernel, double bytes);
struct LedgerScope {
    hipStream_t s;
    bool on;
    std::string kernel;
    double bytes = 0.0;
    explicit LedgerScope(hipStream_t st) : s(st), on(ledger_active()) {
        if (on)
            ledger_open(s);
    }
    ~LedgerScope() {
        if (on)
            ledger_close(s, kernel, bytes);
    }
};
const char *prec_name(int prec);
const char *idx_name(int idx);
namespace unet {
void set_fuse_up(int on); // 0: U-Net inference materialises the upsampled tensors
void set_conv_tile(int on); // 0: no LDS-staged convolution tiles
}


// Scalar state of a Krylov solve, resident in device memory.  One slot per RHS.
struct SolverState {
    double rho[kMaxNV], rho0[kMaxNV], alpha[kMaxNV], nalpha[kMaxNV], omega[kMaxNV];
    double nomega[kMaxNV], beta[kMaxNV], ac[kMaxNV], oc[kMaxNV], pcoef[kMaxNV];
    double phi[kMaxNV], psi[kMaxNV], theta[kMaxNV], res[kMaxNV], scal1[kMaxNV];
    double iconv[kMaxNV], abs_res[kMaxNV], rel_res[kMaxNV];
    int conv[kMaxNV];
    int iters;    // reference stats.iters
    int done;     // solve finished (converged or max_iters): every later kernel is a no-op
    int conv_all; // the last convergence check found every RHS converged
    int err;      // 1: division by a zero-flagged value (reference assert)
    // parameters (copied once per solve)
    int nv, max_iters, conv_check, hist_len;
    float abs_tol, rel_tol;
    double *hist; // [hist_len][nv] sqrt(res) per convergence check, or nullptr
    unsigned long long t_done; // device wall clock (wall_clock64) when `done` was set
};

// ---- SpMV family (xh_spmv.hip) ----
enum SpmvOp : int {
    OP_SET = 0,      // y = A x                                   (csr_Ax_y)
    OP_ADD = 1,      // y += A x, accumulated into y              (csr_Axpy)
    OP_RESID = 2,    // y = b - A x  (optionally stored as fp32)  (Ax_y + axpby(1,b,-1,y))
    OP_DOT1 = 3,     // y = A x and partial (y, o1)
    OP_DOT3 = 4,     // y = A x and partials (y,o1), (y,y), (o1,o1)
    OP_CHEB0 = 5,    // r = isd*(b - A x); u = (c*r)*isd  [or x += (c*r)*isd when last]
    OP_CHEB = 6,     // t1 = c*r + (A u)*isd; u' = t1*isd [or x (+)= t1*isd when last]
    OP_PROLONG32 = 7,// x64 = (double)(P32 xH32) + x64
    OP_OFFBLOCK = 8, // y = b - (off-block part of A) x, HSGS block partition
    OP_RESID0 = 9,   // y = b - A x, b and y f64, written both f64 and group1-fused
    OP_RDOT4 = 10,   // t = A t_ ; r^ = r - alpha v (r: out2, v: b, alpha: rr) ; partials
                     // (t, r^), (t, t), (t, r0), (r^, r^) (r0: o1)  (RBiCGStab steps 11-13)
    OP_CHEBZ = 11,   // the first two Chebyshev steps from a zero-flagged x in one launch:
                     // the operand is b, each gathered entry u[c] = (c2*(b[c]*isd[c]))*isd[c]
                     // (the zero-x first step), r = b*isd written unless last, then as
                     // OP_CHEB: t1 = c*r + (A u)*isd ; u' = t1*isd [or x = t1*isd when last]
    OP_CHEBR = 12,   // OP_CHEB with r = b*isd formed from b (the steps after OP_CHEBZ,
                     // which then need not store r)
};

struct SpmvArgs {
    const void *b = nullptr;    // RESID/CHEB0/OFFBLOCK right-hand side (matrix prec)
    const void *o1 = nullptr;   // DOT1/DOT3 second vector
    const void *isd = nullptr;  // CHEB*: inverse sqrt diagonal
    const void *r = nullptr;    // CHEB: scaled residual r
    void *out2 = nullptr;       // CHEB0: r output; RESID0: p;
    void *out3 = nullptr;       // RESID0: r
    void *xupd = nullptr;       // CHEB*: x to update when `last`
    double c = 0.0;             // CHEB*: polynomial coefficient (cast to the matrix prec)
    double c2 = 0.0;            // CHEBZ: coefficient of the zero-x first step
    int last = 0;               // CHEB*: this step ends the polynomial (updates x)
    int x_zero = 0;             // CHEB*: x is zero-flagged (x = t1 instead of x + t1)
    int out_f32 = 0;            // RESID: store y as float (mixed-precision restriction)
    double *partials = nullptr; // DOT*: [grid][ndot*nv]
    const int *done = nullptr;  // skip the launch when *done
    int64_t blk_rows = 0;       // OFFBLOCK: HSGS core block size (rows)
    int grid_cap = 0;           // 0: default
};

// returns the number of partial blocks written (DOT ops) or 0
int spmv(const Mat &A, int nv, const void *x, void *y, SpmvOp op, const SpmvArgs &a,
         hipStream_t s);
// whether op on A runs on the row-parallel kernel (neither the stencil nor the grid path)
bool spmv_row_parallel(const Mat &A, SpmvOp op, int nv);

// ---- vector kernels, finalize, smoothers (xh_vec.hip) ----
enum VecOp : int {
    V_S_UPDATE = 0,   // s = r + nalpha*v
    V_XR_UPDATE = 1,  // x = ac*p_ + oc*s_ + x ; r = s - omega*t ; partial (r, r0)
    V_P_UPDATE = 2,   // p = r + pcoef*v + beta*p
    V_GROUP1 = 3,     // r0 = b - r0 ; p = r = r0 ; partial (r0, r0)
    V_X_UPDATE = 4,   // x = ac*p_ + oc*s_ + x           (BiCGStab x before check)
    V_R_UPDATE = 5,   // r = s - omega*t ; partial (r, r0)
    V_GROUP5 = 6,     // x += ac*p + oc*s ; p = r + beta*p - (beta*omega)*v   (merged)
    V_DOT3 = 7,       // partials (t,s), (t,t), (s,s) of existing vectors
    V_DOT1 = 8,       // partial (a, b)
    // PCG (pcg.inl:129-160)
    V_CG_XR = 9,      // x = ac*p + x ; r = nalpha*Ap + r        (p: p_, Ap: t)
    V_DOT2 = 10,      // partials (r, z), (r, r)                 (z: r0)
    V_CG_P = 11,      // p = z + beta*p                          (z: v)
    V_COPY_DOT = 12,  // p = z ; partial (r, z)                  (z: v)
    V_COPY = 13,      // p = v
    // RBiCGStab steps 16, 23-25 (rbicgstab.inl:186-216) and merged_rbicgstab_group2
    // (blas_merged_rbicgstab.inl:120-196): x:x, v_:p, z:s, r:r, t:t, t_:p_, s:v
    V_RB_UPDATE = 14,
    V_RB_UPDATE_MERGED = 15,
};

struct VecArgs {
    void *x = nullptr, *p = nullptr, *s = nullptr, *r = nullptr, *r0 = nullptr;
    const void *v = nullptr, *t = nullptr, *b = nullptr, *p_ = nullptr, *s_ = nullptr;
    SolverState *st = nullptr;
    double *partials = nullptr;
    int gate = 0; // 0: run unless done; 1: run only if the last check converged every RHS
};

int vec_op(VecOp op, int64_t n, int nv, const VecArgs &a, hipStream_t s);

enum FinOp : int {
    FIN_INIT = 0,   // rho0 = sum ; convergence check on (rho0, rho0) ; rho = rho0
    FIN_ALPHA = 1,  // ++iters ; scal1 = sum ; alpha = rho / scal1 ; nalpha = -alpha
    FIN_OMEGA = 2,  // phi, psi, theta ; omega ; res ; ac, oc ; convergence check
    FIN_BETA = 3,   // rho_new ; beta ; pcoef ; rho = rho_new ; max_iters check
    FIN_OMEGA_MERGED = 4, // as FIN_OMEGA, but ac/oc recomputed with the updated iconv
    FIN_CG_RHO = 5,   // rho = (r, z)                                        (PCG start)
    FIN_CG_ALPHA = 6, // ++iters ; alpha = rho / (p, Ap) ; nalpha ; ac
    FIN_CG_BETA = 7,  // beta = (r,z)/rho ; rho = (r,z) ; res = (r,r) ; check ; max_iters
    FIN_RB_OMEGA = 8, // theta, phi, psi, eta ; omega ; res ; ac, oc ; check ; rho ; beta
    FIN_RB_OMEGA_MERGED = 9, // as FIN_RB_OMEGA, ac/oc recomputed with the updated iconv
    FIN_END = 10,     // done when the last check converged every RHS or at max_iters
};
// dots reduced by each finalize op
__host__ __device__ inline int fin_ndots(int op) {
    return (op == FIN_OMEGA || op == FIN_OMEGA_MERGED) ? 3
           : (op == FIN_CG_BETA)                       ? 2
           : (op == FIN_RB_OMEGA || op == FIN_RB_OMEGA_MERGED) ? 4
                                                       : 1;
}
void finalize(FinOp op, SolverState *st, const double *partials, int nparts, int nv,
              hipStream_t s);

// smoother / transfer helpers
// `done` (nullable): device flag; when set the launch is a no-op
// bump right-hand sides on the device (count[nv], P[nv][maxb][5] device arrays)
void eval_bumps(int nx, int ny, int nz, int nv, int maxb, const int *count, const double *P,
                double *b, hipStream_t s);

// error message for xh_last_error(); returns `code`
int report_error(int code, const std::string &msg);
// selects device 0 on first use
int runtime_init();

// ---- batched independent systems (xh_batch.hip) ----
constexpr int kChebMax = 16;          // Chebyshev coefficients per system
constexpr int kBatchPart = 3 * 16;    // dot partial stride per work item (3 dots x nv <= 16)
constexpr int kBatchMaxNV = 16;

struct BatchItem {        // rows [r0, r1) of system `sys`, which has nv RHS
    int32_t sys, nv;
    int64_t r0, r1;
    int32_t wpl;           // RHS per lane (1 or 2: nv/wpl lanes per row)
    int32_t slot;          // the item's dot-partial slot (its index in the full item list)
};

struct BatchDev {
    const uint32_t *row = nullptr;   // concatenated row pointers (global nonzero offsets)
    const uint32_t *col = nullptr;   // columns local to the system
    const double *val = nullptr;
    const double *isd = nullptr;     // inverse square-root diagonal per global row
    const int64_t *rbase = nullptr;  // first global row of each system
    const int64_t *voff = nullptr;   // first multivector element of each system
    SolverState *st = nullptr;       // one state per system
    const double *cheb = nullptr;    // kChebMax coefficients per system
};

// multivector operands of one batched launch (all in the concatenated layout)
struct BatchVecs {
    const double *in = nullptr;  // SpMV operand
    const double *b = nullptr;   // RESID0, CHEB0Z: b; COPY: source; S_/P_UPDATE: r
    const double *o1 = nullptr;  // DOT1, XR_UPDATE: r0; DOT3: s; CHEB: b; S_/P_UPDATE: v
    double *out = nullptr;       // result (P_UPDATE: p, read and written)
    double *out2 = nullptr;      // RESID0: p; CHEB0Z/CHEB with `last`: x
    double *out3 = nullptr;      // RESID0: r
    double *x = nullptr;         // XR_UPDATE
    const double *p_ = nullptr, *s_ = nullptr, *s = nullptr, *t = nullptr; // XR_UPDATE
};

enum BatchOp : int {
    B_RESID0 = 0,    // out = b - A in ; out2 = out3 = out ; (out, out)
    B_DOT1 = 1,      // out = A in ; (out, o1)
    B_DOT3 = 2,      // out = A in ; (out, o1), (out, out), (o1, o1)
    B_CHEB0Z = 3,    // first Chebyshev step from a zero x (coefficient ci)
    B_CHEB = 4,      // one Horner step (coefficient ci) with r = o1*isd formed from o1 = b,
                     // x = t1*isd when `last`
    B_COPY = 5,      // out = b
    B_S_UPDATE = 6,  // out = b + nalpha o1
    B_XR_UPDATE = 7, // x = ac p_ + oc s_ + x ; out = s - omega t ; (out, o1)
    B_P_UPDATE = 8,  // out = b + pcoef o1 + beta out
    B_CHEBZ = 9,     // B_CHEB0Z (coefficient ci + 1) and the B_CHEB step after it
                     // (coefficient ci) in one launch: gathers in = b, each entry as
                     // (c'*(b*isd))*isd
};

// Operands of the single-launch batched solve (k_batch_solve): the systems, their work
// items and item ranges, the solver vectors, and the preconditioner (Chebyshev order, or 0
// for a copy).
// arrival counter and error flag of the single-launch solve's grid barriers (zeroed
// before each launch)
struct BatchBarrier {
    unsigned long long arrived;
    int err;
};
struct BatchSolve {
    BatchDev d;
    BatchBarrier *bar = nullptr;
    unsigned long long wait_ticks = 0; // longest barrier wait, wall-clock ticks
    const BatchItem *items = nullptr;
    int nitems = 0, nsys = 0, max_iters = 0, pre = 0, order = 0;
    const int *item_start = nullptr, *item_count = nullptr, *nvs = nullptr;
    double *partials = nullptr;
    double *x = nullptr, *r0 = nullptr, *r = nullptr, *p = nullptr, *v = nullptr, *s = nullptr;
    double *t = nullptr, *p_ = nullptr, *s_ = nullptr, *cu1 = nullptr, *cu2 = nullptr;
    const double *b = nullptr;
};
// workgroups of k_batch_solve that are resident at once (0 when the query fails)
int batch_solve_blocks();
hipError_t batch_solve_launch(const BatchSolve &P, int grid, hipStream_t s);

// nact (nullable, device): the launch walks only the first *nact of the nitems items
void batch_launch(BatchOp op, const BatchDev &d, const BatchItem *items, int nitems, int grid,
                  const BatchVecs &v, int ci, int last, double *partials, hipStream_t s,
                  const int *nact = nullptr);
// items_act / *nact = the items of the systems not done, in item order
void batch_compact(const BatchItem *items, const int *item_start, const int *item_count,
                   const SolverState *st, int nsys, BatchItem *items_act, int *nact,
                   hipStream_t s);
void batch_finalize(FinOp op, SolverState *st, const double *partials, const int *item_start,
                    const int *item_count, const int *nvs, int nsys, hipStream_t s);
// flag[0] = every system done, flag[1] = systems not done
void batch_all_done(const SolverState *st, int nsys, int *flag, hipStream_t s);
void stamp_clock(unsigned long long *t, hipStream_t s);

// x_c[c][k] = x_full[col_ind[c] + off][k], c < nc (the compressed block's column decode)
void gather_cols(int64_t nc, int nv, int prec, const uint32_t *col_ind, int64_t off,
                 const void *x_full, void *x_c, hipStream_t s);
void cheb_init_zero(int64_t n, int nv, int prec, const void *b, const void *isd, double c,
                    void *r, void *u_or_x, int last, const int *done, hipStream_t s);
void jacobi_update(int64_t n, int nv, int prec, void *x, const void *b, const void *r,
                   const void *inv_diag, double relax, int x_zero, const int *done,
                   hipStream_t s);
void hsgs_sweeps(const Mat &A, int nv, const void *t, void *x, int x_zero, double relax,
                 int64_t blk_rows, const int *done, hipStream_t s);
int64_t hsgs_block_rows(int64_t n, int nblocks);
void dense_matvec(int64_t n, int nv, int prec, const void *D, const void *x, void *y,
                  const int *done, hipStream_t s);
void convert(int64_t count, int from_prec, const void *src, void *dst, const int *done,
             hipStream_t s);
void copy_vec(int64_t count, int prec, const void *src, void *dst, const int *done,
              hipStream_t s);
// x = u + x
void add_vec(int64_t count, int prec, const void *u, void *x, const int *done, hipStream_t s);
void fill_zero(int64_t count, int prec, void *dst, const int *done, hipStream_t s);

// ---- lane helpers shared by the kernels ----

// Broadcast of lane S of each L-lane group (L consecutive lanes, L | 64) to the whole
// group, S a compile-time constant.  DPP (a VALU operand modifier, no LDS traffic):
// quad_perm for L = 2, 4; row_newbcast (gfx90a+: every lane of a 16-lane row reads
// lane S of that row) for L = 16, and two of them merged for L = 8.  L = 32 goes
// through ds_bpermute.  The lanes of a group share one row, so a group is active or
// inactive as a whole and every source lane read by an active lane is active: the DPP
// moves need no "old" operand.
template <int L, int S>
__device__ __forceinline__ int bcast_i32(int v) {
    static_assert(S >= 0 && S < L, "source lane outside the group");
    if constexpr (L == 2) {
        constexpr int q = S | (S << 2) | ((2 + S) << 4) | ((2 + S) << 6);
        return __builtin_amdgcn_mov_dpp(v, q, 0xF, 0xF, false);
    } else if constexpr (L == 4) {
        constexpr int q = S | (S << 2) | (S << 4) | (S << 6);
        return __builtin_amdgcn_mov_dpp(v, q, 0xF, 0xF, false);
    } else if constexpr (L == 16) {
        return __builtin_amdgcn_mov_dpp(v, 0x150 + S, 0xF, 0xF, false);
    } else if constexpr (L == 8) {
        const int lo = __builtin_amdgcn_mov_dpp(v, 0x150 + S, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_mov_dpp(v, 0x158 + S, 0xF, 0xF, false);
        return (threadIdx.x & 8) ? hi : lo;
    } else {
        const int lane = threadIdx.x & 63;
        return __shfl(v, lane - (lane % L) + S, 64);
    }
}

template <int L, int S, typename T>
__device__ __forceinline__ T bcast(T v) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, bcast_i32<L, S>(__builtin_bit_cast(int, v)));
    } else {
        static_assert(sizeof(T) == 8, "4- or 8-byte values");
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const int lo = bcast_i32<L, S>((int)(unsigned)u);
        const int hi = bcast_i32<L, S>((int)(unsigned)(u >> 32));
        return __builtin_bit_cast(T, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
    }
}

// compile-time loop: f(std::integral_constant<int, 0>), ..., f(<N-1>)
template <int I, int N, typename F>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>());
        static_for<I + 1, N>(f);
    }
}

} // namespace xh
