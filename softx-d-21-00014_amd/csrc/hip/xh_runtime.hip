// C ABI of libxamg_hip.so: device objects, hierarchy upload and the GPU solve driver.
//
// The solve driver restates the reference solve loops
//   PBiCGStab::solve        src/solvers/detail/pbicgstab.inl:44-237
//   MergedPBiCGStab::solve  src/solvers/detail/merged/merged_pbicgstab.inl:44-225
//   BiCGStab::solve         src/solvers/detail/bicgstab.inl:44-221
//   MergedBiCGStab::solve   src/solvers/detail/merged/merged_bicgstab.inl:44-208
//   MultiGrid::MG_cycle     src/solvers/detail/mg.inl:166-260
//   Chebyshev / Jacobi / HSGS / Direct smoothers (chebyshev.inl, jacobi.inl, hsgs.inl,
//   direct.inl)
// as a stream of fused kernels.  All Krylov scalars live in device memory
// (SolverState) and the convergence test runs in a one-workgroup kernel, so the host
// never waits inside an iteration: it enqueues whole iterations (or replays one
// captured hipGraph per iteration) and polls the device `done` flag one iteration
// behind.  Kernels launched after `done` is set return immediately, so x is frozen at
// exactly the iteration where the reference returns.
#include "../../../include/xamg_hip.h"
#include "xh_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

// Staging area of xh_eval_bumps_noalloc: the per-RHS bump counts (int32) in the first
// kScratchCountBytes bytes, then the bump parameters (5 doubles per bump, nv * max_bumps
// bumps, nv and max_bumps at most 64).  Every GPU the library runs on holds one copy until
// the process ends; host code reaches it only through the per-device address that xh_init
// looks up (g_scratch), never through this name.
constexpr int kScratchMaxNv = 64;
constexpr int kScratchMaxBumps = 64;
constexpr size_t kScratchCountBytes = 256;
constexpr size_t kScratchBytes =
    kScratchCountBytes + (size_t)kScratchMaxNv * kScratchMaxBumps * 5 * sizeof(double);
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace xh {
namespace {

thread_local std::string g_err;
int g_device = -1;
hipStream_t g_stream = nullptr;
// device address of `scratch` on each device, set by xh_init with that device current
std::vector<void *> g_scratch;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define XH_HIP(call)                                                                              \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(XH_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));              \
    } while (0)

bool nv_ok(int nv) {
    return nv == 1 || nv == 2 || nv == 4 || nv == 8 || nv == 16 || nv == 32 || nv == 64;
}

int ensure_init() {
    if (g_device >= 0)
        return 0;
    return xh_init(0);
}

// Launch records of a profile solve: kernel, level, iteration, algorithmic bytes and the
// two events around the launch.  Events come from a pool that lives as long as the ledger.
struct LedgerRec {
    std::string kernel;
    int level, iter;
    double bytes;
    hipEvent_t e0, e1;
};
struct Ledger {
    std::vector<LedgerRec> recs;
    std::vector<hipEvent_t> pool;
    size_t next = 0;
    int level = -1, iter = 0;
    hipEvent_t open = nullptr;
    bool failed = false;
    hipEvent_t take() {
        if (next == pool.size()) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) {
                failed = true;
                return nullptr;
            }
            pool.push_back(e);
        }
        return pool[next++];
    }
    ~Ledger() {
        for (hipEvent_t e : pool)
            (void)hipEventDestroy(e);
    }
};
Ledger *g_ledger = nullptr;

void ledger_level(int l) {
    if (g_ledger)
        g_ledger->level = l;
}

} // namespace

// shared with the other sources of the library (xh_internal.hpp)
int report_error(int code, const std::string &msg) { return fail(code, msg); }
int runtime_init() { return ensure_init(); }

bool ledger_active() { return g_ledger != nullptr; }

void ledger_open(hipStream_t s) {
    hipEvent_t e = g_ledger->take();
    if (e && hipEventRecord(e, s) != hipSuccess)
        g_ledger->failed = true;
    g_ledger->open = e;
}

void ledger_close(hipStream_t s, const std::string &kernel, double bytes) {
    Ledger &L = *g_ledger;
    hipEvent_t e = L.take();
    if (!e || !L.open || hipEventRecord(e, s) != hipSuccess) {
        L.failed = true;
        return;
    }
    L.recs.push_back({kernel, L.level, L.iter, bytes, L.open, e});
    L.open = nullptr;
}

namespace {

// ---- host CSR ----
struct HostCsr {
    int64_t nrows = 0, ncols = 0;
    std::vector<uint32_t> row, col;
    std::vector<double> val;
};

// largest panel shift (rows per panel = 2^s, s >= 6) whose every panel spans at most
// 65536 columns; -1 when none does
int find_panel_shift(const HostCsr &m) {
    for (int s = 20; s >= 6; --s) {
        const int64_t pr = (int64_t)1 << s;
        bool ok = true;
        for (int64_t p0 = 0; p0 < m.nrows && ok; p0 += pr) {
            const int64_t p1 = std::min(m.nrows, p0 + pr);
            uint32_t lo = UINT32_MAX, hi = 0;
            for (uint32_t j = m.row[p0]; j < m.row[p1]; ++j) {
                lo = std::min(lo, m.col[j]);
                hi = std::max(hi, m.col[j]);
            }
            if (m.row[p1] > m.row[p0] && (uint64_t)hi - lo > 65535u)
                ok = false;
        }
        if (ok)
            return s;
    }
    return -1;
}

} // namespace

int64_t csr_fold(int64_t nrows, int64_t ncols, const uint32_t *row, const uint32_t *col,
                 int64_t min_fold) {
    if (nrows != ncols || nrows <= 0)
        return 0;
    std::unordered_map<int64_t, int64_t> cnt;
    for (int64_t i = 0; i < nrows; ++i) {
        int64_t mx = 0;
        for (uint32_t j = row[i]; j < row[i + 1]; ++j)
            mx = std::max<int64_t>(mx, (int64_t)col[j] - i);
        if (mx > 0)
            ++cnt[mx];
    }
    int64_t best = 0, bc = 0;
    for (const auto &kv : cnt)
        if (kv.second > bc || (kv.second == bc && kv.first > best)) {
            best = kv.first;
            bc = kv.second;
        }
    return (bc * 4 >= nrows && best >= min_fold) ? best : 0;
}

// Grid view of a square matrix: rows = points of an nx x ny x nz grid in natural order.
// nx and F = nx*ny are the most frequent pair of column offsets (sampled rows) with
// F % nx == 0, nx >= 3, F / nx >= 3, n % F == 0 and each offset on at least a quarter
// of the rows; both are the most frequent offsets of their neighbourhoods in a 7- or
// 27-point operator.  False when no pair qualifies.
bool csr_grid(int64_t n, const uint32_t *row, const uint32_t *col, int32_t &nx, int32_t &ny,
              int32_t &nz) {
    if (n < 27)
        return false;
    std::unordered_map<int64_t, int64_t> cnt;
    const int64_t step = std::max<int64_t>(1, n / 65536);
    int64_t sampled = 0;
    for (int64_t i = 0; i < n; i += step, ++sampled)
        for (uint32_t j = row[i]; j < row[i + 1]; ++j) {
            const int64_t d = (int64_t)col[j] - i;
            if (d >= 3)
                ++cnt[d];
        }
    std::vector<std::pair<int64_t, int64_t>> top(cnt.begin(), cnt.end());
    std::sort(top.begin(), top.end(), [](const auto &p, const auto &q) {
        return p.second != q.second ? p.second > q.second : p.first < q.first;
    });
    if (top.size() > 16)
        top.resize(16);
    int64_t bx = 0, bF = 0, bs = 0;
    for (const auto &p : top)
        for (const auto &q : top) {
            const int64_t a = p.first, F = q.first;
            // both offsets on at least a quarter of the rows, as in a grid operator
            if (F <= a || F % a != 0 || F / a < 3 || n % F != 0 || F > INT32_MAX ||
                4 * p.second < sampled || 4 * q.second < sampled)
                continue;
            const int64_t sc = p.second + q.second;
            if (sc > bs) {
                bs = sc;
                bx = a;
                bF = F;
            }
        }
    if (bs == 0)
        return false;
    nx = (int32_t)bx;
    ny = (int32_t)(bF / bx);
    nz = (int32_t)(n / bF);
    return true;
}

// gcode of every entry for the grid view (Mat::gcode); returns the coded count
int64_t grid_codes(int64_t n, const uint32_t *row, const uint32_t *col, int32_t nx, int32_t ny,
                   int32_t nz, uint8_t *code) {
    const int64_t F = (int64_t)nx * ny;
    int64_t coded = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t c = i % nx, y = (i / nx) % ny, z = i / F;
        for (uint32_t j = row[i]; j < row[i + 1]; ++j) {
            const int64_t dl = (int64_t)col[j] - i;
            uint8_t cd = 255;
            for (int a = -1; a <= 1 && cd == 255; ++a)
                for (int b = -1; b <= 1 && cd == 255; ++b)
                    for (int d = -1; d <= 1; ++d)
                        if (dl == a * F + b * (int64_t)nx + d && c + d >= 0 && c + d < nx &&
                            y + b >= 0 && y + b < ny && z + a >= 0 && z + a < nz) {
                            cd = (uint8_t)(9 * (a + 1) + 3 * (b + 1) + (d + 1));
                            break;
                        }
            code[j] = cd;
            coded += cd != 255;
        }
    }
    return coded;
}

namespace {

template <typename T>
int dmalloc(T **p, size_t count, size_t &acc) {
    if (count == 0) {
        *p = nullptr;
        return 0;
    }
    XH_HIP(hipMalloc((void **)p, count * sizeof(T)));
    acc += count * sizeof(T);
    return 0;
}

// The st7 value layout of a grid matrix (gx x gy x gz, entry codes `code`) when every row
// holds exactly the diagonal and its in-grid neighbours in the order -z, -y, -x, +x, +y,
// +z: v8[i*8 + s], slot s of {C, -z, -y, -x, +x, +y, +z, unused}, 0 for a neighbour
// outside the grid.  Returns false for any other matrix.
bool stencil7_values(const HostCsr &h, int32_t gx, int32_t gy, int32_t gz, const uint8_t *code,
                     std::vector<double> &v8) {
    // entry code 9(dz+1) + 3(dy+1) + (dx+1) of each slot
    static const uint8_t slot_code[7] = {13, 4, 10, 12, 14, 16, 22};
    const int64_t n = h.nrows;
    if ((int64_t)gx * gy * gz != n)
        return false;
    v8.assign((size_t)n * 8, 0.0);
    for (int64_t i = 0; i < n; ++i) {
        const int64_t cx = i % gx, cy = (i / gx) % gy, cz = i / ((int64_t)gx * gy);
        const bool present[7] = {true,        cz > 0,      cy > 0,     cx > 0,
                                 cx + 1 < gx, cy + 1 < gy, cz + 1 < gz};
        uint32_t j = h.row[i];
        for (int s2 = 0; s2 < 7; ++s2) {
            if (!present[s2])
                continue;
            if (j >= h.row[i + 1] || code[j] != slot_code[s2])
                return false;
            v8[(size_t)i * 8 + s2] = h.val[j];
            ++j;
        }
        if (j != h.row[i + 1])
            return false;
    }
    return true;
}

int upload_mat(const HostCsr &h, int prec, int index_mode, Mat &M) {
    M = Mat();
    M.nrows = h.nrows;
    M.ncols = h.ncols;
    M.nnz = (int64_t)h.col.size();
    M.prec = prec;
    const int64_t n = h.nrows, nnz = M.nnz;
    int shift = -1;
    if (index_mode != 1)
        shift = find_panel_shift(h);
    if (index_mode == 2 && shift < 0)
        return fail(XH_EINVAL, "u16 column panels impossible for this matrix");
    M.idx = shift >= 0 ? IDX_U16 : IDX_U32;
    M.panel_shift = shift >= 0 ? shift : 0;
    size_t acc = 0;
    int rc;
    if ((rc = dmalloc(&M.row, n + 1, acc)))
        return rc;
    XH_HIP(hipMemcpy(M.row, h.row.data(), (n + 1) * 4, hipMemcpyHostToDevice));
    if (M.idx == IDX_U16) {
        const int64_t pr = (int64_t)1 << shift;
        const int64_t np = (n + pr - 1) / pr;
        std::vector<uint32_t> base(np, 0);
        std::vector<uint16_t> c16(nnz);
        for (int64_t p = 0; p < np; ++p) {
            const int64_t p0 = p * pr, p1 = std::min(n, p0 + pr);
            uint32_t lo = UINT32_MAX;
            for (uint32_t j = h.row[p0]; j < h.row[p1]; ++j)
                lo = std::min(lo, h.col[j]);
            base[p] = lo == UINT32_MAX ? 0 : lo;
            for (uint32_t j = h.row[p0]; j < h.row[p1]; ++j)
                c16[j] = (uint16_t)(h.col[j] - base[p]);
        }
        uint16_t *dc = nullptr;
        if ((rc = dmalloc(&dc, nnz, acc)) || (rc = dmalloc(&M.panel_base, np, acc)))
            return rc;
        M.col = dc;
        if (nnz)
            XH_HIP(hipMemcpy(dc, c16.data(), nnz * 2, hipMemcpyHostToDevice));
        XH_HIP(hipMemcpy(M.panel_base, base.data(), np * 4, hipMemcpyHostToDevice));
    } else {
        uint32_t *dc = nullptr;
        if ((rc = dmalloc(&dc, nnz, acc)))
            return rc;
        M.col = dc;
        if (nnz)
            XH_HIP(hipMemcpy(dc, h.col.data(), nnz * 4, hipMemcpyHostToDevice));
    }
    // values and diagonal inverses in the matrix precision
    // (csr_matrix::alloc_inv_diag / alloc_inv_sqrt_diag, csr_matrix.h:302-368)
    if (prec == F64) {
        std::vector<double> id(n, 0.0), isd(n, 0.0);
        for (int64_t i = 0; i < n; ++i)
            for (uint32_t j = h.row[i]; j < h.row[i + 1]; ++j)
                if (h.col[j] == (uint32_t)i) {
                    id[i] = 1.0 / h.val[j];
                    isd[i] = std::sqrt(1.0 / h.val[j]);
                }
        double *v, *a, *b;
        if ((rc = dmalloc(&v, nnz, acc)) || (rc = dmalloc(&a, n, acc)) ||
            (rc = dmalloc(&b, n, acc)))
            return rc;
        if (nnz)
            XH_HIP(hipMemcpy(v, h.val.data(), nnz * 8, hipMemcpyHostToDevice));
        XH_HIP(hipMemcpy(a, id.data(), n * 8, hipMemcpyHostToDevice));
        XH_HIP(hipMemcpy(b, isd.data(), n * 8, hipMemcpyHostToDevice));
        M.val = v;
        M.inv_diag = a;
        M.inv_sqrt_diag = b;
    } else {
        std::vector<float> vf(nnz), id(n, 0.f), isd(n, 0.f);
        for (int64_t j = 0; j < nnz; ++j)
            vf[j] = (float)h.val[j];
        for (int64_t i = 0; i < n; ++i)
            for (uint32_t j = h.row[i]; j < h.row[i + 1]; ++j)
                if (h.col[j] == (uint32_t)i) {
                    id[i] = (float)(1.0 / (double)vf[j]);
                    isd[i] = (float)std::sqrt(1.0 / (double)vf[j]);
                }
        float *v, *a, *b;
        if ((rc = dmalloc(&v, nnz, acc)) || (rc = dmalloc(&a, n, acc)) ||
            (rc = dmalloc(&b, n, acc)))
            return rc;
        if (nnz)
            XH_HIP(hipMemcpy(v, vf.data(), nnz * 4, hipMemcpyHostToDevice));
        XH_HIP(hipMemcpy(a, id.data(), n * 4, hipMemcpyHostToDevice));
        XH_HIP(hipMemcpy(b, isd.data(), n * 4, hipMemcpyHostToDevice));
        M.val = v;
        M.inv_diag = a;
        M.inv_sqrt_diag = b;
    }
    if (h.nrows == h.ncols)
        M.fold = csr_fold(h.nrows, h.ncols, h.row.data(), h.col.data(), 256);
    // structured-grid view when every entry is a grid neighbour
    int32_t gx, gy, gz;
    if (h.nrows == h.ncols && nnz > 0 && csr_grid(n, h.row.data(), h.col.data(), gx, gy, gz)) {
        std::vector<uint8_t> code(nnz);
        const int64_t coded = grid_codes(n, h.row.data(), h.col.data(), gx, gy, gz, code.data());
        if (coded == nnz) {
            if ((rc = dmalloc(&M.gcode, nnz, acc)))
                return rc;
            XH_HIP(hipMemcpy(M.gcode, code.data(), nnz, hipMemcpyHostToDevice));
            M.gx = gx;
            M.gy = gy;
            M.gz = gz;
            std::vector<double> v8;
            if (stencil7_values(h, gx, gy, gz, code.data(), v8)) {
                if (prec == F64) {
                    double *d = nullptr;
                    if ((rc = dmalloc(&d, v8.size(), acc)))
                        return rc;
                    XH_HIP(hipMemcpy(d, v8.data(), v8.size() * 8, hipMemcpyHostToDevice));
                    M.st7 = d;
                } else {
                    std::vector<float> f8(v8.size());
                    for (size_t q = 0; q < v8.size(); ++q)
                        f8[q] = (float)v8[q];
                    float *d = nullptr;
                    if ((rc = dmalloc(&d, f8.size(), acc)))
                        return rc;
                    XH_HIP(hipMemcpy(d, f8.data(), f8.size() * 4, hipMemcpyHostToDevice));
                    M.st7 = d;
                }
            }
        }
    }
    M.y_rows = h.nrows;
    M.bytes = acc;
    return 0;
}

// The core-block part of h for hybrid GS with blocks of B rows: row i keeps the entries
// whose column lies in [B*(i/B), B*(i/B) + B), in stored order (the only ones the sweep of
// csr_SGS on the core block reads, csr_gs.inl:75-130), with u32 columns and the values
// in the matrix precision; inv_diag is taken from A (not owned).  Each row's list is
// padded to a multiple of KB entries (at least KB) with column = the block's end and
// value +0: the sweep kernel reads a zero x there, so a pad subtracts +0 * +0 = +0 from
// the row sum, which leaves every value (-0 included) unchanged.  KB (3 or 4) is the width
// that needs fewer pads.
// The entries of each row i outside its core block [lo, lo + B), lo = (i / B) * B, in
// stored order: b - (this matrix) x sums the products of the off-block residual of hybrid
// GS (csr_gs.inl:75-130, the columns outside the calling core's rows) in their order.
HostCsr off_block_csr(const HostCsr &h, int64_t B) {
    HostCsr o;
    o.nrows = h.nrows;
    o.ncols = h.ncols;
    o.row.assign((size_t)h.nrows + 1, 0);
    for (int64_t i = 0; i < h.nrows; ++i) {
        const int64_t lo = (i / B) * B;
        for (uint32_t j = h.row[i]; j < h.row[i + 1]; ++j)
            if ((int64_t)h.col[j] < lo || (int64_t)h.col[j] >= lo + B) {
                o.col.push_back(h.col[j]);
                o.val.push_back(h.val[j]);
            }
        o.row[i + 1] = (uint32_t)o.col.size();
    }
    return o;
}

int upload_core(const HostCsr &h, int64_t B, int prec, const Mat &A, Mat &M) {
    M = Mat();
    const int64_t n = h.nrows;
    std::vector<int64_t> cnt(n, 0);
    int64_t nreal = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t lo = (i / B) * B, hi = std::min(n, lo + B);
        for (uint32_t j = h.row[i]; j < h.row[i + 1]; ++j)
            cnt[i] += (int64_t)h.col[j] >= lo && (int64_t)h.col[j] < hi;
        nreal += cnt[i];
    }
    auto padded = [](int64_t c, int kb) { return std::max<int64_t>(kb, (c + kb - 1) / kb * kb); };
    int64_t tot3 = 0, tot4 = 0;
    for (int64_t i = 0; i < n; ++i) {
        tot3 += padded(cnt[i], 3);
        tot4 += padded(cnt[i], 4);
    }
    const int KB = tot3 < tot4 ? 3 : 4;
    std::vector<uint32_t> row(n + 1, 0), col;
    std::vector<double> val;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t lo = (i / B) * B, hi = std::min(n, lo + B);
        for (uint32_t j = h.row[i]; j < h.row[i + 1]; ++j)
            if ((int64_t)h.col[j] >= lo && (int64_t)h.col[j] < hi) {
                col.push_back(h.col[j]);
                val.push_back(h.val[j]);
            }
        for (int64_t q = cnt[i]; q < padded(cnt[i], KB); ++q) {
            col.push_back((uint32_t)hi);
            val.push_back(0.0);
        }
        row[i + 1] = (uint32_t)col.size();
    }
    const int64_t nnz = (int64_t)col.size();
    M.nrows = M.ncols = n;
    M.nnz = nreal;
    M.core_kb = KB;
    M.prec = prec;
    M.idx = IDX_U32;
    size_t acc = 0;
    int rc;
    uint32_t *dc = nullptr;
    if ((rc = dmalloc(&M.row, n + 1, acc)) || (rc = dmalloc(&dc, std::max<int64_t>(nnz, 1), acc)))
        return rc;
    M.col = dc;
    XH_HIP(hipMemcpy(M.row, row.data(), (n + 1) * 4, hipMemcpyHostToDevice));
    if (nnz)
        XH_HIP(hipMemcpy(dc, col.data(), nnz * 4, hipMemcpyHostToDevice));
    if (prec == F64) {
        double *v = nullptr;
        if ((rc = dmalloc(&v, std::max<int64_t>(nnz, 1), acc)))
            return rc;
        M.val = v;
        if (nnz)
            XH_HIP(hipMemcpy(v, val.data(), nnz * 8, hipMemcpyHostToDevice));
    } else {
        std::vector<float> vf(nnz);
        for (int64_t j = 0; j < nnz; ++j)
            vf[j] = (float)val[j]; // as upload_mat converts A
        float *v = nullptr;
        if ((rc = dmalloc(&v, std::max<int64_t>(nnz, 1), acc)))
            return rc;
        M.val = v;
        if (nnz)
            XH_HIP(hipMemcpy(v, vf.data(), nnz * 4, hipMemcpyHostToDevice));
    }
    M.inv_diag = A.inv_diag;
    M.core_B = B;
    for (int64_t lo = 0; lo < n; lo += B)
        M.core_emax = std::max<int64_t>(M.core_emax, row[std::min(n, lo + B)] - row[lo]);
    for (int64_t i = 0; i < n; ++i)
        M.core_rowmax = std::max<int64_t>(M.core_rowmax, row[i + 1] - row[i]);
    M.bytes = acc;
    return 0;
}

void free_mat(Mat &M) {
    for (void *p : {(void *)M.row, M.col, (void *)M.panel_base, M.val, M.inv_diag,
                    M.inv_sqrt_diag, (void *)M.row_ind, (void *)M.col_ind, (void *)M.gcode,
                    M.st7})
        if (p)
            (void)hipFree(p);
    M = Mat();
}

// Gauss-Jordan without pivoting (src/primitives/matrix/operations/inverse.inl:31-106)
template <typename T>
void gj_inverse(int64_t n, const std::vector<T> &A, std::vector<T> &inv) {
    const int64_t nc = 2 * n;
    std::vector<T> e((size_t)n * nc, T(0));
    for (int64_t i = 0; i < n; ++i) {
        e[i * nc + i + n] = 1.0;
        for (int64_t j = 0; j < n; ++j)
            e[i * nc + j] = A[i * n + j];
    }
    for (int64_t i = 0; i < n; ++i) {
        T *ri = &e[i * nc];
        for (int64_t k = nc - 1; k > i; --k)
            ri[k] /= ri[i];
        ri[i] = 1.0;
        for (int64_t j = i + 1; j < n; ++j) {
            T *rj = &e[j * nc];
            for (int64_t k = nc - 1; k > i; --k)
                rj[k] -= rj[i] * ri[k];
            rj[i] = 0.0;
        }
    }
    for (int64_t ii = 0; ii < n; ++ii) {
        const int64_t row = n - 1 - ii;
        T *ri = &e[row * nc];
        for (int64_t j = 0; j < row; ++j) {
            T *rj = &e[j * nc];
            for (int64_t k = nc / 2; k < nc; ++k)
                rj[k] -= rj[row] * ri[k];
            rj[row] = 0.0;
        }
    }
    inv.assign((size_t)n * n, T(0));
    for (int64_t i = 0; i < n; ++i)
        for (int64_t j = 0; j < n; ++j)
            inv[i * n + j] = e[i * nc + n + j];
}

// Chebyshev_polynomial_coeffs (src/solvers/detail/chebyshev.inl:50-121); orders above 4
// (the reference asserts) use the same polynomial p(t) = (1 - T_k((th-t)/de)/T_k(th/de))/t
// expanded by the three-term recurrence
int cheby_coeffs(double mn, double mx, int order, double fraction, std::vector<double> &c) {
    const double upper = mx * 1.1;
    const double lower = (upper - mn) * fraction + mn;
    const double th = (upper + lower) / 2., de = (upper - lower) / 2.;
    c.assign(order > 0 ? order : 1, 0.0);
    double den;
    switch (order) {
    case 1:
        c[0] = 1. / th;
        return 0;
    case 2:
        den = de * de - 2 * th * th;
        c[0] = -4 * th / den;
        c[1] = 2 / den;
        return 0;
    case 3:
        den = 3 * (de * de) * th - 4 * (th * th * th);
        c[0] = (3 * de * de - 12 * th * th) / den;
        c[1] = 12 * th / den;
        c[2] = -4 / den;
        return 0;
    case 4:
        den = pow(de, 4) - 8 * de * de * th * th + 8 * pow(th, 4);
        c[0] = (32 * pow(th, 3) - 16 * de * de * th) / den;
        c[1] = (8 * de * de - 48 * th * th) / den;
        c[2] = 32 * th / den;
        c[3] = -8 / den;
        return 0;
    default:
        break;
    }
    if (order < 1 || order > 16)
        return -1;
    std::vector<double> tm1(order + 2, 0.0), tm(order + 2, 0.0), tn(order + 2, 0.0);
    tm1[0] = 1.0;
    tm[0] = th / de;
    tm[1] = -1.0 / de;
    for (int m = 1; m < order; ++m) {
        std::fill(tn.begin(), tn.end(), 0.0);
        for (int i = 0; i <= m; ++i) {
            tn[i] += 2 * (th / de) * tm[i];
            tn[i + 1] += 2 * (-1.0 / de) * tm[i];
        }
        for (int i = 0; i <= m - 1; ++i)
            tn[i] -= tm1[i];
        tm1 = tm;
        tm = tn;
    }
    for (int i = 0; i < order; ++i)
        c[i] = -tm[i + 1] / tm[0];
    return 0;
}

} // namespace
} // namespace xh

using namespace xh;

struct xh_matrix_s {
    Mat m;
};

struct xh_hier_s {
    int nlevels = 0;
    std::vector<HostCsr> A, P, R;
    std::vector<double> mn, mx;
};

namespace xh {
namespace {

struct Level {
    int64_t n = 0;
    int prec = F64, tprec = F64; // operator / transfer precision
    Mat A, P, R;
    bool own_A = true;
    double mn = 0, mx = 0;
    // the smoother roles of this level: params.get("pre_smoother", l) and
    // params.get("post_smoother", l) (mg.inl:113-131)
    xh_role pre{}, post{};
    int64_t hsgs_B = 0;
    // the entries of A inside each row's core block of hsgs_B rows (stored order, u32
    // columns; inv_diag is A's): what the hybrid GS sweeps read, built when a role of
    // the level is HSGS
    Mat core;
    // the other entries of A (outside each row's core block, stored order): b - off x is
    // the off-block residual of the sweeps; built with `core` when hsgs_offblock_csr is on
    Mat off;
    void *tmp = nullptr, *tmp_t = nullptr, *b = nullptr, *x = nullptr;
    void *sr = nullptr, *su1 = nullptr, *su2 = nullptr;
    void *dinv = nullptr;
    int dinv_prec = F64;
    void *cb64 = nullptr, *cx64 = nullptr; // fp64 coarse solve in an fp32 hierarchy
    std::vector<double> cheb_pre, cheb_post, cheb_coarse;
};

} // namespace
} // namespace xh

struct xh_solver_s {
    xh_params prm;
    int nv = 0;
    int64_t n = 0;
    std::vector<Level> lv;
    Mat A0; // fp64 finest operator of the Krylov iteration
    hipStream_t stream = nullptr;
    // Krylov vectors (fp64)
    double *r0 = nullptr, *r = nullptr, *p = nullptr, *v = nullptr, *s = nullptr, *t = nullptr;
    double *p_ = nullptr, *s_ = nullptr, *xi = nullptr, *bi = nullptr;
    void *b32 = nullptr, *x32 = nullptr; // level-0 fp32 b/x (mg_prec 2)
    SolverState *st = nullptr;
    SolverState *hst = nullptr; // pinned host mirror
    int *hdone = nullptr;       // pinned [2]
    double *partials = nullptr;
    double *hist = nullptr;
    int hist_len = 0;
    const int *dflag = nullptr; // `done` gate of the launches being enqueued
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipGraphExec_t gexec = nullptr, gexec_rep = nullptr; // plain / residual-replacing
    hipGraph_t graph = nullptr, graph_rep = nullptr;
    const double *g_x = nullptr, *g_b = nullptr; // pointers the graphs were captured with
    const double *bcur = nullptr;                // b of the solve being enqueued
    size_t bytes = 0;
    int err = 0;
};

namespace xh {
namespace {

template <typename T>
int alloc_vec(xh_solver_s *S, T **p, size_t count) {
    return dmalloc(p, count, S->bytes);
}

int alloc_prec(xh_solver_s *S, void **p, size_t count, int prec) {
    if (prec == F64)
        return alloc_vec(S, (double **)p, count);
    return alloc_vec(S, (float **)p, count);
}

SpmvArgs args(const xh_solver_s *S) {
    SpmvArgs a;
    a.done = S->dflag;
    return a;
}

void smooth(xh_solver_s *S, const xh_role &role, Level &L, const std::vector<double> &cheb,
            void *x, const void *b, bool x_zero) {
    const int nv = S->nv;
    hipStream_t st = S->stream;
    switch (role.method) {
    case XH_CHEBYSHEV: {
        const int k = role.poly_order;
        void *uc = L.su1, *un = L.su2;
        int first = k - 2; // the first coefficient index left for the OP_CHEB loop
        bool r_from_b = false; // the later steps form r = b*isd themselves (OP_CHEBR)
        if (x_zero && k >= 2 && g_tune.chebz_enable && L.A.idx != IDX_U8 &&
            spmv_row_parallel(L.A, OP_CHEB, nv)) {
            // the zero-x first step and the next one in one launch: the SpMV gathers b and
            // forms u = (c*(b*isd))*isd per entry, bit-identical to the two launches; r is
            // not stored, the later steps form it from b
            SpmvArgs a = args(S);
            a.b = b;
            a.isd = L.A.inv_sqrt_diag;
            a.out2 = nullptr;
            a.c = cheb[k - 2];
            a.c2 = cheb[k - 1];
            a.last = k == 2;
            a.xupd = x;
            a.x_zero = 1;
            spmv(L.A, nv, b, uc, OP_CHEBZ, a, st);
            first = k - 3;
            r_from_b = true;
        } else if (x_zero) {
            cheb_init_zero(L.n, nv, L.prec, b, L.A.inv_sqrt_diag, cheb[k - 1], L.sr,
                           k == 1 ? x : uc, k == 1, S->dflag, st);
        } else {
            // r = isd*(b - A x), u = (c*r)*isd; for order 1 then x = u + x in a launch of
            // its own: x is this SpMV's operand, which other rows still read, so it cannot
            // be updated in place by the same launch
            SpmvArgs a = args(S);
            a.b = b;
            a.isd = L.A.inv_sqrt_diag;
            a.out2 = L.sr;
            a.c = cheb[k - 1];
            a.last = 0;
            a.x_zero = 0;
            spmv(L.A, nv, x, uc, OP_CHEB0, a, st);
            if (k == 1)
                add_vec(L.n * nv, L.prec, uc, x, S->dflag, st);
        }
        for (int i = first; i >= 0; --i) {
            SpmvArgs a = args(S);
            a.isd = L.A.inv_sqrt_diag;
            a.r = L.sr;
            a.b = b;
            a.c = cheb[i];
            a.last = i == 0;
            a.xupd = x;
            a.x_zero = x_zero ? 1 : 0;
            spmv(L.A, nv, uc, un, r_from_b ? OP_CHEBR : OP_CHEB, a, st);
            std::swap(uc, un);
        }
        return;
    }
    case XH_JACOBI: {
        const double relax = L.prec == F64 ? (double)role.relax_factor
                                           : (double)(float)role.relax_factor;
        for (int it = 0; it < std::max(1, role.max_iters); ++it) {
            if (x_zero) {
                jacobi_update(L.n, nv, L.prec, x, b, nullptr, L.A.inv_diag, relax, 1, S->dflag,
                              st);
            } else {
                spmv(L.A, nv, x, L.sr, OP_SET, args(S), st);
                jacobi_update(L.n, nv, L.prec, x, b, L.sr, L.A.inv_diag, relax, 0, S->dflag,
                              st);
            }
            x_zero = false;
        }
        return;
    }
    case XH_HSGS: {
        const double relax = (double)role.relax_factor;
        for (int it = 0; it < std::max(1, role.max_iters); ++it) {
            const Mat &G = L.core.row ? L.core : L.A;
            if (x_zero) {
                hsgs_sweeps(G, nv, b, x, 1, relax, L.hsgs_B, S->dflag, st);
            } else {
                SpmvArgs a = args(S);
                a.b = b;
                if (L.off.row && g_tune.hsgs_off_csr) {
                    spmv(L.off, nv, x, L.sr, OP_RESID, a, st);
                } else {
                    a.blk_rows = L.hsgs_B;
                    spmv(L.A, nv, x, L.sr, OP_OFFBLOCK, a, st);
                }
                hsgs_sweeps(G, nv, L.sr, x, 0, relax, L.hsgs_B, S->dflag, st);
            }
            x_zero = false;
        }
        return;
    }
    case XH_DIRECT: {
        if (L.cb64) {
            convert(L.n * nv, F32, b, L.cb64, S->dflag, st);
            dense_matvec(L.n, nv, F64, L.dinv, L.cb64, L.cx64, S->dflag, st);
            convert(L.n * nv, F64, L.cx64, x, S->dflag, st);
        } else {
            dense_matvec(L.n, nv, L.dinv_prec, L.dinv, b, x, S->dflag, st);
        }
        return;
    }
    case XH_IDENTITY:
    default:
        copy_vec(L.n * nv, L.prec, b, x, S->dflag, st);
        return;
    }
}

// MultiGrid::MG_cycle (mg.inl:166-260); x_zero mirrors the reference's if_zero flag
void mg_cycle(xh_solver_s *S, int l, void *x, const void *b, bool x_zero, int cyc) {
    Level &L = S->lv[l];
    Level &H = S->lv[l + 1];
    const int nv = S->nv;
    hipStream_t st = S->stream;
    const int nl = (int)S->lv.size();
    ledger_level(l);
    if (cyc != 2) {
        smooth(S, L.pre, L, L.cheb_pre, x, b, x_zero);
        x_zero = false;
    }
    void *rt = (L.tprec == L.prec) ? L.tmp : L.tmp_t;
    if (x_zero) {
        if (L.tprec == L.prec)
            copy_vec(L.n * nv, L.prec, b, rt, S->dflag, st);
        else
            convert(L.n * nv, L.prec, b, rt, S->dflag, st);
    } else {
        SpmvArgs a = args(S);
        a.b = b;
        a.out_f32 = (L.tprec != L.prec) ? 1 : 0;
        spmv(L.A, nv, x, rt, OP_RESID, a, st);
    }
    spmv(L.R, nv, rt, H.b, OP_SET, args(S), st);
    if (l + 1 == nl - 1) {
        ledger_level(l + 1);
        smooth(S, S->prm.coarse, H, H.cheb_coarse, H.x, H.b, true);
    } else {
        mg_cycle(S, l + 1, H.x, H.b, true, cyc);
        if (cyc == 1)
            mg_cycle(S, l + 1, H.x, H.b, false, 1);
        else if (cyc == 2)
            mg_cycle(S, l + 1, H.x, H.b, false, 0);
    }
    ledger_level(l);
    if (L.tprec == L.prec) {
        spmv(L.P, nv, H.x, x, x_zero ? OP_SET : OP_ADD, args(S), st);
    } else {
        SpmvArgs a = args(S);
        a.x_zero = x_zero ? 1 : 0;
        spmv(L.P, nv, H.x, x, OP_PROLONG32, a, st);
    }
    smooth(S, L.post, L, L.cheb_post, x, b, false);
}

// preconditioner application x = M^{-1} b, x zero-flagged on entry
void precond_body(xh_solver_s *S, const double *b, double *x);

void precond(xh_solver_s *S, const double *b, double *x) {
    ledger_level(0);
    precond_body(S, b, x);
    ledger_level(-1);
}

void precond_body(xh_solver_s *S, const double *b, double *x) {
    const xh_role &pr = S->prm.precond;
    const int nv = S->nv;
    hipStream_t st = S->stream;
    Level &L0 = S->lv[0];
    if (pr.method == XH_MULTIGRID) {
        const int nl = (int)S->lv.size();
        const int iters = std::max(1, pr.max_iters);
        if (S->prm.mg_prec == 2) {
            convert(S->n * nv, F64, b, S->b32, S->dflag, st);
            for (int it = 0; it < iters; ++it) {
                if (nl == 1)
                    smooth(S, S->prm.coarse, L0, L0.cheb_coarse, S->x32, S->b32, it == 0);
                else
                    mg_cycle(S, 0, S->x32, S->b32, it == 0, S->prm.mg_cycle);
            }
            convert(S->n * nv, F32, S->x32, x, S->dflag, st);
        } else {
            for (int it = 0; it < iters; ++it) {
                if (nl == 1)
                    smooth(S, S->prm.coarse, L0, L0.cheb_coarse, x, b, it == 0);
                else
                    mg_cycle(S, 0, x, b, it == 0, S->prm.mg_cycle);
            }
        }
        return;
    }
    smooth(S, pr, L0, L0.cheb_pre, x, b, true);
}

bool is_preconditioned(int m) {
    return m == XH_PBICGSTAB || m == XH_MERGED_PBICGSTAB || m == XH_PCG || m == XH_RBICGSTAB ||
           m == XH_MERGED_RBICGSTAB;
}

bool is_rbicgstab(int m) { return m == XH_RBICGSTAB || m == XH_MERGED_RBICGSTAB; }

// Vector roles of PCG (pcg.inl:77-80): r, z -> p_, p, Ap -> v.
// RBiCGStab (rbicgstab.inl:85-92): r0, r, z -> p_, v_ -> p, v, s, t_ -> s_, t.

// get_residual + z = M^{-1} r + p = z + rho = (r, z)   (pcg.inl:84-96, 114-115)
void enqueue_init_pcg(xh_solver_s *S, const double *b, double *x) {
    SpmvArgs a = args(S);
    a.b = b;
    a.out2 = S->t; // scratch copies, overwritten before use
    a.out3 = S->s;
    a.partials = S->partials;
    int np = spmv(S->A0, S->nv, x, S->r, OP_RESID0, a, S->stream);
    finalize(FIN_INIT, S->st, S->partials, np, S->nv, S->stream);
    precond(S, S->r, S->p_);
    VecArgs vc;
    vc.st = S->st;
    vc.p = S->p;
    vc.v = S->p_;
    vc.r = S->r;
    vc.partials = S->partials;
    np = vec_op(V_COPY_DOT, S->n, S->nv, vc, S->stream);
    finalize(FIN_CG_RHO, S->st, S->partials, np, S->nv, S->stream);
}

// r0 = b - A x ; r = r0 ; rho = rho0 ; z = M^{-1} r0 ; v_ = z   (rbicgstab.inl:96-113)
void enqueue_init_rbicgstab(xh_solver_s *S, const double *b, double *x) {
    SpmvArgs a = args(S);
    a.b = b;
    a.out2 = S->r;
    a.out3 = S->t; // scratch copy, overwritten before use
    a.partials = S->partials;
    int np = spmv(S->A0, S->nv, x, S->r0, OP_RESID0, a, S->stream);
    finalize(FIN_INIT, S->st, S->partials, np, S->nv, S->stream);
    precond(S, S->r0, S->p_);
    VecArgs vc;
    vc.st = S->st;
    vc.p = S->p;
    vc.v = S->p_;
    vec_op(V_COPY, S->n, S->nv, vc, S->stream);
}

void enqueue_init(xh_solver_s *S, const double *b, double *x) {
    const int m = S->prm.solver.method;
    if (m == XH_PCG)
        return enqueue_init_pcg(S, b, x);
    if (is_rbicgstab(m))
        return enqueue_init_rbicgstab(S, b, x);
    SpmvArgs a = args(S);
    a.b = b;
    a.out2 = S->p;
    a.out3 = S->r;
    a.partials = S->partials;
    int np = spmv(S->A0, S->nv, x, S->r0, OP_RESID0, a, S->stream);
    finalize(FIN_INIT, S->st, S->partials, np, S->nv, S->stream);
}

// r = b - A x (fp64) and partials of (r, r0), replacing the recurrence residual and the
// rho_new partials of the r update just enqueued; returns the partial count
int enqueue_replacement(xh_solver_s *S, double *x) {
    SpmvArgs a = args(S);
    a.b = S->bcur;
    spmv(S->A0, S->nv, x, S->r, OP_RESID, a, S->stream);
    VecArgs vd;
    vd.st = S->st;
    vd.r = S->r;
    vd.r0 = S->r0;
    vd.partials = S->partials;
    return vec_op(V_DOT1, S->n, S->nv, vd, S->stream);
}

// one PCG iteration (pcg.inl:101-160)
void enqueue_iteration_pcg(xh_solver_s *S, double *x) {
    const int nv = S->nv;
    hipStream_t st = S->stream;
    SpmvArgs a = args(S);
    a.o1 = S->p;
    a.partials = S->partials;
    int np = spmv(S->A0, nv, S->p, S->v, OP_DOT1, a, st); // Ap ; (p, Ap)
    finalize(FIN_CG_ALPHA, S->st, S->partials, np, nv, st);
    VecArgs vx;
    vx.st = S->st;
    vx.x = x;
    vx.p_ = S->p;
    vx.t = S->v;
    vx.r = S->r;
    vec_op(V_CG_XR, S->n, nv, vx, st);
    precond(S, S->r, S->p_);
    VecArgs vd;
    vd.st = S->st;
    vd.r = S->r;
    vd.r0 = S->p_;
    vd.partials = S->partials;
    np = vec_op(V_DOT2, S->n, nv, vd, st); // (r, z), (r, r)
    finalize(FIN_CG_BETA, S->st, S->partials, np, nv, st);
    VecArgs vp;
    vp.st = S->st;
    vp.p = S->p;
    vp.v = S->p_;
    vec_op(V_CG_P, S->n, nv, vp, st);
}

// one RBiCGStab / MergedRBiCGStab iteration (rbicgstab.inl:124-216,
// merged_rbicgstab.inl:124-216)
void enqueue_iteration_rbicgstab(xh_solver_s *S, double *x, bool merged) {
    const int nv = S->nv;
    hipStream_t st = S->stream;
    SpmvArgs a = args(S);
    a.o1 = S->r0;
    a.partials = S->partials;
    int np = spmv(S->A0, nv, S->p, S->v, OP_DOT1, a, st); // 6-7: v = A v_ ; delta
    finalize(FIN_ALPHA, S->st, S->partials, np, nv, st);  // 9: alpha = rho / delta
    precond(S, S->v, S->s);                               // 8: s = M^{-1} v
    VecArgs vt;
    vt.st = S->st;
    vt.s = S->s_;
    vt.r = S->p_;
    vt.v = S->s;
    vec_op(V_S_UPDATE, S->n, nv, vt, st);                 // 10: t_ = z - alpha s
    SpmvArgs a4 = args(S);
    a4.b = S->v;
    a4.o1 = S->r0;
    a4.r = S->st->alpha;
    a4.out2 = S->r;
    a4.partials = S->partials;
    np = spmv(S->A0, nv, S->s_, S->t, OP_RDOT4, a4, st);  // 11-13
    finalize(merged ? FIN_RB_OMEGA_MERGED : FIN_RB_OMEGA, S->st, S->partials, np, nv, st);
    precond(S, S->t, S->p_);                              // 14: z = M^{-1} t
    VecArgs vu;
    vu.st = S->st;
    vu.x = x;
    vu.p = S->p;
    vu.s = S->p_;
    vu.r = S->r;
    vu.t = S->t;
    vu.p_ = S->s_;
    vu.v = S->s;
    vec_op(merged ? V_RB_UPDATE_MERGED : V_RB_UPDATE, S->n, nv, vu, st); // 16, 23-25
    finalize(FIN_END, S->st, S->partials, 0, nv, st);
}

void enqueue_iteration(xh_solver_s *S, double *x, bool replace) {
    const int m = S->prm.solver.method;
    if (m == XH_PCG)
        return enqueue_iteration_pcg(S, x);
    if (is_rbicgstab(m))
        return enqueue_iteration_rbicgstab(S, x, m == XH_MERGED_RBICGSTAB);
    const int nv = S->nv;
    hipStream_t st = S->stream;
    const bool pre = is_preconditioned(m);
    double *pp = pre ? S->p_ : S->p;
    double *ss = pre ? S->s_ : S->s;
    if (pre)
        precond(S, S->p, S->p_);
    SpmvArgs a = args(S);
    a.o1 = S->r0;
    a.partials = S->partials;
    int np = spmv(S->A0, nv, pp, S->v, OP_DOT1, a, st);
    finalize(FIN_ALPHA, S->st, S->partials, np, nv, st);
    VecArgs va;
    va.st = S->st;
    va.s = S->s;
    va.r = S->r;
    va.v = S->v;
    vec_op(V_S_UPDATE, S->n, nv, va, st);
    if (pre)
        precond(S, S->s, S->s_);
    SpmvArgs a3 = args(S);
    a3.o1 = S->s;
    a3.partials = S->partials;
    np = spmv(S->A0, nv, ss, S->t, OP_DOT3, a3, st);
    if (m == XH_MERGED_BICGSTAB) {
        finalize(FIN_OMEGA_MERGED, S->st, S->partials, np, nv, st);
        VecArgs vx;
        vx.st = S->st;
        vx.x = x;
        vx.p_ = S->p;
        vx.s_ = S->s;
        vx.gate = 1;
        vec_op(V_X_UPDATE, S->n, nv, vx, st);
        VecArgs vr;
        vr.st = S->st;
        vr.s = S->s;
        vr.t = S->t;
        vr.r = S->r;
        vr.r0 = S->r0;
        vr.partials = S->partials;
        np = vec_op(V_R_UPDATE, S->n, nv, vr, st);
        finalize(FIN_BETA, S->st, S->partials, np, nv, st);
        VecArgs v5;
        v5.st = S->st;
        v5.x = x;
        v5.p = S->p;
        v5.s = S->s;
        v5.r = S->r;
        v5.v = S->v;
        vec_op(V_GROUP5, S->n, nv, v5, st);
        return;
    }
    finalize(FIN_OMEGA, S->st, S->partials, np, nv, st);
    VecArgs vx;
    vx.st = S->st;
    vx.x = x;
    vx.p_ = pp;
    vx.s_ = ss;
    vx.s = S->s;
    vx.t = S->t;
    vx.r = S->r;
    vx.r0 = S->r0;
    vx.partials = S->partials;
    np = vec_op(V_XR_UPDATE, S->n, nv, vx, st);
    if (replace)
        np = enqueue_replacement(S, x);
    finalize(FIN_BETA, S->st, S->partials, np, nv, st);
    VecArgs vp;
    vp.st = S->st;
    vp.p = S->p;
    vp.r = S->r;
    vp.v = S->v;
    vec_op(V_P_UPDATE, S->n, nv, vp, st);
}

void free_graphs(xh_solver_s *S) {
    for (hipGraphExec_t *e : {&S->gexec, &S->gexec_rep})
        if (*e) {
            (void)hipGraphExecDestroy(*e);
            *e = nullptr;
        }
    for (hipGraph_t *g : {&S->graph, &S->graph_rep})
        if (*g) {
            (void)hipGraphDestroy(*g);
            *g = nullptr;
        }
    S->g_x = nullptr;
    S->g_b = nullptr;
}

int check_launch() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return fail(XH_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return 0;
}

int solve_impl(xh_solver_s *S, const double *db, double *dx) {
    const int nv = S->nv;
    // initial state (solver_stats::reset + iconv = ones, solver.h:54-65)
    SolverState &h = *S->hst;
    std::memset(&h, 0, sizeof(SolverState));
    for (int k = 0; k < nv; ++k)
        h.iconv[k] = 1.0;
    h.nv = nv;
    h.max_iters = S->prm.solver.max_iters;
    h.conv_check = S->prm.solver.conv_check;
    h.abs_tol = S->prm.solver.abs_tol;
    h.rel_tol = S->prm.solver.rel_tol;
    h.hist = S->hist;
    h.hist_len = S->hist_len;
    XH_HIP(hipMemcpyAsync(S->st, S->hst, sizeof(SolverState), hipMemcpyHostToDevice, S->stream));
    S->dflag = &S->st->done;
    if (g_ledger) {
        g_ledger->iter = 0;
        g_ledger->level = -1;
    }
    enqueue_init(S, db, dx);
    int rc = check_launch();
    if (rc)
        return rc;
    const bool use_graph = S->prm.graph != 0;
    const int period = S->prm.true_res_period;
    S->bcur = db;
    if (use_graph && (S->gexec == nullptr || S->g_x != dx || S->g_b != db)) {
        free_graphs(S);
        XH_HIP(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
        enqueue_iteration(S, dx, false);
        XH_HIP(hipStreamEndCapture(S->stream, &S->graph));
        XH_HIP(hipGraphInstantiate(&S->gexec, S->graph, nullptr, nullptr, 0));
        if (period > 0) {
            XH_HIP(hipStreamBeginCapture(S->stream, hipStreamCaptureModeThreadLocal));
            enqueue_iteration(S, dx, true);
            XH_HIP(hipStreamEndCapture(S->stream, &S->graph_rep));
            XH_HIP(hipGraphInstantiate(&S->gexec_rep, S->graph_rep, nullptr, nullptr, 0));
        }
        S->g_x = dx;
        S->g_b = db;
    }
    const int maxit = std::max(1, S->prm.solver.max_iters) + 1;
    for (int it = 0; it < maxit; ++it) {
        // replay `it` runs iteration it + 1 (or nothing once the solve is done)
        const bool replace = period > 0 && (it + 1) % period == 0;
        if (use_graph) {
            XH_HIP(hipGraphLaunch(replace ? S->gexec_rep : S->gexec, S->stream));
        } else {
            if (g_ledger)
                g_ledger->iter = it + 1;
            enqueue_iteration(S, dx, replace);
            if ((rc = check_launch()))
                return rc;
        }
        XH_HIP(hipMemcpyAsync(&S->hdone[it & 1], &S->st->done, sizeof(int),
                              hipMemcpyDeviceToHost, S->stream));
        XH_HIP(hipEventRecord(S->ev[it & 1], S->stream));
        if (it > 0) {
            XH_HIP(hipEventSynchronize(S->ev[(it - 1) & 1]));
            if (S->hdone[(it - 1) & 1])
                break;
        }
    }
    XH_HIP(hipMemcpyAsync(S->hst, S->st, sizeof(SolverState), hipMemcpyDeviceToHost, S->stream));
    XH_HIP(hipStreamSynchronize(S->stream));
    S->dflag = nullptr;
    if (!S->hst->done)
        return fail(XH_EINVAL, "solve did not terminate within max_iters");
    return 0;
}

} // namespace
} // namespace xh

extern "C" {

int xh_init(int device) {
    int n = 0;
    XH_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fail(XH_EINVAL, "no such HIP device");
    XH_HIP(hipSetDevice(device));
    if (g_stream == nullptr)
        XH_HIP(hipStreamCreate(&g_stream));
    if ((int)g_scratch.size() < n)
        g_scratch.resize(n, nullptr);
    if (g_scratch[device] == nullptr) {
        void *p = nullptr;
        XH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[device] = p;
    }
    g_device = device;
    return 0;
}

int xh_finalize(void) {
    if (g_stream) {
        (void)hipStreamDestroy(g_stream);
        g_stream = nullptr;
    }
    g_device = -1;
    return 0;
}

const char *xh_last_error(void) { return g_err.c_str(); }

int xh_device_info(char *name, int name_len, int *compute_units, double *hbm_bytes) {
    int rc = ensure_init();
    if (rc)
        return rc;
    hipDeviceProp_t p;
    XH_HIP(hipGetDeviceProperties(&p, g_device));
    if (name && name_len > 0) {
        std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (compute_units)
        *compute_units = p.multiProcessorCount;
    if (hbm_bytes)
        *hbm_bytes = (double)p.totalGlobalMem;
    return 0;
}

int xh_malloc(void **dptr, size_t bytes) {
    int rc = ensure_init();
    if (rc)
        return rc;
    XH_HIP(hipMalloc(dptr, bytes));
    return 0;
}

int xh_free(void *dptr) {
    XH_HIP(hipFree(dptr));
    return 0;
}

int xh_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    XH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return 0;
}

int xh_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    XH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return 0;
}

int xh_synchronize(void) {
    XH_HIP(hipDeviceSynchronize());
    return 0;
}

int xh_fill_zero(void *dptr, size_t bytes) {
    int rc = ensure_init();
    if (rc)
        return rc;
    if (!dptr)
        return fail(XH_EINVAL, "xh_fill_zero: null pointer");
    XH_HIP(hipMemsetAsync(dptr, 0, bytes, g_stream));
    XH_HIP(hipStreamSynchronize(g_stream));
    return 0;
}

int xh_matrix_create(int64_t nrows, int64_t ncols, int64_t nnz, const uint32_t *row,
                     const uint32_t *col, const double *val, int prec, int index_mode,
                     xh_matrix *out) {
    int rc = ensure_init();
    if (rc)
        return rc;
    if (!out || !row || (nnz > 0 && (!col || !val)) || nrows < 0 || ncols < 0)
        return fail(XH_EINVAL, "xh_matrix_create: bad arguments");
    if (prec != F64 && prec != F32)
        return fail(XH_EINVAL, "xh_matrix_create: prec must be 0 (f64) or 1 (f32)");
    if ((int64_t)row[nrows] != nnz || row[0] != 0)
        return fail(XH_EINVAL, "xh_matrix_create: row pointer does not match nnz");
    HostCsr h;
    h.nrows = nrows;
    h.ncols = ncols;
    h.row.assign(row, row + nrows + 1);
    h.col.assign(col, col + nnz);
    h.val.assign(val, val + nnz);
    for (int64_t i = 0; i < nrows; ++i)
        if (h.row[i + 1] < h.row[i])
            return fail(XH_EINVAL, "xh_matrix_create: row pointer not monotone");
    for (int64_t j = 0; j < nnz; ++j)
        if ((int64_t)h.col[j] >= ncols)
            return fail(XH_EINVAL, "xh_matrix_create: column index out of range");
    auto *m = new xh_matrix_s();
    rc = upload_mat(h, prec, index_mode, m->m);
    if (rc) {
        free_mat(m->m);
        delete m;
        return rc;
    }
    *out = m;
    return 0;
}

int xh_matrix_info(xh_matrix A, int64_t *nrows, int64_t *ncols, int64_t *nnz, int *prec,
                   int *index_bits, int64_t *panel_rows, size_t *device_bytes) {
    if (!A)
        return fail(XH_EINVAL, "null matrix");
    if (nrows)
        *nrows = A->m.nrows;
    if (ncols)
        *ncols = A->m.ncols;
    if (nnz)
        *nnz = A->m.nnz;
    if (prec)
        *prec = A->m.prec;
    if (index_bits)
        *index_bits = A->m.idx == IDX_U8 ? 8 : (A->m.idx == IDX_U16 ? 16 : 32);
    if (panel_rows)
        *panel_rows = A->m.idx == IDX_U16 ? ((int64_t)1 << A->m.panel_shift) : 0;
    if (device_bytes)
        *device_bytes = A->m.bytes;
    return 0;
}

int xh_matrix_destroy(xh_matrix A) {
    if (A) {
        free_mat(A->m);
        delete A;
    }
    return 0;
}

int xh_spmv(xh_matrix A, int nv, const void *dx, void *dy, int accumulate) {
    if (!A || !nv_ok(nv) || !dx || !dy)
        return fail(XH_EINVAL, "xh_spmv: bad arguments");
    SpmvArgs a;
    if (spmv(A->m, nv, dx, dy, accumulate ? OP_ADD : OP_SET, a, g_stream) < 0)
        return fail(XH_ENOTSUP, "xh_spmv: unsupported configuration");
    XH_HIP(hipGetLastError());
    XH_HIP(hipStreamSynchronize(g_stream));
    return 0;
}

int xh_spmv_timed(xh_matrix A, int nv, const void *dx, void *dy, int accumulate, int reps,
                  double *ms_per_launch) {
    if (!A || !nv_ok(nv) || !dx || !dy || reps < 1)
        return fail(XH_EINVAL, "xh_spmv_timed: bad arguments");
    hipEvent_t e0, e1;
    XH_HIP(hipEventCreate(&e0));
    XH_HIP(hipEventCreate(&e1));
    SpmvArgs a;
    // one untimed launch first: the first launch of a kernel also loads its code object
    spmv(A->m, nv, dx, dy, accumulate ? OP_ADD : OP_SET, a, g_stream);
    XH_HIP(hipEventRecord(e0, g_stream));
    for (int i = 0; i < reps; ++i)
        spmv(A->m, nv, dx, dy, accumulate ? OP_ADD : OP_SET, a, g_stream);
    XH_HIP(hipEventRecord(e1, g_stream));
    XH_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    XH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms_per_launch)
        *ms_per_launch = (double)ms / reps;
    XH_HIP(hipGetLastError());
    return 0;
}

int xh_dense_matvec(int64_t n, int nv, int prec, const void *dD, const void *dx, void *dy) {
    if (n <= 0 || !nv_ok(nv) || !dD || !dx || !dy)
        return fail(XH_EINVAL, "xh_dense_matvec: bad arguments");
    dense_matvec(n, nv, prec, dD, dx, dy, nullptr, g_stream);
    XH_HIP(hipGetLastError());
    XH_HIP(hipStreamSynchronize(g_stream));
    return 0;
}

int xh_hier_create(int nlevels, xh_hier *out) {
    if (nlevels < 1 || !out)
        return fail(XH_EINVAL, "xh_hier_create: nlevels must be >= 1");
    auto *H = new xh_hier_s();
    H->nlevels = nlevels;
    H->A.resize(nlevels);
    H->P.resize(nlevels);
    H->R.resize(nlevels);
    H->mn.assign(nlevels, 0.0);
    H->mx.assign(nlevels, 0.0);
    *out = H;
    return 0;
}

static int copy_csr(HostCsr &d, int64_t nr, int64_t nc, int64_t nnz, const uint32_t *row,
                    const uint32_t *col, const double *val) {
    if (!row || (nnz > 0 && (!col || !val)))
        return fail(XH_EINVAL, "xh_hier_set_level: missing CSR array");
    if ((int64_t)row[nr] != nnz)
        return fail(XH_EINVAL, "xh_hier_set_level: row pointer does not match nnz");
    d.nrows = nr;
    d.ncols = nc;
    d.row.assign(row, row + nr + 1);
    d.col.assign(col, col + nnz);
    d.val.assign(val, val + nnz);
    for (int64_t j = 0; j < nnz; ++j)
        if ((int64_t)d.col[j] >= nc)
            return fail(XH_EINVAL, "xh_hier_set_level: column index out of range");
    return 0;
}

int xh_hier_set_level(xh_hier H, int level, int64_t n, int64_t nnzA, const uint32_t *rowA,
                      const uint32_t *colA, const double *valA, int64_t nc, int64_t nnzP,
                      const uint32_t *rowP, const uint32_t *colP, const double *valP,
                      int64_t nnzR, const uint32_t *rowR, const uint32_t *colR,
                      const double *valR, double min_eig, double max_eig) {
    if (!H || level < 0 || level >= H->nlevels)
        return fail(XH_EINVAL, "xh_hier_set_level: bad level");
    int rc = copy_csr(H->A[level], n, n, nnzA, rowA, colA, valA);
    if (rc)
        return rc;
    if (level + 1 < H->nlevels) {
        if ((rc = copy_csr(H->P[level], n, nc, nnzP, rowP, colP, valP)))
            return rc;
        if ((rc = copy_csr(H->R[level], nc, n, nnzR, rowR, colR, valR)))
            return rc;
    }
    H->mn[level] = min_eig;
    H->mx[level] = max_eig;
    return 0;
}

int xh_hier_destroy(xh_hier H) {
    delete H;
    return 0;
}

int xh_solver_destroy(xh_solver S);

int xh_solver_create(xh_hier H, const xh_params *prm, int nv, xh_solver *out) {
    int rc = ensure_init();
    if (rc)
        return rc;
    if (!H || !prm || !out || !nv_ok(nv))
        return fail(XH_EINVAL, "xh_solver_create: bad arguments (nv must be 1,2,4,...,64)");
    const int m = prm->solver.method;
    if (m != XH_BICGSTAB && m != XH_PBICGSTAB && m != XH_MERGED_BICGSTAB &&
        m != XH_MERGED_PBICGSTAB && m != XH_PCG && m != XH_RBICGSTAB && m != XH_MERGED_RBICGSTAB)
        return fail(XH_ENOTSUP, "xh_solver_create: solver method must be a Krylov method "
                                "(PCG or a BiCGStab variant)");
    if (is_preconditioned(m) && prm->precond.method == XH_NONE)
        return fail(XH_EINVAL, "xh_solver_create: preconditioned method without preconditioner");
    if (prm->true_res_period < 0)
        return fail(XH_EINVAL, "xh_solver_create: true_res_period must be >= 0");
    if (prm->true_res_period > 0 &&
        !(m == XH_BICGSTAB || m == XH_PBICGSTAB || m == XH_MERGED_PBICGSTAB))
        return fail(XH_ENOTSUP, "xh_solver_create: true_res_period is available for BiCGStab, "
                                "PBiCGStab and MergedPBiCGStab only");
    const int nl = H->nlevels;
    for (int l = 0; l < nl; ++l)
        if (H->A[l].nrows == 0 || (l + 1 < nl && H->P[l].ncols != H->A[l + 1].nrows))
            return fail(XH_EINVAL, "xh_solver_create: inconsistent hierarchy");
    auto check_role = [](const xh_role &r, bool smoother) {
        if (r.method == XH_CHEBYSHEV && (r.poly_order < 1 || r.poly_order > 16))
            return false;
        if (smoother && !(r.method == XH_JACOBI || r.method == XH_HSGS ||
                          r.method == XH_CHEBYSHEV || r.method == XH_DIRECT ||
                          r.method == XH_IDENTITY))
            return false;
        return true;
    };
    const bool mg = is_preconditioned(m) && prm->precond.method == XH_MULTIGRID;
    if (prm->nlev_roles < 0 || prm->nlev_roles > XH_MAX_LEVELS)
        return fail(XH_EINVAL, "xh_solver_create: nlev_roles must be in [0, XH_MAX_LEVELS]");
    // the smoother roles of level l
    auto pre_of = [&](int l) -> const xh_role & {
        return l < prm->nlev_roles ? prm->pre_lev[l] : prm->pre;
    };
    auto post_of = [&](int l) -> const xh_role & {
        return l < prm->nlev_roles ? prm->post_lev[l] : prm->post;
    };
    if (mg && !check_role(prm->coarse, true))
        return fail(XH_ENOTSUP, "xh_solver_create: unsupported coarse solver");
    if (mg)
        for (int l = 0; l + 1 < nl; ++l)
            if (!check_role(pre_of(l), true) || !check_role(post_of(l), true))
                return fail(XH_ENOTSUP, "xh_solver_create: unsupported smoother on level " +
                                            std::to_string(l));
    if (is_preconditioned(m) && !mg && !check_role(prm->precond, true))
        return fail(XH_ENOTSUP, "xh_solver_create: unsupported preconditioner");
    if (prm->mg_prec < 0 || prm->mg_prec > 2)
        return fail(XH_EINVAL, "xh_solver_create: mg_prec must be 0, 1 or 2");

    auto *S = new xh_solver_s();
    S->prm = *prm;
    S->nv = nv;
    S->n = H->A[0].nrows;
    XH_HIP(hipStreamCreate(&S->stream));
    XH_HIP(hipEventCreateWithFlags(&S->ev[0], hipEventDisableTiming));
    XH_HIP(hipEventCreateWithFlags(&S->ev[1], hipEventDisableTiming));
    auto bail = [&](int code) {
        xh_solver_destroy(S);
        return code;
    };
    // finest fp64 operator of the Krylov loop
    if ((rc = upload_mat(H->A[0], F64, prm->index_mode, S->A0)))
        return bail(rc);
    S->bytes += S->A0.bytes;
    const int nlev = mg ? nl : 1;
    S->lv.resize(nlev);
    const int mp = prm->mg_prec;
    for (int l = 0; l < nlev; ++l) {
        Level &L = S->lv[l];
        L.n = H->A[l].nrows;
        L.prec = (mg && (mp == 2 || (mp == 1 && l >= 1))) ? F32 : F64;
        L.tprec = (mg && mp != 0) ? F32 : F64;
        L.mn = H->mn[l];
        L.mx = H->mx[l];
        if (l == 0 && L.prec == F64) {
            L.A = S->A0;
            L.own_A = false;
        } else {
            if ((rc = upload_mat(H->A[l], L.prec, l == 0 ? prm->index_mode : 0, L.A)))
                return bail(rc);
            S->bytes += L.A.bytes;
        }
        if (mg && l + 1 < nlev) {
            if ((rc = upload_mat(H->P[l], L.tprec, 0, L.P)) ||
                (rc = upload_mat(H->R[l], L.tprec, 0, L.R)))
                return bail(rc);
            S->bytes += L.P.bytes + L.R.bytes;
        }
        int nb = std::max(1, prm->hsgs_ncores);
        if (prm->hsgs_block_rows > 0)
            nb = (int)std::max<int64_t>(1, (L.n + prm->hsgs_block_rows - 1) / prm->hsgs_block_rows);
        L.hsgs_B = hsgs_block_rows(L.n, nb);
        const size_t cnt = (size_t)L.n * nv;
        if ((rc = alloc_prec(S, &L.sr, cnt, L.prec)) || (rc = alloc_prec(S, &L.su1, cnt, L.prec)) ||
            (rc = alloc_prec(S, &L.su2, cnt, L.prec)) || (rc = alloc_prec(S, &L.tmp, cnt, L.prec)))
            return bail(rc);
        if (L.tprec != L.prec && (rc = alloc_prec(S, &L.tmp_t, cnt, L.tprec)))
            return bail(rc);
        if (l >= 1 && ((rc = alloc_prec(S, &L.b, cnt, L.prec)) ||
                       (rc = alloc_prec(S, &L.x, cnt, L.prec))))
            return bail(rc);
        // Chebyshev coefficients per role (the smoother objects of MultiGrid::setup,
        // mg.inl:113-131, each evaluate them from the level's eigenvalue estimates)
        L.pre = mg ? pre_of(l) : prm->precond;
        L.post = mg ? post_of(l) : prm->post;
        const xh_role *roles[3] = {&L.pre, &L.post, &prm->coarse};
        std::vector<double> *dst[3] = {&L.cheb_pre, &L.cheb_post, &L.cheb_coarse};
        for (int q = 0; q < 3; ++q)
            if (roles[q]->method == XH_CHEBYSHEV &&
                cheby_coeffs(L.mn, L.mx, roles[q]->poly_order, (double)roles[q]->spectrum_fraction,
                             *dst[q]))
                return bail(fail(XH_EINVAL, "Chebyshev order out of range"));
        if (roles[0]->method == XH_HSGS || roles[1]->method == XH_HSGS ||
            roles[2]->method == XH_HSGS) {
            if ((rc = upload_core(H->A[l], L.hsgs_B, L.prec, L.A, L.core)))
                return bail(rc);
            S->bytes += L.core.bytes;
            if (g_tune.hsgs_off_csr) {
                if ((rc = upload_mat(off_block_csr(H->A[l], L.hsgs_B), L.prec, 0, L.off)))
                    return bail(rc);
                S->bytes += L.off.bytes;
            }
        }
    }
    // coarsest direct solve: dense Gauss-Jordan inverse on the host (Direct::init)
    {
        Level &C = S->lv[nlev - 1];
        const bool need = (mg && prm->coarse.method == XH_DIRECT) ||
                          (!mg && prm->precond.method == XH_DIRECT);
        if (need) {
            const HostCsr &Ac = H->A[nlev - 1];
            const int64_t nc = Ac.nrows;
            if (nc > 8192)
                return bail(fail(XH_EINVAL, "Direct coarse solve on a level larger than 8192"));
            C.dinv_prec = C.prec;
            if (mg && mp == 2 && prm->coarse_fp64) {
                C.dinv_prec = F64;
                if ((rc = alloc_vec(S, (double **)&C.cb64, (size_t)nc * nv)) ||
                    (rc = alloc_vec(S, (double **)&C.cx64, (size_t)nc * nv)))
                    return bail(rc);
            }
            if (C.dinv_prec == F64) {
                std::vector<double> dense((size_t)nc * nc, 0.0), inv;
                for (int64_t i = 0; i < nc; ++i)
                    for (uint32_t j = Ac.row[i]; j < Ac.row[i + 1]; ++j)
                        dense[i * nc + Ac.col[j]] = Ac.val[j];
                gj_inverse<double>(nc, dense, inv);
                if ((rc = alloc_vec(S, (double **)&C.dinv, (size_t)nc * nc)))
                    return bail(rc);
                XH_HIP(hipMemcpy(C.dinv, inv.data(), inv.size() * 8, hipMemcpyHostToDevice));
            } else {
                std::vector<float> dense((size_t)nc * nc, 0.f), inv;
                for (int64_t i = 0; i < nc; ++i)
                    for (uint32_t j = Ac.row[i]; j < Ac.row[i + 1]; ++j)
                        dense[i * nc + Ac.col[j]] = (float)Ac.val[j];
                gj_inverse<float>(nc, dense, inv);
                if ((rc = alloc_vec(S, (float **)&C.dinv, (size_t)nc * nc)))
                    return bail(rc);
                XH_HIP(hipMemcpy(C.dinv, inv.data(), inv.size() * 4, hipMemcpyHostToDevice));
            }
        }
    }
    const size_t cnt = (size_t)S->n * nv;
    for (double **q : {&S->r0, &S->r, &S->p, &S->v, &S->s, &S->t, &S->p_, &S->s_, &S->xi, &S->bi})
        if ((rc = alloc_vec(S, q, cnt)))
            return bail(rc);
    if (mg && mp == 2) {
        if ((rc = alloc_prec(S, &S->b32, cnt, F32)) || (rc = alloc_prec(S, &S->x32, cnt, F32)))
            return bail(rc);
    }
    if ((rc = alloc_vec(S, &S->st, 1)) ||
        (rc = alloc_vec(S, &S->partials, (size_t)kMaxPartials * 4 * kMaxNV)))
        return bail(rc);
    S->hist_len = std::max(1, prm->solver.max_iters) + 2;
    if ((rc = alloc_vec(S, &S->hist, (size_t)S->hist_len * nv)))
        return bail(rc);
    XH_HIP(hipHostMalloc((void **)&S->hst, sizeof(SolverState), hipHostMallocDefault));
    XH_HIP(hipHostMalloc((void **)&S->hdone, 2 * sizeof(int), hipHostMallocDefault));
    *out = S;
    return 0;
}

int xh_solver_solve_device(xh_solver S, const double *db, double *dx) {
    if (!S || !db || !dx)
        return fail(XH_EINVAL, "xh_solver_solve_device: bad arguments");
    return solve_impl(S, db, dx);
}

int xh_solver_solve(xh_solver S, const double *b, double *x) {
    if (!S || !b || !x)
        return fail(XH_EINVAL, "xh_solver_solve: bad arguments");
    const size_t bytes = (size_t)S->n * S->nv * sizeof(double);
    XH_HIP(hipMemcpy(S->bi, b, bytes, hipMemcpyHostToDevice));
    XH_HIP(hipMemcpy(S->xi, x, bytes, hipMemcpyHostToDevice));
    int rc = solve_impl(S, S->bi, S->xi);
    if (rc)
        return rc;
    XH_HIP(hipMemcpy(x, S->xi, bytes, hipMemcpyDeviceToHost));
    return 0;
}

int xh_solver_stats(xh_solver S, int32_t *iters, int32_t *converged, double *abs_res,
                    double *rel_res) {
    if (!S)
        return fail(XH_EINVAL, "null solver");
    const SolverState &h = *S->hst;
    if (iters)
        *iters = h.iters;
    for (int k = 0; k < S->nv; ++k) {
        if (converged)
            converged[k] = h.conv[k];
        if (abs_res)
            abs_res[k] = h.abs_res[k];
        if (rel_res)
            rel_res[k] = h.rel_res[k];
    }
    return 0;
}

int xh_solver_history(xh_solver S, double *hist, int32_t max_checks, int32_t *nchecks) {
    if (!S || !hist)
        return fail(XH_EINVAL, "xh_solver_history: bad arguments");
    const int checks = std::min<int>(S->hst->iters + 1, std::min(max_checks, S->hist_len));
    XH_HIP(hipMemcpy(hist, S->hist, (size_t)checks * S->nv * sizeof(double),
                     hipMemcpyDeviceToHost));
    if (nchecks)
        *nchecks = checks;
    return 0;
}

int xh_solver_precond_apply(xh_solver S, const double *db, double *dx) {
    if (!S || !db || !dx)
        return fail(XH_EINVAL, "xh_solver_precond_apply: bad arguments");
    if (!is_preconditioned(S->prm.solver.method))
        return fail(XH_EINVAL, "xh_solver_precond_apply: solver has no preconditioner");
    S->dflag = nullptr;
    precond(S, db, dx);
    int rc = check_launch();
    if (rc)
        return rc;
    XH_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

int xh_solver_smooth(xh_solver S, const double *db, double *dx, int32_t sweeps) {
    if (!S || !db || !dx || sweeps < 1)
        return fail(XH_EINVAL, "xh_solver_smooth: bad arguments");
    if (S->prm.precond.method != XH_MULTIGRID || S->lv.empty())
        return fail(XH_ENOTSUP, "xh_solver_smooth: needs a MultiGrid preconditioner");
    Level &L0 = S->lv[0];
    if (L0.prec != F64)
        return fail(XH_ENOTSUP, "xh_solver_smooth: the finest level is fp32 (mg_prec 2)");
    const xh_role role = L0.pre;
    if (role.method != XH_CHEBYSHEV && role.method != XH_JACOBI && role.method != XH_HSGS)
        return fail(XH_ENOTSUP, "xh_solver_smooth: the finest pre-smoother is not a smoother");
    S->dflag = nullptr;
    ledger_level(0);
    for (int k = 0; k < sweeps; ++k)
        smooth(S, role, L0, L0.cheb_pre, dx, db, false);
    ledger_level(-1);
    int rc = check_launch();
    if (rc)
        return rc;
    XH_HIP(hipStreamSynchronize(S->stream));
    return 0;
}

int xh_solver_profile(xh_solver S, const double *db, double *dx, int32_t iters,
                      xh_kernel_stat *stats, int32_t max_stats, int32_t *nstats) {
    if (!S || !db || !dx || iters < 1 || (max_stats > 0 && !stats) || !nstats)
        return fail(XH_EINVAL, "xh_solver_profile: bad arguments");
    if (g_ledger)
        return fail(XH_EINVAL, "xh_solver_profile: a profile solve is already running");
    // exactly `iters` iterations, launched one by one: no tolerance, no graph
    const xh_params saved = S->prm;
    S->prm.graph = 0;
    S->prm.solver.max_iters = iters;
    S->prm.solver.abs_tol = 0.0f;
    S->prm.solver.rel_tol = 0.0f;
    Ledger L;
    g_ledger = &L;
    int rc = solve_impl(S, db, dx);
    g_ledger = nullptr;
    S->prm = saved;
    if (rc)
        return rc;
    if (L.failed)
        return fail(XH_EHIP, "xh_solver_profile: event record failed");
    XH_HIP(hipStreamSynchronize(S->stream));
    // per (kernel, level) sums over the launches of the init step and iterations 1..iters
    // (launches enqueued behind the last iteration found the solve done and did nothing)
    std::vector<xh_kernel_stat> out;
    std::unordered_map<std::string, size_t> at;
    for (const LedgerRec &r : L.recs) {
        if (r.iter > iters)
            continue;
        float ms = 0.0f;
        XH_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
        const std::string key = r.kernel + "@" + std::to_string(r.level);
        auto it = at.find(key);
        if (it == at.end()) {
            xh_kernel_stat z;
            std::memset(&z, 0, sizeof(z));
            std::snprintf(z.kernel, sizeof(z.kernel), "%s", r.kernel.c_str());
            z.level = r.level;
            it = at.emplace(key, out.size()).first;
            out.push_back(z);
        }
        xh_kernel_stat &k = out[it->second];
        k.launches += 1;
        k.bytes += r.bytes;
        k.ms += ms;
    }
    std::stable_sort(out.begin(), out.end(),
                     [](const xh_kernel_stat &a, const xh_kernel_stat &b) { return a.ms > b.ms; });
    *nstats = (int32_t)out.size();
    for (int i = 0; i < std::min<int>(max_stats, (int)out.size()); ++i)
        stats[i] = out[i];
    return 0;
}

int xh_solver_renew(xh_solver S, const xh_role *solver) {
    if (!S || !solver)
        return fail(XH_EINVAL, "xh_solver_renew: bad arguments");
    if (solver->method != S->prm.solver.method)
        return fail(XH_EINVAL, "xh_solver_renew: the solver method cannot change");
    const int need = std::max(1, solver->max_iters) + 2;
    if (need > S->hist_len) {
        double *h = nullptr;
        XH_HIP(hipMalloc((void **)&h, (size_t)need * S->nv * sizeof(double)));
        (void)hipFree(S->hist);
        S->hist = h;
        S->hist_len = need;
    }
    S->prm.solver = *solver;
    return 0;
}

int xh_solver_info(xh_solver S, int *nlevels, size_t *device_bytes) {
    if (!S)
        return fail(XH_EINVAL, "null solver");
    if (nlevels)
        *nlevels = (int)S->lv.size();
    if (device_bytes)
        *device_bytes = S->bytes;
    return 0;
}

int xh_solver_destroy(xh_solver S) {
    if (!S)
        return 0;
    if (S->stream)
        (void)hipStreamSynchronize(S->stream);
    free_graphs(S);
    for (Level &L : S->lv) {
        if (L.own_A)
            free_mat(L.A);
        L.core.inv_diag = nullptr; // A's
        free_mat(L.core);
        free_mat(L.off);
        free_mat(L.P);
        free_mat(L.R);
        for (void *q : {L.tmp, L.tmp_t, L.b, L.x, L.sr, L.su1, L.su2, L.dinv, L.cb64, L.cx64})
            if (q)
                (void)hipFree(q);
    }
    free_mat(S->A0);
    for (void *q : {(void *)S->r0, (void *)S->r, (void *)S->p, (void *)S->v, (void *)S->s,
                    (void *)S->t, (void *)S->p_, (void *)S->s_, (void *)S->xi, (void *)S->bi,
                    S->b32, S->x32, (void *)S->st, (void *)S->partials, (void *)S->hist})
        if (q)
            (void)hipFree(q);
    if (S->hst)
        (void)hipHostFree(S->hst);
    if (S->hdone)
        (void)hipHostFree(S->hdone);
    for (hipEvent_t e : S->ev)
        if (e)
            (void)hipEventDestroy(e);
    if (S->stream)
        (void)hipStreamDestroy(S->stream);
    delete S;
    return 0;
}

} // extern "C"

extern "C" int xh_copy_device(double *dst, const double *src, size_t count) {
    int rc = ensure_init();
    if (rc)
        return rc;
    if (!dst || !src)
        return fail(XH_EINVAL, "xh_copy_device: null pointer");
    copy_vec((int64_t)count, F64, src, dst, nullptr, g_stream);
    XH_HIP(hipGetLastError());
    XH_HIP(hipStreamSynchronize(g_stream));
    return 0;
}

extern "C" int xh_copy_device_timed(double *dst, const double *src, size_t count, int reps,
                                    double *ms_per_launch) {
    int rc = ensure_init();
    if (rc)
        return rc;
    if (!dst || !src || reps < 1 || !ms_per_launch)
        return fail(XH_EINVAL, "xh_copy_device_timed: bad arguments");
    hipEvent_t e0 = nullptr, e1 = nullptr;
    XH_HIP(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        copy_vec((int64_t)count, F64, src, dst, nullptr, g_stream); // warm-up
        e = hipEventRecord(e0, g_stream);
    }
    for (int i = 0; e == hipSuccess && i < reps; ++i) {
        copy_vec((int64_t)count, F64, src, dst, nullptr, g_stream);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipEventRecord(e1, g_stream);
    if (e == hipSuccess)
        e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess)
        e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    if (e != hipSuccess)
        return fail(XH_EHIP, std::string("xh_copy_device_timed: ") + hipGetErrorString(e));
    *ms_per_launch = (double)ms / reps;
    return 0;
}

extern "C" int xh_set_option(const char *key, int64_t value) {
    if (!key)
        return fail(XH_EINVAL, "xh_set_option: null key");
    const std::string k(key);
    if (k == "spmv_strip_budget" && value > 0)
        g_tune.strip_budget = value;
    else if (k == "spmv_tile_passes" && value >= 1 && value <= 1024)
        g_tune.tile_passes = (int)value;
    else if (k == "spmv_min_passes" && value >= 1 && value <= 1024)
        g_tune.min_passes = (int)value;
    else if (k == "dot_groups" && value >= 8 && value <= kMaxPartials && value % 8 == 0)
        g_tune.dot_groups = (int)value;
    else if (k == "grid_dot_groups" && value >= 8 && value <= kMaxPartials && value % 8 == 0)
        g_tune.grid_dot_groups = (int)value;
    else if (k == "vec_dot_groups" && value >= 8 && value <= kMaxPartials && value % 8 == 0)
        g_tune.vec_dot_groups = (int)value;
    else if (k == "spmv_fold_enable" && (value == 0 || value == 1))
        g_tune.fold_enable = (int)value;
    else if (k == "spmv_fold_min_bytes" && value >= 0)
        g_tune.fold_min_bytes = value;
    else if (k == "cheb_fuse_zero" && (value == 0 || value == 1))
        g_tune.chebz_enable = (int)value;
    else if (k == "hsgs_offblock_csr" && (value == 0 || value == 1))
        g_tune.hsgs_off_csr = (int)value;
    else if (k == "spmv_wide_batches" && (value == 0 || value == 1))
        g_tune.wide_batches = (int)value;
    else if (k == "spmv_grid_enable" && value >= 0 && value <= 2)
        g_tune.grid_enable = (int)value;
    else if (k == "spmv_st7_enable" && (value == 0 || value == 1))
        g_tune.st7_enable = (int)value;
    else if (k == "spmv_st7_waves" && value >= 8 && value <= (1 << 20))
        g_tune.st7_waves = (int)value;
    else if (k == "spmv_st7_min_seg" && value >= 1 && value <= (1 << 24))
        g_tune.st7_min_seg = (int)value;
    else if (k == "spmv_grid_items" && value >= 1 && value <= (1 << 24))
        g_tune.grid_items = (int)value;
    else if (k == "spmv_grid_min_seg" && value >= 1 && value <= (1 << 24))
        g_tune.grid_min_seg = (int)value;
    else if (k == "unet_fuse_up" && (value == 0 || value == 1))
        unet::set_fuse_up((int)value);
    else if (k == "unet_conv_tile" && (value == 0 || value == 1))
        unet::set_conv_tile((int)value);

    else
        return fail(XH_EINVAL, "xh_set_option: unknown key or value out of range: " + k);
    return 0;
}

extern "C" int xh_csr_fold(int64_t nrows, int64_t ncols, const uint32_t *row, const uint32_t *col,
                           int64_t *fold) {
    if (!row || !fold || (nrows > 0 && row[nrows] > 0 && !col))
        return fail(XH_EINVAL, "xh_csr_fold: null pointer");
    *fold = csr_fold(nrows, ncols, row, col, 256);
    return 0;
}

extern "C" int xh_spmv_schedule(int64_t nrows, int64_t nnz, int64_t fold, int prec, int nv,
                                int dots, int64_t *grid, int64_t *ntiles, int64_t *tpw,
                                int64_t *rows) {
    if (nrows < 0 || nnz < 0 || fold < 0 || (prec != F64 && prec != F32) || !nv_ok(nv))
        return fail(XH_EINVAL, "xh_spmv_schedule: bad argument");
    const int vb = prec == F64 ? 8 : 4;
    const int w = std::min(16 / vb, nv);
    const int rpb = kBlock / (nv / w);
    const Launch L = spmv_plan(nrows, nnz, fold, rpb, dots ? g_tune.dot_groups : 16384,
                               (int64_t)nv * vb);
    if (grid)
        *grid = L.grid;
    if (ntiles)
        *ntiles = L.sc.ntiles;
    if (tpw)
        *tpw = L.sc.tpw;
    if (rows)
        for (int64_t k = 0; k < L.sc.ntiles; ++k)
            tile_rows(L.sc, k, rows[2 * k], rows[2 * k + 1]);
    return 0;
}

extern "C" int xh_csr_grid(int64_t nrows, const uint32_t *row, const uint32_t *col,
                           int32_t *dims, uint8_t *code, int64_t *coded) {
    if (!row || !dims || (nrows > 0 && row[nrows] > 0 && !col))
        return fail(XH_EINVAL, "xh_csr_grid: null pointer");
    dims[0] = dims[1] = dims[2] = 0;
    if (coded)
        *coded = 0;
    if (!csr_grid(nrows, row, col, dims[0], dims[1], dims[2]))
        return 0;
    std::vector<uint8_t> tmp;
    if (!code) {
        tmp.resize(row[nrows]);
        code = tmp.data();
    }
    const int64_t c = grid_codes(nrows, row, col, dims[0], dims[1], dims[2], code);
    if (coded)
        *coded = c;
    return 0;
}

extern "C" int xh_grid_schedule(int32_t nx, int32_t ny, int32_t nz, int prec, int nv, int dots,
                                int64_t *grid, int64_t *nitems, int64_t *ipw, int32_t *items) {
    if (nx < 1 || ny < 1 || nz < 1 || (prec != F64 && prec != F32) || !nv_ok(nv))
        return fail(XH_EINVAL, "xh_grid_schedule: bad argument");
    const int vb = prec == F64 ? 8 : 4;
    const int w = std::min(16 / vb, nv);
    const int rpb = kBlock / (nv / w);
    const GridLaunch G = grid_plan(nx, ny, nz, rpb, dots ? g_tune.dot_groups : 16384);
    if (grid)
        *grid = G.grid;
    if (nitems)
        *nitems = G.g.nitems;
    if (ipw)
        *ipw = G.g.ipw;
    if (items)
        for (int k = 0; k < G.g.nitems; ++k) {
            int yb, zb, c0, c1;
            grid_item(G.g, k, yb, zb, c0, c1);
            items[4 * k] = yb * G.g.ty;
            items[4 * k + 1] = zb * G.g.tz;
            items[4 * k + 2] = c0;
            items[4 * k + 3] = c1;
        }
    return 0;
}

extern "C" int xh_matrix_layouts(xh_matrix A, int32_t *flags) {
    if (!A || !flags)
        return fail(XH_EINVAL, "xh_matrix_layouts: bad arguments");
    *flags = (A->m.gcode ? 1 : 0) | (A->m.st7 ? 2 : 0) | (A->m.idx == IDX_U16 ? 4 : 0);
    return 0;
}

extern "C" int xh_matrix_grid(xh_matrix A, int32_t *dims) {
    if (!A || !dims)
        return fail(XH_EINVAL, "xh_matrix_grid: null argument");
    dims[0] = A->m.gx;
    dims[1] = A->m.gy;
    dims[2] = A->m.gz;
    return 0;
}

extern "C" int xh_matrix_fold(xh_matrix A, int64_t set, int64_t *fold) {
    if (!A)
        return fail(XH_EINVAL, "xh_matrix_fold: null matrix");
    if (set >= 0) {
        if (set > 0 && (A->m.nrows != A->m.ncols || set >= A->m.nrows))
            return fail(XH_EINVAL, "xh_matrix_fold: fold must be below nrows of a square matrix");
        A->m.fold = set;
    }
    if (fold)
        *fold = A->m.fold;
    return 0;
}

// Compressed block upload (csr_matrix::compress, csr_matrix.inl:86-170): indices are kept
// at the width given (1, 2 or 4 bytes) and decoded in the SpMM kernel.
extern "C" int xh_matrix_create_compressed(int64_t y_rows, int64_t nrows, int64_t ncols,
                                           int64_t nnz, const uint32_t *row, const void *col,
                                           int col_bytes, const uint32_t *row_ind,
                                           const uint32_t *col_ind, int64_t col_offset,
                                           int64_t x_rows, const double *val, int prec,
                                           xh_matrix *out) {
    int rc = ensure_init();
    if (rc)
        return rc;
    if (!out || !row || (nnz > 0 && (!col || !val)) || nrows < 0 || ncols < 0 || y_rows < 0 ||
        col_offset < 0 || x_rows < 0)
        return fail(XH_EINVAL, "xh_matrix_create_compressed: bad arguments");
    if (col_bytes != 1 && col_bytes != 2 && col_bytes != 4)
        return fail(XH_EINVAL, "xh_matrix_create_compressed: col_bytes must be 1, 2 or 4");
    if (prec != F64 && prec != F32)
        return fail(XH_EINVAL, "xh_matrix_create_compressed: prec must be 0 (f64) or 1 (f32)");
    if ((int64_t)row[nrows] != nnz || row[0] != 0)
        return fail(XH_EINVAL, "xh_matrix_create_compressed: row pointer does not match nnz");
    for (int64_t i = 0; i < nrows; ++i)
        if (row[i + 1] < row[i])
            return fail(XH_EINVAL, "xh_matrix_create_compressed: row pointer not monotone");
    if (row_ind == nullptr && nrows != y_rows)
        return fail(XH_EINVAL, "xh_matrix_create_compressed: row_ind needed when rows were dropped");
    if (row_ind != nullptr)
        for (int64_t i = 0; i < nrows; ++i)
            if ((int64_t)row_ind[i] >= y_rows || (i > 0 && row_ind[i] <= row_ind[i - 1]))
                return fail(XH_EINVAL,
                            "xh_matrix_create_compressed: row_ind must ascend below y_rows");
    if (col_ind != nullptr)
        for (int64_t c = 0; c < ncols; ++c)
            if ((int64_t)col_ind[c] + col_offset >= x_rows)
                return fail(XH_EINVAL,
                            "xh_matrix_create_compressed: col_ind + col_offset beyond x_rows");
    std::vector<uint32_t> cv(nnz);
    for (int64_t j = 0; j < nnz; ++j) {
        cv[j] = col_bytes == 1   ? ((const uint8_t *)col)[j]
                : col_bytes == 2 ? ((const uint16_t *)col)[j]
                                 : ((const uint32_t *)col)[j];
        if ((int64_t)cv[j] >= ncols)
            return fail(XH_EINVAL, "xh_matrix_create_compressed: column index out of range");
    }
    auto *m = new xh_matrix_s();
    Mat &M = m->m;
    M.nrows = nrows;
    M.ncols = ncols;
    M.nnz = nnz;
    M.prec = prec;
    M.y_rows = y_rows;
    M.col_offset = col_offset;
    M.idx = col_bytes == 1 ? IDX_U8 : (col_bytes == 2 ? IDX_U16 : IDX_U32);
    M.panel_shift = 32; // one panel with base 0 for every row (rows < 2^32)
    size_t acc = 0;
    auto fail_free = [&](int code) {
        free_mat(M);
        delete m;
        return code;
    };
    if ((rc = dmalloc(&M.row, nrows + 1, acc)))
        return fail_free(rc);
    hipError_t e = hipMemcpy(M.row, row, (nrows + 1) * 4, hipMemcpyHostToDevice);
    if (M.idx != IDX_U32 && e == hipSuccess) {
        if ((rc = dmalloc(&M.panel_base, 1, acc)))
            return fail_free(rc);
        const uint32_t zero = 0;
        e = hipMemcpy(M.panel_base, &zero, 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && nnz > 0) {
        if (M.idx == IDX_U8) {
            uint8_t *d = nullptr;
            if ((rc = dmalloc(&d, nnz, acc)))
                return fail_free(rc);
            M.col = d;
            std::vector<uint8_t> c8(cv.begin(), cv.end());
            e = hipMemcpy(d, c8.data(), nnz, hipMemcpyHostToDevice);
        } else if (M.idx == IDX_U16) {
            uint16_t *d = nullptr;
            if ((rc = dmalloc(&d, nnz, acc)))
                return fail_free(rc);
            M.col = d;
            std::vector<uint16_t> c16(cv.begin(), cv.end());
            e = hipMemcpy(d, c16.data(), nnz * 2, hipMemcpyHostToDevice);
        } else {
            uint32_t *d = nullptr;
            if ((rc = dmalloc(&d, nnz, acc)))
                return fail_free(rc);
            M.col = d;
            e = hipMemcpy(d, cv.data(), nnz * 4, hipMemcpyHostToDevice);
        }
    }
    if (e == hipSuccess && nnz > 0) {
        if (prec == F64) {
            double *d = nullptr;
            if ((rc = dmalloc(&d, nnz, acc)))
                return fail_free(rc);
            M.val = d;
            e = hipMemcpy(d, val, nnz * 8, hipMemcpyHostToDevice);
        } else {
            std::vector<float> vf(val, val + nnz);
            float *d = nullptr;
            if ((rc = dmalloc(&d, nnz, acc)))
                return fail_free(rc);
            M.val = d;
            e = hipMemcpy(d, vf.data(), nnz * 4, hipMemcpyHostToDevice);
        }
    }
    if (e == hipSuccess && row_ind != nullptr && nrows > 0) {
        if ((rc = dmalloc(&M.row_ind, nrows, acc)))
            return fail_free(rc);
        e = hipMemcpy(M.row_ind, row_ind, nrows * 4, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && col_ind != nullptr && ncols > 0) {
        if ((rc = dmalloc(&M.col_ind, ncols, acc)))
            return fail_free(rc);
        e = hipMemcpy(M.col_ind, col_ind, ncols * 4, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess)
        return fail_free(fail(XH_EHIP, std::string("xh_matrix_create_compressed: ") +
                                           hipGetErrorString(e)));
    M.bytes = acc;
    *out = m;
    return 0;
}

extern "C" int xh_matrix_gather(xh_matrix A, int nv, const void *dx_full, void *dx_c) {
    if (!A || !nv_ok(nv) || !dx_full || !dx_c)
        return fail(XH_EINVAL, "xh_matrix_gather: bad arguments");
    if (A->m.col_ind == nullptr)
        return fail(XH_EINVAL, "xh_matrix_gather: the matrix has no col_ind");
    gather_cols(A->m.ncols, nv, A->m.prec, A->m.col_ind, A->m.col_offset, dx_full, dx_c,
                g_stream);
    XH_HIP(hipGetLastError());
    XH_HIP(hipStreamSynchronize(g_stream));
    return 0;
}

// ---------------------------------------------------------------------------------------
// Batched independent systems
// ---------------------------------------------------------------------------------------

struct xh_batch_s {
    int nsys = 0;
    xh_params prm;
    std::vector<int64_t> n, rbase, voff;
    std::vector<int> nv;
    int64_t rows = 0, nnz = 0, velems = 0;
    hipStream_t stream = nullptr;
    // device
    uint32_t *row = nullptr, *col = nullptr;
    double *val = nullptr, *isd = nullptr, *cheb = nullptr;
    int64_t *d_rbase = nullptr, *d_voff = nullptr;
    int *d_nv = nullptr, *d_item_start = nullptr, *d_item_count = nullptr, *d_flag = nullptr;
    BatchItem *d_items = nullptr;
    // the items of the systems not yet done (rebuilt after every iteration) and their count
    BatchItem *d_items_act = nullptr;
    int *d_nact = nullptr;
    SolverState *st = nullptr;
    double *partials = nullptr;
    double *x = nullptr, *b = nullptr, *r0 = nullptr, *r = nullptr, *p = nullptr, *v = nullptr;
    double *s = nullptr, *t = nullptr, *p_ = nullptr, *s_ = nullptr, *cu1 = nullptr,
           *cu2 = nullptr;
    unsigned long long *d_t0 = nullptr;
    BatchBarrier *d_bar = nullptr; // grid barriers of the single-launch solve
    // host
    std::vector<BatchItem> items;
    std::vector<int> item_start, item_count;
    int tile_passes[5] = {8, 8, 8, 8, 8}; // per nv class 1, 2, 4, 8, 16
    int tile_w[5] = {1, 1, 1, 1, 1};      // RHS per lane per nv class (1 or 2)
    int grid = 2048;
    int persistent = 0; // 1: the whole solve is one cooperative launch (k_batch_solve)
    SolverState *hst = nullptr; // pinned [nsys]
    int *hflag = nullptr;       // pinned [2][2]
    unsigned long long *ht0 = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    double wall_khz = 0.0;
    size_t bytes = 0;
};

namespace xh {
namespace {

int nv_class(int nv) { return nv == 1 ? 0 : nv == 2 ? 1 : nv == 4 ? 2 : nv == 8 ? 3 : 4; }

// work items of the systems whose nv class is in `mask`, tile_passes[class] passes of
// 256/nv rows each; item_start/count filled when all classes are included
void build_items(xh_batch_s *B, unsigned mask, std::vector<BatchItem> &items,
                 std::vector<int> *start, std::vector<int> *count) {
    items.clear();
    if (start) {
        start->assign(B->nsys, 0);
        count->assign(B->nsys, 0);
    }
    for (int s = 0; s < B->nsys; ++s) {
        const int c = nv_class(B->nv[s]);
        if (!(mask & (1u << c)))
            continue;
        const int wpl = B->nv[s] >= 2 ? B->tile_w[c] : 1;
        const int64_t tile = (int64_t)B->tile_passes[c] * (kBlock * wpl / B->nv[s]);
        if (start)
            (*start)[s] = (int)items.size();
        for (int64_t r0 = 0; r0 < B->n[s]; r0 += tile)
            items.push_back({s, B->nv[s], r0, std::min(B->n[s], r0 + tile), wpl,
                             (int32_t)items.size()});
        if (count)
            (*count)[s] = (int)items.size() - (*start)[s];
    }
}

int upload_items(xh_batch_s *B) {
    build_items(B, 0x1f, B->items, &B->item_start, &B->item_count);
    (void)hipFree(B->d_items);
    (void)hipFree(B->d_items_act);
    (void)hipFree(B->partials);
    B->d_items = nullptr;
    B->d_items_act = nullptr;
    B->partials = nullptr;
    size_t acc = 0;
    int rc;
    if ((rc = dmalloc(&B->d_items, B->items.size(), acc)) ||
        (rc = dmalloc(&B->d_items_act, B->items.size(), acc)) ||
        (rc = dmalloc(&B->partials, B->items.size() * kBatchPart, acc)))
        return rc;
    XH_HIP(hipMemcpy(B->d_items, B->items.data(), B->items.size() * sizeof(BatchItem),
                     hipMemcpyHostToDevice));
    XH_HIP(hipMemcpy(B->d_item_start, B->item_start.data(), B->nsys * sizeof(int),
                     hipMemcpyHostToDevice));
    XH_HIP(hipMemcpy(B->d_item_count, B->item_count.data(), B->nsys * sizeof(int),
                     hipMemcpyHostToDevice));
    if (B->gexec) {
        (void)hipGraphExecDestroy(B->gexec);
        (void)hipGraphDestroy(B->graph);
        B->gexec = nullptr;
        B->graph = nullptr;
    }
    return 0;
}

BatchDev batch_dev(const xh_batch_s *B) {
    BatchDev d;
    d.row = B->row;
    d.col = B->col;
    d.val = B->val;
    d.isd = B->isd;
    d.rbase = B->d_rbase;
    d.voff = B->d_voff;
    d.st = B->st;
    d.cheb = B->cheb;
    return d;
}

// x = M^-1 b from a zero x: Chebyshev polynomial (smooth() with x_zero) or a copy, over
// the work list of the systems not yet done
void batch_precond(xh_batch_s *B, const double *b, double *x) {
    const BatchDev d = batch_dev(B);
    const int nit = (int)B->items.size();
    const BatchItem *items = B->d_items_act;
    const int *na = B->d_nact;
    hipStream_t st = B->stream;
    if (B->prm.precond.method == XH_CHEBYSHEV) {
        const int k = B->prm.precond.poly_order;
        BatchVecs v;
        v.b = b;
        v.out = k == 1 ? x : B->cu1;
        v.out2 = x;
        int first = k - 2;
        if (k >= 2) {
            // the zero-x first step and the next one in one launch (bit-identical)
            BatchVecs z;
            z.in = b;
            z.out = B->cu1;
            z.out2 = x;
            batch_launch(B_CHEBZ, d, items, nit, B->grid, z, k - 2, k == 2, nullptr, st, na);
            first = k - 3;
        } else {
            batch_launch(B_CHEB0Z, d, items, nit, B->grid, v, k - 1, k == 1, nullptr, st, na);
        }
        double *uc = B->cu1, *un = B->cu2;
        for (int i = first; i >= 0; --i) {
            BatchVecs w;
            w.in = uc;
            w.o1 = b;
            w.out = un;
            w.out2 = x;
            batch_launch(B_CHEB, d, items, nit, B->grid, w, i, i == 0, nullptr, st, na);
            std::swap(uc, un);
        }
        return;
    }
    BatchVecs v;
    v.b = b;
    v.out = x;
    batch_launch(B_COPY, d, items, nit, B->grid, v, 0, 0, nullptr, st, na);
}

// one iteration over the work list of the systems not yet done (items of a system that
// finishes during the iteration are skipped by its done flag); the list for the next
// iteration is rebuilt at the end
void batch_iteration(xh_batch_s *B) {
    const BatchDev d = batch_dev(B);
    const int nit = (int)B->items.size();
    const BatchItem *items = B->d_items_act;
    const int *na = B->d_nact;
    hipStream_t st = B->stream;
    const bool pre = B->prm.solver.method != XH_BICGSTAB;
    double *pp = pre ? B->p_ : B->p, *ss = pre ? B->s_ : B->s;
    auto fin = [&](FinOp op) {
        batch_finalize(op, B->st, B->partials, B->d_item_start, B->d_item_count, B->d_nv,
                       B->nsys, st);
    };
    if (pre)
        batch_precond(B, B->p, B->p_);
    BatchVecs v1;
    v1.in = pp;
    v1.o1 = B->r0;
    v1.out = B->v;
    batch_launch(B_DOT1, d, items, nit, B->grid, v1, 0, 0, B->partials, st, na);
    fin(FIN_ALPHA);
    BatchVecs vs;
    vs.b = B->r;
    vs.o1 = B->v;
    vs.out = B->s;
    batch_launch(B_S_UPDATE, d, items, nit, B->grid, vs, 0, 0, nullptr, st, na);
    if (pre)
        batch_precond(B, B->s, B->s_);
    BatchVecs v3;
    v3.in = ss;
    v3.o1 = B->s;
    v3.out = B->t;
    batch_launch(B_DOT3, d, items, nit, B->grid, v3, 0, 0, B->partials, st, na);
    fin(FIN_OMEGA);
    BatchVecs vx;
    vx.x = B->x;
    vx.p_ = pp;
    vx.s_ = ss;
    vx.s = B->s;
    vx.t = B->t;
    vx.o1 = B->r0;
    vx.out = B->r;
    batch_launch(B_XR_UPDATE, d, items, nit, B->grid, vx, 0, 0, B->partials, st, na);
    fin(FIN_BETA);
    BatchVecs vp;
    vp.b = B->r;
    vp.o1 = B->v;
    vp.out = B->p;
    batch_launch(B_P_UPDATE, d, items, nit, B->grid, vp, 0, 0, nullptr, st, na);
    batch_all_done(B->st, B->nsys, B->d_flag, st);
    batch_compact(B->d_items, B->d_item_start, B->d_item_count, B->st, B->nsys,
                  B->d_items_act, B->d_nact, st);
}

void free_batch(xh_batch_s *B) {
    if (B->stream)
        (void)hipStreamSynchronize(B->stream);
    if (B->gexec)
        (void)hipGraphExecDestroy(B->gexec);
    if (B->graph)
        (void)hipGraphDestroy(B->graph);
    for (void *q : {(void *)B->row, (void *)B->col, (void *)B->val, (void *)B->isd,
                    (void *)B->cheb, (void *)B->d_rbase, (void *)B->d_voff, (void *)B->d_nv,
                    (void *)B->d_item_start, (void *)B->d_item_count, (void *)B->d_flag,
                    (void *)B->d_items, (void *)B->d_items_act, (void *)B->d_nact,
                    (void *)B->st, (void *)B->partials, (void *)B->x,
                    (void *)B->b, (void *)B->r0, (void *)B->r, (void *)B->p, (void *)B->v,
                    (void *)B->s, (void *)B->t, (void *)B->p_, (void *)B->s_,
                    (void *)B->cu1, (void *)B->cu2, (void *)B->d_t0, (void *)B->d_bar})
        if (q)
            (void)hipFree(q);
    if (B->hst)
        (void)hipHostFree(B->hst);
    if (B->hflag)
        (void)hipHostFree(B->hflag);
    if (B->ht0)
        (void)hipHostFree(B->ht0);
    for (hipEvent_t e : B->ev)
        if (e)
            (void)hipEventDestroy(e);
    if (B->stream)
        (void)hipStreamDestroy(B->stream);
    delete B;
}

} // namespace
} // namespace xh

extern "C" int xh_batch_create(int nsys, const int64_t *nrows, const int32_t *nvs,
                               const uint32_t *const *row, const uint32_t *const *col,
                               const double *const *val, const double *lmax,
                               const xh_params *prm, xh_batch *out) {
    int rc = ensure_init();
    if (rc)
        return rc;
    if (nsys < 1 || !nrows || !nvs || !row || !col || !val || !prm || !out)
        return fail(XH_EINVAL, "xh_batch_create: bad arguments");
    const int m = prm->solver.method;
    if (m != XH_BICGSTAB && m != XH_PBICGSTAB && m != XH_MERGED_PBICGSTAB)
        return fail(XH_ENOTSUP, "xh_batch_create: solver must be BiCGStab, PBiCGStab or "
                                "MergedPBiCGStab");
    const int pm = prm->precond.method;
    if (m != XH_BICGSTAB && pm != XH_CHEBYSHEV && pm != XH_IDENTITY)
        return fail(XH_ENOTSUP, "xh_batch_create: batched preconditioner must be Chebyshev "
                                "or Identity");
    if (m != XH_BICGSTAB && pm == XH_CHEBYSHEV &&
        (prm->precond.poly_order < 1 || prm->precond.poly_order > kChebMax || !lmax))
        return fail(XH_EINVAL, "xh_batch_create: Chebyshev order 1..16 and lmax needed");
    auto *B = new xh_batch_s();
    B->nsys = nsys;
    B->prm = *prm;
    int64_t rows = 0, nnz = 0, ve = 0;
    for (int s = 0; s < nsys; ++s) {
        const int v = nvs[s];
        if (!(v == 1 || v == 2 || v == 4 || v == 8 || v == 16) || nrows[s] < 1 || !row[s] ||
            row[s][0] != 0) {
            free_batch(B);
            return fail(XH_EINVAL, "xh_batch_create: system " + std::to_string(s) +
                                       ": nv must be 1, 2, 4, 8 or 16 and rows >= 1");
        }
        const int64_t n = nrows[s], z = row[s][n];
        for (int64_t i = 0; i < n; ++i)
            if (row[s][i + 1] < row[s][i]) {
                free_batch(B);
                return fail(XH_EINVAL, "xh_batch_create: row pointer not monotone");
            }
        for (int64_t j = 0; j < z; ++j)
            if ((int64_t)col[s][j] >= n) {
                free_batch(B);
                return fail(XH_EINVAL, "xh_batch_create: column index out of range");
            }
        B->n.push_back(n);
        B->nv.push_back(v);
        B->rbase.push_back(rows);
        ve = (ve + 31) / 32 * 32; // each system's vectors start on 256 B (16-B vector loads)
        B->voff.push_back(ve);
        rows += n;
        nnz += z;
        ve += n * v;
    }
    if (nnz >= ((int64_t)1 << 32)) {
        free_batch(B);
        return fail(XH_EINVAL, "xh_batch_create: more than 2^32-1 nonzeros in the batch");
    }
    B->rows = rows;
    B->nnz = nnz;
    B->velems = ve;
    // concatenated CSR, inverse square-root diagonals, Chebyshev coefficients
    std::vector<uint32_t> hrow(rows + 1), hcol(nnz);
    std::vector<double> hval(nnz), hisd(rows, 0.0), hcheb((size_t)nsys * kChebMax, 0.0);
    int64_t zr = 0;
    for (int s = 0; s < nsys; ++s) {
        const int64_t n = B->n[s], g0 = B->rbase[s];
        for (int64_t i = 0; i < n; ++i) {
            hrow[g0 + i] = (uint32_t)(zr + row[s][i]);
            for (uint32_t j = row[s][i]; j < row[s][i + 1]; ++j)
                if (col[s][j] == (uint32_t)i)
                    hisd[g0 + i] = std::sqrt(1.0 / val[s][j]);
        }
        std::copy(col[s], col[s] + row[s][n], hcol.begin() + zr);
        std::copy(val[s], val[s] + row[s][n], hval.begin() + zr);
        zr += row[s][n];
        if (m != XH_BICGSTAB && pm == XH_CHEBYSHEV) {
            std::vector<double> c;
            cheby_coeffs(0.0, lmax[s], prm->precond.poly_order,
                         (double)prm->precond.spectrum_fraction, c);
            std::copy(c.begin(), c.end(), hcheb.begin() + (size_t)s * kChebMax);
        }
    }
    hrow[rows] = (uint32_t)zr;
    size_t &acc = B->bytes;
    auto bail = [&](int code) {
        free_batch(B);
        return code;
    };
    if ((rc = dmalloc(&B->row, rows + 1, acc)) || (rc = dmalloc(&B->col, std::max<int64_t>(1, nnz), acc)) ||
        (rc = dmalloc(&B->val, std::max<int64_t>(1, nnz), acc)) || (rc = dmalloc(&B->isd, rows, acc)) ||
        (rc = dmalloc(&B->cheb, (size_t)nsys * kChebMax, acc)) ||
        (rc = dmalloc(&B->d_rbase, nsys, acc)) || (rc = dmalloc(&B->d_voff, nsys, acc)) ||
        (rc = dmalloc(&B->d_nv, nsys, acc)) || (rc = dmalloc(&B->d_item_start, nsys, acc)) ||
        (rc = dmalloc(&B->d_item_count, nsys, acc)) || (rc = dmalloc(&B->d_flag, 2, acc)) ||
        (rc = dmalloc(&B->d_nact, 1, acc)) ||
        (rc = dmalloc(&B->st, nsys, acc)) || (rc = dmalloc(&B->d_t0, 1, acc)) ||
        (rc = dmalloc(&B->d_bar, 1, acc)))
        return bail(rc);
    for (double **q : {&B->x, &B->b, &B->r0, &B->r, &B->p, &B->v, &B->s, &B->t, &B->p_, &B->s_,
                       &B->cu1, &B->cu2})
        if ((rc = dmalloc(q, ve, acc)))
            return bail(rc);
    hipError_t e = hipSuccess;
    auto cp = [&](void *dst, const void *src, size_t bytes) {
        if (e == hipSuccess && bytes)
            e = hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice);
    };
    cp(B->row, hrow.data(), (rows + 1) * 4);
    cp(B->col, hcol.data(), nnz * 4);
    cp(B->val, hval.data(), nnz * 8);
    cp(B->isd, hisd.data(), rows * 8);
    cp(B->cheb, hcheb.data(), hcheb.size() * 8);
    cp(B->d_rbase, B->rbase.data(), nsys * 8);
    cp(B->d_voff, B->voff.data(), nsys * 8);
    cp(B->d_nv, B->nv.data(), nsys * 4);
    if (e == hipSuccess)
        e = hipMemset(B->b, 0, ve * 8);
    if (e == hipSuccess)
        e = hipStreamCreate(&B->stream);
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&B->hst, sizeof(SolverState) * nsys);
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&B->hflag, sizeof(int) * 4);
    if (e == hipSuccess)
        e = hipHostMalloc((void **)&B->ht0, sizeof(unsigned long long));
    for (hipEvent_t &ev : B->ev)
        if (e == hipSuccess)
            e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    int khz = 0;
    if (e == hipSuccess)
        e = hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, g_device);
    if (e != hipSuccess)
        return bail(fail(XH_EHIP, std::string("xh_batch_create: ") + hipGetErrorString(e)));
    B->wall_khz = khz;
    if ((rc = upload_items(B)))
        return bail(rc);
    *out = B;
    return 0;
}

extern "C" int xh_batch_set_rhs(xh_batch B, int s, const double *b) {
    if (!B || s < 0 || s >= B->nsys || !b)
        return fail(XH_EINVAL, "xh_batch_set_rhs: bad arguments");
    XH_HIP(hipMemcpy(B->b + B->voff[s], b, B->n[s] * B->nv[s] * 8, hipMemcpyHostToDevice));
    return 0;
}

extern "C" int xh_batch_get_x(xh_batch B, int s, double *x) {
    if (!B || s < 0 || s >= B->nsys || !x)
        return fail(XH_EINVAL, "xh_batch_get_x: bad arguments");
    XH_HIP(hipMemcpy(x, B->x + B->voff[s], B->n[s] * B->nv[s] * 8, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int xh_batch_solve(xh_batch B) {
    if (!B)
        return fail(XH_EINVAL, "xh_batch_solve: null batch");
    const int nsys = B->nsys;
    for (int s = 0; s < nsys; ++s) {
        SolverState &h = B->hst[s];
        std::memset(&h, 0, sizeof(SolverState));
        for (int k = 0; k < B->nv[s]; ++k)
            h.iconv[k] = 1.0;
        h.nv = B->nv[s];
        h.max_iters = B->prm.solver.max_iters;
        h.conv_check = B->prm.solver.conv_check;
        h.abs_tol = B->prm.solver.abs_tol;
        h.rel_tol = B->prm.solver.rel_tol;
    }
    hipStream_t st = B->stream;
    XH_HIP(hipMemcpyAsync(B->st, B->hst, sizeof(SolverState) * nsys, hipMemcpyHostToDevice, st));
    fill_zero(B->velems, F64, B->x, nullptr, st);
    stamp_clock(B->d_t0, st);
    if (B->persistent) {
        // init and every iteration in one cooperative launch; no host round trip
        BatchSolve P;
        P.d = batch_dev(B);
        P.items = B->d_items;
        P.nitems = (int)B->items.size();
        P.nsys = nsys;
        P.max_iters = std::max(1, B->prm.solver.max_iters);
        P.pre = B->prm.solver.method != XH_BICGSTAB;
        P.order = B->prm.precond.method == XH_CHEBYSHEV ? B->prm.precond.poly_order : 0;
        P.item_start = B->d_item_start;
        P.item_count = B->d_item_count;
        P.nvs = B->d_nv;
        P.partials = B->partials;
        P.x = B->x;
        P.b = B->b;
        P.r0 = B->r0;
        P.r = B->r;
        P.p = B->p;
        P.v = B->v;
        P.s = B->s;
        P.t = B->t;
        P.p_ = B->p_;
        P.s_ = B->s_;
        P.cu1 = B->cu1;
        P.cu2 = B->cu2;
        const int cap = batch_solve_blocks();
        if (cap < 1)
            return fail(XH_EHIP, "xh_batch_solve: no resident workgroup count for the "
                                 "single-launch solve");
        const int grid = std::max(1, std::min(cap, std::max(P.nitems, nsys)));
        P.bar = B->d_bar;
        P.wait_ticks = (unsigned long long)(B->wall_khz * 2000.0); // 2 s per barrier
        XH_HIP(hipMemsetAsync(B->d_bar, 0, sizeof(BatchBarrier), st));
        XH_HIP(batch_solve_launch(P, grid, st));
        BatchBarrier hb;
        XH_HIP(hipMemcpyAsync(&hb, B->d_bar, sizeof(BatchBarrier), hipMemcpyDeviceToHost, st));
        XH_HIP(hipStreamSynchronize(st));
        if (hb.err)
            return fail(XH_EHIP, "xh_batch_solve: a grid barrier of the single-launch solve "
                                 "timed out (the kernel ended; the solution is not valid)");
    } else {
    // r0 = b - A x ; p = r = r0 ; rho0 (get_residual + copies, pbicgstab.inl:111-117)
    BatchVecs v;
    v.in = B->x;
    v.b = B->b;
    v.out = B->r0;
    v.out2 = B->p;
    v.out3 = B->r;
    const BatchDev d = batch_dev(B);
    batch_launch(B_RESID0, d, B->d_items, (int)B->items.size(), B->grid, v, 0, 0, B->partials, st);
    batch_finalize(FIN_INIT, B->st, B->partials, B->d_item_start, B->d_item_count, B->d_nv,
                   nsys, st);
    batch_compact(B->d_items, B->d_item_start, B->d_item_count, B->st, nsys, B->d_items_act,
                  B->d_nact, st);
    int rc = check_launch();
    if (rc)
        return rc;
    const bool use_graph = B->prm.graph != 0;
    if (use_graph && B->gexec == nullptr) {
        XH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        batch_iteration(B);
        XH_HIP(hipStreamEndCapture(st, &B->graph));
        XH_HIP(hipGraphInstantiate(&B->gexec, B->graph, nullptr, nullptr, 0));
    }
    const int maxit = std::max(1, B->prm.solver.max_iters) + 1;
    for (int it = 0; it < maxit; ++it) {
        if (use_graph) {
            XH_HIP(hipGraphLaunch(B->gexec, st));
        } else {
            batch_iteration(B);
            if ((rc = check_launch()))
                return rc;
        }
        XH_HIP(hipMemcpyAsync(&B->hflag[2 * (it & 1)], B->d_flag, 2 * sizeof(int),
                              hipMemcpyDeviceToHost, st));
        XH_HIP(hipEventRecord(B->ev[it & 1], st));
        if (it > 0) {
            XH_HIP(hipEventSynchronize(B->ev[(it - 1) & 1]));
            if (B->hflag[2 * ((it - 1) & 1)])
                break;
        }
    }
    }
    XH_HIP(hipMemcpyAsync(B->hst, B->st, sizeof(SolverState) * nsys, hipMemcpyDeviceToHost, st));
    XH_HIP(hipMemcpyAsync(B->ht0, B->d_t0, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    XH_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < nsys; ++s)
        if (!B->hst[s].done)
            return fail(XH_EINVAL, "xh_batch_solve: system " + std::to_string(s) +
                                       " did not terminate within max_iters");
    return 0;
}

extern "C" int xh_batch_stats(xh_batch B, int s, int32_t *iters, int32_t *conv,
                              double *abs_res, double *rel_res, double *latency_ms) {
    if (!B || s < 0 || s >= B->nsys)
        return fail(XH_EINVAL, "xh_batch_stats: bad arguments");
    const SolverState &h = B->hst[s];
    if (iters)
        *iters = h.iters;
    for (int k = 0; k < B->nv[s]; ++k) {
        if (conv)
            conv[k] = h.conv[k];
        if (abs_res)
            abs_res[k] = h.abs_res[k];
        if (rel_res)
            rel_res[k] = h.rel_res[k];
    }
    if (latency_ms)
        *latency_ms = B->wall_khz > 0 ? (double)(h.t_done - *B->ht0) / B->wall_khz : 0.0;
    return 0;
}

extern "C" int xh_batch_autotune(xh_batch B, int reps, int32_t *tile_passes) {
    if (!B || reps < 1)
        return fail(XH_EINVAL, "xh_batch_autotune: bad arguments");
    static const int cand[] = {1, 2, 4, 8, 16, 32};
    hipStream_t st = B->stream;
    hipEvent_t e0, e1;
    XH_HIP(hipEventCreate(&e0));
    XH_HIP(hipEventCreate(&e1));
    // every system "not done" so the timed launches do their work
    for (int s = 0; s < B->nsys; ++s) {
        std::memset(&B->hst[s], 0, sizeof(SolverState));
        B->hst[s].nv = B->nv[s];
    }
    XH_HIP(hipMemcpy(B->st, B->hst, sizeof(SolverState) * B->nsys, hipMemcpyHostToDevice));
    std::vector<BatchItem> items;
    BatchItem *d_items = nullptr;
    double *d_part = nullptr;
    int rc = 0;
    for (int c = 0; c < 5 && rc == 0; ++c) {
        bool present = false;
        for (int s = 0; s < B->nsys; ++s)
            present = present || nv_class(B->nv[s]) == c;
        if (!present)
            continue;
        double best = 1e30;
        int best_p = B->tile_passes[c], best_w = B->tile_w[c];
        for (int cw = 0; cw < (c == 0 ? 6 : 12); ++cw) {
            const int p = cand[cw % 6], wl = 1 + cw / 6;
            B->tile_passes[c] = p;
            B->tile_w[c] = wl;
            build_items(B, 1u << c, items, nullptr, nullptr);
            (void)hipFree(d_items);
            (void)hipFree(d_part);
            size_t acc = 0;
            if ((rc = dmalloc(&d_items, items.size(), acc)) ||
                (rc = dmalloc(&d_part, items.size() * kBatchPart, acc)))
                break;
            XH_HIP(hipMemcpy(d_items, items.data(), items.size() * sizeof(BatchItem),
                             hipMemcpyHostToDevice));
            BatchVecs v;
            v.in = B->b;
            v.o1 = B->b;
            v.out = B->t;
            const BatchDev d = batch_dev(B);
            batch_launch(B_DOT1, d, d_items, (int)items.size(), B->grid, v, 0, 0, d_part, st);
            XH_HIP(hipEventRecord(e0, st));
            for (int q = 0; q < reps; ++q)
                batch_launch(B_DOT1, d, d_items, (int)items.size(), B->grid, v, 0, 0, d_part, st);
            XH_HIP(hipEventRecord(e1, st));
            XH_HIP(hipEventSynchronize(e1));
            float ms = 0;
            XH_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) {
                best = ms;
                best_p = p;
                best_w = wl;
            }
        }
        B->tile_passes[c] = best_p;
        B->tile_w[c] = best_w;
    }
    (void)hipFree(d_items);
    (void)hipFree(d_part);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc)
        return rc;
    if (tile_passes)
        for (int c = 0; c < 5; ++c)
            tile_passes[c] = B->tile_passes[c];
    return upload_items(B);
}

// Times the batched Chebyshev step (B_CHEB: y = isd (c (b isd) + isd (A u)) over every work item
// of every system, as in the preconditioner) with every system active: reps launches
// between two events on the batch stream.  The systems' solver state (device and its host
// copy) is saved first and restored afterwards; the work vector t is overwritten.
extern "C" int xh_batch_spmv_timed(xh_batch B, int reps, double *us_per_launch,
                                   double *alg_bytes) {
    if (!B || reps < 1 || !us_per_launch || !alg_bytes)
        return fail(XH_EINVAL, "xh_batch_spmv_timed: bad arguments");
    hipStream_t st = B->stream;
    XH_HIP(hipStreamSynchronize(st));
    std::vector<SolverState> saved(B->nsys), saved_host(B->hst, B->hst + B->nsys);
    XH_HIP(hipMemcpy(saved.data(), B->st, sizeof(SolverState) * B->nsys, hipMemcpyDeviceToHost));
    for (int s = 0; s < B->nsys; ++s) {
        std::memset(&B->hst[s], 0, sizeof(SolverState));
        B->hst[s].nv = B->nv[s];
    }
    XH_HIP(hipMemcpy(B->st, B->hst, sizeof(SolverState) * B->nsys, hipMemcpyHostToDevice));
    hipEvent_t e0, e1;
    XH_HIP(hipEventCreate(&e0));
    XH_HIP(hipEventCreate(&e1));
    BatchVecs v;
    v.in = B->b;
    v.o1 = B->b;
    v.out = B->t;
    const BatchDev d = batch_dev(B);
    const int nit = (int)B->items.size();
    batch_launch(B_CHEB, d, B->d_items, nit, B->grid, v, 0, 0, nullptr, st);
    XH_HIP(hipEventRecord(e0, st));
    for (int q = 0; q < reps; ++q)
        batch_launch(B_CHEB, d, B->d_items, nit, B->grid, v, 0, 0, nullptr, st);
    XH_HIP(hipEventRecord(e1, st));
    XH_HIP(hipEventSynchronize(e1));
    float ms = 0;
    XH_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    XH_HIP(hipMemcpy(B->st, saved.data(), sizeof(SolverState) * B->nsys, hipMemcpyHostToDevice));
    std::copy(saved_host.begin(), saved_host.end(), B->hst);
    double vec = 0.0;
    for (int s = 0; s < B->nsys; ++s)
        vec += (double)B->n[s] * B->nv[s] * 8.0;
    // matrix values and columns, row pointers, inverse square-root diagonal, the work
    // items; u and r read, y written
    *alg_bytes = (double)B->nnz * 12.0 + (double)B->rows * (4.0 + 8.0) +
                 (double)nit * sizeof(BatchItem) + 3.0 * vec;
    *us_per_launch = 1000.0 * ms / reps;
    return 0;
}

extern "C" int xh_batch_set_tiles(xh_batch B, const int32_t *tile_passes) {
    if (!B || !tile_passes)
        return fail(XH_EINVAL, "xh_batch_set_tiles: bad arguments");
    for (int c = 0; c < 5; ++c)
        if (tile_passes[c] < 1 || tile_passes[c] > 1024)
            return fail(XH_EINVAL, "xh_batch_set_tiles: tile passes must be 1..1024");
    for (int c = 0; c < 5; ++c)
        B->tile_passes[c] = tile_passes[c];
    return upload_items(B);
}

extern "C" int xh_batch_get_layout(xh_batch B, int32_t *tile_passes, int32_t *rhs_per_lane) {
    if (!B)
        return fail(XH_EINVAL, "xh_batch_get_layout: null batch");
    for (int c = 0; c < 5; ++c) {
        if (tile_passes)
            tile_passes[c] = B->tile_passes[c];
        if (rhs_per_lane)
            rhs_per_lane[c] = c == 0 ? 1 : B->tile_w[c];
    }
    return 0;
}

extern "C" int xh_batch_set_layout(xh_batch B, const int32_t *tile_passes,
                                   const int32_t *rhs_per_lane) {
    if (!B || !tile_passes || !rhs_per_lane)
        return fail(XH_EINVAL, "xh_batch_set_layout: bad arguments");
    for (int c = 0; c < 5; ++c)
        if (tile_passes[c] < 1 || tile_passes[c] > 1024 ||
            (rhs_per_lane[c] != 1 && rhs_per_lane[c] != 2) || (c == 0 && rhs_per_lane[c] != 1))
            return fail(XH_EINVAL, "xh_batch_set_layout: passes 1..1024, 1 or 2 RHS per lane "
                                   "(1 for nv = 1)");
    for (int c = 0; c < 5; ++c) {
        B->tile_passes[c] = tile_passes[c];
        B->tile_w[c] = rhs_per_lane[c];
    }
    return upload_items(B);
}

extern "C" int xh_batch_set_persistent(xh_batch B, int on) {
    if (!B || (on != 0 && on != 1))
        return fail(XH_EINVAL, "xh_batch_set_persistent: bad arguments");
    B->persistent = on;
    return 0;
}

extern "C" int xh_batch_destroy(xh_batch B) {
    if (B)
        free_batch(B);
    return 0;
}

extern "C" int xh_eval_bumps(int nx, int ny, int nz, int nv, int max_bumps,
                             const int32_t *count, const double *params, double *db) {
    int rc = ensure_init();
    if (rc)
        return rc;
    if (nx < 1 || ny < 1 || nz < 1 || nv < 1 || max_bumps < 1 || !count || !params || !db)
        return fail(XH_EINVAL, "xh_eval_bumps: bad arguments");
    for (int k = 0; k < nv; ++k)
        if (count[k] < 0 || count[k] > max_bumps)
            return fail(XH_EINVAL, "xh_eval_bumps: count out of range");
    int *dc = nullptr;
    double *dp = nullptr;
    XH_HIP(hipMalloc((void **)&dc, nv * sizeof(int)));
    hipError_t e = hipMalloc((void **)&dp, (size_t)nv * max_bumps * 5 * sizeof(double));
    if (e == hipSuccess)
        e = hipMemcpy(dc, count, nv * sizeof(int), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(dp, params, (size_t)nv * max_bumps * 5 * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        eval_bumps(nx, ny, nz, nv, max_bumps, dc, dp, db, g_stream);
        e = hipStreamSynchronize(g_stream);
    }
    (void)hipFree(dc);
    (void)hipFree(dp);
    if (e != hipSuccess)
        return fail(XH_EHIP, std::string("xh_eval_bumps: ") + hipGetErrorString(e));
    return 0;
}

extern "C" int xh_eval_bumps_noalloc(int nx, int ny, int nz, int nv, int max_bumps,
                                     const int32_t *count, const double *params, double *db) {
    int rc = ensure_init();
    if (rc)
        return rc;
    if (nx < 1 || ny < 1 || nz < 1 || nv < 1 || max_bumps < 1 || !count || !params || !db)
        return fail(XH_EINVAL, "xh_eval_bumps_noalloc: bad arguments");
    if (nv > kScratchMaxNv || max_bumps > kScratchMaxBumps)
        return fail(XH_EINVAL, "xh_eval_bumps_noalloc: nv and max_bumps must be at most 64 "
                               "(the library's static staging area)");
    for (int k = 0; k < nv; ++k)
        if (count[k] < 0 || count[k] > max_bumps)
            return fail(XH_EINVAL, "xh_eval_bumps_noalloc: count out of range");
    unsigned char *base = (unsigned char *)g_scratch[g_device];
    if (base == nullptr)
        return fail(XH_EHIP, "xh_eval_bumps_noalloc: staging area not mapped on this device");
    int *dc = (int *)base;
    double *dp = (double *)(base + kScratchCountBytes);
    // every byte the kernel reads (counts of RHS < nv, parameters of bumps < count) is
    // written here first, so nothing left by an earlier call is seen
    XH_HIP(hipMemcpyAsync(dc, count, (size_t)nv * sizeof(int), hipMemcpyHostToDevice, g_stream));
    XH_HIP(hipMemcpyAsync(dp, params, (size_t)nv * max_bumps * 5 * sizeof(double),
                          hipMemcpyHostToDevice, g_stream));
    eval_bumps(nx, ny, nz, nv, max_bumps, dc, dp, db, g_stream);
    rc = check_launch();
    if (rc)
        return rc;
    XH_HIP(hipStreamSynchronize(g_stream));
    return 0;
}
