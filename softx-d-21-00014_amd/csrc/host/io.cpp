// Binary system files in the layout of the reference's reader (src/io/detail/reader.inl:
// 47-177): a u32 header, u32 row pointers and columns, f64 values, then the optional b and
// x multivectors.  Reads take one row block, as one MPI rank of the reference does.
#include "../../../include/xamg_io.h"

#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

namespace xamg {
int api_fail(int code, const std::string &m);
}
using xamg::api_fail;

namespace {

struct File {
    FILE *f = nullptr;
    int64_t size = 0;
    ~File() {
        if (f)
            std::fclose(f);
    }
};

int open_file(const char *path, const char *mode, File &F) {
    if (!path)
        return api_fail(-1, "xamg_io: null path");
    F.f = std::fopen(path, mode);
    if (!F.f)
        return api_fail(-1, std::string("xamg_io: cannot open ") + path);
    if (mode[0] == 'r') {
        std::fseek(F.f, 0, SEEK_END);
        F.size = (int64_t)std::ftell(F.f);
        std::fseek(F.f, 0, SEEK_SET);
    }
    return 0;
}

// count items of `sz` bytes at byte offset `off`; false when the file is too short
bool read_at(File &F, int64_t off, void *dst, size_t sz, int64_t count) {
    if (count == 0)
        return true;
    if (off < 0 || off + (int64_t)sz * count > F.size)
        return false;
    if (std::fseek(F.f, (long)off, SEEK_SET) != 0)
        return false;
    return std::fread(dst, sz, (size_t)count, F.f) == (size_t)count;
}

struct Header {
    int64_t n = 0, nnz = 0;
    int64_t off_row = 8, off_col = 0, off_val = 0, off_b = 0;
};

int header(File &F, Header &H) {
    uint32_t h[2];
    if (!read_at(F, 0, h, 4, 2))
        return api_fail(-1, "xamg_io: file shorter than its header");
    H.n = h[0];
    H.nnz = h[1];
    H.off_col = H.off_row + 4 * (H.n + 1);
    H.off_val = H.off_col + 4 * H.nnz;
    H.off_b = H.off_val + 8 * H.nnz;
    if (F.size < H.off_b)
        return api_fail(-1, "xamg_io: file shorter than its matrix section");
    return 0;
}

int block_ptrs(File &F, const Header &H, int64_t rb, int64_t nr, uint32_t &lo, uint32_t &hi) {
    if (rb < 0 || nr < 0 || rb + nr > H.n)
        return api_fail(-1, "xamg_io: row block outside the matrix");
    if (!read_at(F, H.off_row + 4 * rb, &lo, 4, 1) ||
        !read_at(F, H.off_row + 4 * (rb + nr), &hi, 4, 1))
        return api_fail(-1, "xamg_io: cannot read row pointers");
    if (hi < lo || (int64_t)hi > H.nnz)
        return api_fail(-1, "xamg_io: row pointers out of range");
    return 0;
}

} // namespace

extern "C" {

void xamg_io_block_range(int64_t nrows, int nprocs, int proc, int64_t *row_begin,
                         int64_t *block_rows) {
    const int64_t np = nprocs > 0 ? nprocs : 1;
    int64_t bs = nrows / np;
    const int64_t off = bs * proc;
    if (proc == np - 1)
        bs = nrows - off;
    if (row_begin)
        *row_begin = off;
    if (block_rows)
        *block_rows = bs;
}

int xamg_io_read_header(const char *path, int64_t *nrows, int64_t *nnz) {
    File F;
    Header H;
    int rc;
    if ((rc = open_file(path, "rb", F)) || (rc = header(F, H)))
        return rc;
    if (nrows)
        *nrows = H.n;
    if (nnz)
        *nnz = H.nnz;
    return 0;
}

int xamg_io_sections(const char *path, int nv, int32_t *has_b, int32_t *has_x) {
    File F;
    Header H;
    int rc;
    if ((rc = open_file(path, "rb", F)) || (rc = header(F, H)))
        return rc;
    if (nv < 1)
        return api_fail(-1, "xamg_io_sections: nv must be >= 1");
    const int64_t vb = 8 * H.n * nv;
    const bool b = F.size >= H.off_b + vb;
    if (has_b)
        *has_b = b ? 1 : 0;
    if (has_x)
        *has_x = (b && F.size >= H.off_b + 2 * vb) ? 1 : 0;
    return 0;
}

int xamg_io_block_nnz(const char *path, int64_t row_begin, int64_t block_rows,
                      int64_t *block_nnz) {
    File F;
    Header H;
    int rc;
    if ((rc = open_file(path, "rb", F)) || (rc = header(F, H)))
        return rc;
    uint32_t lo, hi;
    if ((rc = block_ptrs(F, H, row_begin, block_rows, lo, hi)))
        return rc;
    if (block_nnz)
        *block_nnz = (int64_t)hi - (int64_t)lo;
    return 0;
}

int xamg_io_read_system(const char *path, int nv, int64_t row_begin, int64_t block_rows,
                        uint32_t *row, uint32_t *col, double *val, double *b, double *x,
                        int32_t *found) {
    File F;
    Header H;
    int rc;
    if ((rc = open_file(path, "rb", F)) || (rc = header(F, H)))
        return rc;
    if ((b || x) && nv < 1)
        return api_fail(-1, "xamg_io_read_system: nv must be >= 1");
    if (!row || (!col && block_rows > 0) || (!val && block_rows > 0))
        return api_fail(-1, "xamg_io_read_system: null output");
    uint32_t lo, hi;
    if ((rc = block_ptrs(F, H, row_begin, block_rows, lo, hi)))
        return rc;
    // read_matrix (reader.inl:97-123): the block's row pointers, then its slice of the
    // column and value arrays, row pointers rebased to the slice
    if (!read_at(F, H.off_row + 4 * row_begin, row, 4, block_rows + 1))
        return api_fail(-1, "xamg_io_read_system: cannot read row pointers");
    const int64_t bnnz = (int64_t)hi - (int64_t)lo;
    for (int64_t i = 0; i < block_rows; ++i)
        if (row[i + 1] < row[i])
            return api_fail(-1, "xamg_io_read_system: row pointers decrease");
    for (int64_t i = 0; i <= block_rows; ++i)
        row[i] -= lo;
    if (!read_at(F, H.off_col + 4 * (int64_t)lo, col, 4, bnnz) ||
        !read_at(F, H.off_val + 8 * (int64_t)lo, val, 8, bnnz))
        return api_fail(-1, "xamg_io_read_system: cannot read columns / values");
    for (int64_t j = 0; j < bnnz; ++j)
        if ((int64_t)col[j] >= H.n)
            return api_fail(-1, "xamg_io_read_system: column index out of range");
    // read_system (reader.inl:148-168): b, else b = 1; x, else x = 0
    int32_t fl = 0;
    const int64_t cnt = block_rows * (int64_t)(nv > 0 ? nv : 0);
    const int64_t vb = 8 * H.n * (int64_t)(nv > 0 ? nv : 0);
    const int64_t ob = H.off_b + 8 * row_begin * (int64_t)(nv > 0 ? nv : 0);
    const bool have_b = nv > 0 && F.size >= H.off_b + vb;
    const bool have_x = have_b && F.size >= H.off_b + 2 * vb;
    if (b) {
        if (have_b && read_at(F, ob, b, 8, cnt))
            fl |= 1;
        else
            for (int64_t i = 0; i < cnt; ++i)
                b[i] = 1.0;
    }
    if (x) {
        if (have_x && read_at(F, ob + vb, x, 8, cnt))
            fl |= 2;
        else
            std::memset(x, 0, sizeof(double) * (size_t)cnt);
    }
    if (found)
        *found = fl;
    return 0;
}

int xamg_io_write_system(const char *path, int64_t nrows, int64_t nnz, const uint32_t *row,
                         const uint32_t *col, const double *val, int nv, const double *b,
                         const double *x) {
    if (nrows < 0 || nrows > 0xFFFFFFFFll || nnz < 0 || nnz > 0xFFFFFFFFll || !row ||
        (nnz > 0 && (!col || !val)) || ((b || x) && nv < 1) || (x && !b))
        return api_fail(-1, "xamg_io_write_system: bad arguments");
    if ((int64_t)row[nrows] != nnz)
        return api_fail(-1, "xamg_io_write_system: row[nrows] != nnz");
    File F;
    int rc;
    if ((rc = open_file(path, "wb", F)))
        return rc;
    const uint32_t h[2] = {(uint32_t)nrows, (uint32_t)nnz};
    bool ok = std::fwrite(h, 4, 2, F.f) == 2;
    ok = ok && std::fwrite(row, 4, (size_t)nrows + 1, F.f) == (size_t)nrows + 1;
    if (nnz > 0) {
        ok = ok && std::fwrite(col, 4, (size_t)nnz, F.f) == (size_t)nnz;
        ok = ok && std::fwrite(val, 8, (size_t)nnz, F.f) == (size_t)nnz;
    }
    const size_t cnt = (size_t)nrows * (size_t)(nv > 0 ? nv : 0);
    if (b && cnt)
        ok = ok && std::fwrite(b, 8, cnt, F.f) == cnt;
    if (x && cnt)
        ok = ok && std::fwrite(x, 8, cnt, F.f) == cnt;
    if (!ok)
        return api_fail(-1, std::string("xamg_io_write_system: write failed: ") + path);
    return 0;
}

} // extern "C"
