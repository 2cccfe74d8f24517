/*
 * Binary system files (part of libxamg.so): the format of the reference's reader
 * XAMG::io::read_system / read_matrix (src/io/detail/reader.inl:47-177), used by its test
 * harness for `-matrix <file>` (examples/common/system/system.h:33-41):
 *
 *   u32 nrows, u32 nnz, u32 row[nrows+1], u32 col[nnz], f64 val[nnz],
 *   f64 b[nrows*nv], f64 x[nrows*nv]          (b and x optional, nv-interleaved)
 *
 * A reader takes one row block (the reference reads one block per MPI rank, split as in
 * xamg_io_block_range): row pointers rebased to the block's first non-zero, all columns
 * kept (global indices), the b and x rows of the block.  As in the reference, a file
 * without b gives b = 1 and one without x gives x = 0.  Deviation: the reference seeks to
 * the x block with an absolute offset (reader.inl:158-159, SEEK_SET with a length meant
 * relative), which reads x from the start of the file; x is read here from its section
 * right after b, the position that code computes relative to the end of the b block.
 * All functions return 0 on success, < 0 on failure (XAMG_get_last_error()).
 */
#ifndef XAMG_IO_H_
#define XAMG_IO_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* the reference's row split for rank `proc` of `nprocs` (reader.inl:57-60): nrows/nprocs
 * rows each, the last rank takes the remainder */
void xamg_io_block_range(int64_t nrows, int nprocs, int proc, int64_t *row_begin,
                         int64_t *block_rows);

/* nrows and nnz of a file, and the number of RHS columns its b / x sections hold for a
 * given nv (has_b / has_x: 1 when the section is complete) */
int xamg_io_read_header(const char *path, int64_t *nrows, int64_t *nnz);
int xamg_io_sections(const char *path, int nv, int32_t *has_b, int32_t *has_x);

/* non-zeros of rows [row_begin, row_begin + block_rows) */
int xamg_io_block_nnz(const char *path, int64_t row_begin, int64_t block_rows,
                      int64_t *block_nnz);

/* rows [row_begin, row_begin + block_rows): row[block_rows+1] (rebased), col / val
 * [block_nnz], b / x [block_rows*nv] (either may be NULL); found (nullable): bit 0 set when
 * b was read from the file, bit 1 when x was.  Row pointers must be non-decreasing and
 * columns < nrows, else the file is rejected. */
int xamg_io_read_system(const char *path, int nv, int64_t row_begin, int64_t block_rows,
                        uint32_t *row, uint32_t *col, double *val, double *b, double *x,
                        int32_t *found);

/* writes a whole system; b and x may be NULL (x written only when b is given) */
int xamg_io_write_system(const char *path, int64_t nrows, int64_t nnz, const uint32_t *row,
                         const uint32_t *col, const double *val, int nv, const double *b,
                         const double *x);

#ifdef __cplusplus
}
#endif

#endif
