/*
 * XAMG C API, drop-in for the reference api/c/xamg_c_api.h:43-117 (same names, same
 * argument meaning).  Implemented in libxamg.so over the MI355X solve phase of
 * libxamg_hip.so.  Differences from the reference, all documented per function:
 *   - one process, one GPU: XAMG_id_get_proc() is 0, XAMG_id_get_nprocs() is 1;
 *   - NV is chosen at run time (env XAMG_NV, default 2) instead of -DXAMG_NV;
 *   - calls return 0 on success and a negative code on failure instead of asserting
 *     (XAMG_get_last_error() gives the message).
 * Extensions (no reference counterpart) are marked "extension".
 */
#ifndef XAMG_C_API_H_
#define XAMG_C_API_H_

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef size_t XAMG_param_id;
typedef size_t XAMG_param_override_id;
typedef size_t XAMG_part_id;
typedef size_t XAMG_vector_id;
typedef size_t XAMG_csr_matrix_id;
typedef size_t XAMG_matrix_id;
typedef size_t XAMG_solver_id;

typedef enum { XAMG_MEM_LOCAL, XAMG_MEM_SHARED, XAMG_MEM_DISTRIBUTED } vector_alloc_mode;

int XAMG_init(int argc, char **argv, char *conf);
int XAMG_finalize();

int XAMG_get_nv();

int XAMG_id_get_proc();
int XAMG_id_get_nprocs();
void *XAMG_id_get_comm_ptr();

int XAMG_param_create(XAMG_param_id *);
int XAMG_param_add_value_s(char *solver_type, char *key, char *value, XAMG_param_id pid);
int XAMG_param_add_value_i(char *solver_type, char *key, uint16_t value, XAMG_param_id pid);
int XAMG_param_add_value_d(char *solver_type, char *key, float value, XAMG_param_id pid);

int XAMG_param_forced_change_value_s(char *solver_type, char *key, char *value, XAMG_param_id pid);
int XAMG_param_forced_change_value_i(char *solver_type, char *key, uint16_t value,
                                     XAMG_param_id pid);
int XAMG_param_forced_change_value_d(char *solver_type, char *key, float value, XAMG_param_id pid);

int XAMG_param_change_value_s(char *solver_type, char *key, char *value, XAMG_param_id pid);
int XAMG_param_change_value_i(char *solver_type, char *key, uint16_t value, XAMG_param_id pid);
int XAMG_param_change_value_d(char *solver_type, char *key, float value, XAMG_param_id pid);

int XAMG_param_override_create(XAMG_param_override_id *);
int XAMG_param_override_add_value(char *key, char *val, XAMG_param_override_id ovid);
int XAMG_param_override_apply(char *solver_type, XAMG_param_id pid, XAMG_param_override_id ovid);
int XAMG_param_override_destroy(XAMG_param_override_id);

void XAMG_param_set_default(XAMG_param_id);
int XAMG_param_print(XAMG_param_id);
int XAMG_param_destroy(XAMG_param_id);

int XAMG_part_create(XAMG_part_id *);
int XAMG_part_construct(int, XAMG_part_id);
int XAMG_part_get_numa_size(XAMG_part_id);
int XAMG_part_destroy(XAMG_part_id);

int XAMG_vector_create(XAMG_vector_id *, vector_alloc_mode);
int XAMG_vector_offset(uint64_t, XAMG_vector_id);
int XAMG_vector_alloc_d(int, uint16_t, XAMG_vector_id);
int XAMG_vector_set_part(XAMG_part_id, XAMG_vector_id);
int XAMG_vector_set_val_d(double, XAMG_vector_id);
int XAMG_vector_download_d(double *, uint64_t, uint64_t, uint16_t, XAMG_vector_id);
int XAMG_vector_upload_v(XAMG_vector_id, XAMG_vector_id);
int XAMG_vector_upload_d(double *, uint64_t, uint64_t, uint16_t, XAMG_vector_id);
int XAMG_vector_destroy(XAMG_vector_id);

int XAMG_csr_matrix_create(XAMG_csr_matrix_id *);
int XAMG_csr_matrix_alloc(uint64_t, uint64_t, XAMG_csr_matrix_id);
int XAMG_csr_matrix_fill(int *, int *, double *, XAMG_matrix_id);
int XAMG_csr_matrix_offset(int, int, XAMG_matrix_id);
uint32_t XAMG_csr_matrix_get_nrows(XAMG_csr_matrix_id);
int XAMG_csr_matrix_destroy(XAMG_csr_matrix_id);

int XAMG_matrix_create(XAMG_matrix_id *);
int XAMG_matrix_set_part(XAMG_part_id, XAMG_matrix_id);
int XAMG_matrix_construct(XAMG_csr_matrix_id mcsrid, XAMG_matrix_id mid);
int XAMG_matrix_destroy(XAMG_matrix_id);

int XAMG_solver_create1(XAMG_matrix_id, XAMG_param_id, XAMG_solver_id *);
int XAMG_solver_create2(XAMG_matrix_id, XAMG_param_id, XAMG_vector_id, XAMG_vector_id,
                        XAMG_solver_id *);
int XAMG_solver_renew_params(XAMG_param_id, XAMG_solver_id);
int XAMG_solver_solve1(uint16_t, XAMG_solver_id);
int XAMG_solver_solve2(uint16_t, XAMG_vector_id, XAMG_vector_id, XAMG_solver_id);
int XAMG_solver_get_convergence_info(int *, uint16_t, XAMG_solver_id);
int XAMG_solver_destroy(XAMG_solver_id);

double XAMG_timer();
int XAMG_barrier();

/* extension: solver_stats (src/solvers/solver.h:48-66) of the last solve */
int XAMG_solver_get_stats(XAMG_solver_id, int *iters, double *abs_res, double *rel_res);
/* extension: message of the last failing call */
const char *XAMG_get_last_error();

#ifdef __cplusplus
}
#endif

#endif /* XAMG_C_API_H_ */
