/*
 * nomad_tools.h -- host-side data tooling (libnomad_tools.so): the
 * train.dat/test.dat format, synthetic rating matrices and the reference's
 * factor initialisation.  No GPU code.
 *
 *   nomad_tools_read_dat / nomad_tools_write_dat
 *       the PETSc-style binary format read by Code/nomad/nb_data.cpp:19-176 and
 *       written by Scripts/convert.py:101-104 (class ids 1211216 and 1015)
 *   nomad_tools_init_factors
 *       W:  part t of T rows ~ U(0, 1/sqrt(k)) from mt19937_64(seed + t + 1)
 *           (Code/nomad/nb_main.cpp:444-456, one rank)
 *       H:  columns 0..n-1, k draws each, from mt19937_64(seed + 139)
 *           (Code/nomad/nb_main.cpp:70-93, one rank)
 *   nomad_tools_synth / nomad_tools_preset
 *       synthetic matrices with a planted low-rank truth; presets for the
 *       benchmark shapes (small uniform, Netflix, Yahoo! Music, gradient check,
 *       batched-sweep problems)
 *
 * Return codes: 0 on success, negative on error (message via
 * nomad_tools_last_error()).
 */
#ifndef NOMAD_TOOLS_H_
#define NOMAD_TOOLS_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nomad_matrix nomad_matrix; /* CSC, rows ascending in a column */

typedef struct nomad_synth_spec {
    int32_t m, n;
    int64_t nnz_train, nnz_test;
    int32_t rank;
    int32_t degree; /* 0 uniform cells, 1 zipf degrees                       */
    int32_t rating; /* 0 clip(round(center + scale*u.v + eps)), 1 affine      */
    int32_t threads;
    double factor_var, noise_std;
    double row_exp, col_exp;
    double center, scale, lo, hi;
    uint64_t seed;
} nomad_synth_spec;

const char* nomad_tools_last_error(void);

int nomad_tools_read_dat(const char* path, nomad_matrix** out);
int nomad_tools_read_dat_header(const char* path, int32_t* m, int32_t* n, int64_t* nnz);
int nomad_tools_write_dat(const char* path, const nomad_matrix* a, int32_t long_format);

int nomad_tools_matrix_from_csc(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowidx,
                                const double* val, nomad_matrix** out);
int nomad_tools_matrix_info(const nomad_matrix* a, int32_t* m, int32_t* n, int64_t* nnz);
const int64_t* nomad_tools_matrix_colptr(const nomad_matrix* a);
const int32_t* nomad_tools_matrix_rowidx(const nomad_matrix* a);
const double* nomad_tools_matrix_val(const nomad_matrix* a);
void nomad_tools_matrix_free(nomad_matrix* a);

int nomad_tools_synth(const nomad_synth_spec* spec, nomad_matrix** train, nomad_matrix** test);

/* config: 1 small uniform (2,000 x 1,000), 2 Netflix shape, 3 Yahoo! Music
 * shape, 4 gradient check (100,000 x 20,000).  scale in (0, 1] shrinks rows,
 * columns and cell counts proportionally (1 = full size). */
int nomad_tools_preset(int32_t config, double scale, nomad_synth_spec* out);

/* Problem p of the batched sweep: sizes log-uniform, k and reg drawn from
 * {16,32,64,128} and {0.01,0.05,0.1,0.5}. */
int nomad_tools_preset_batch(int32_t p, double scale, nomad_synth_spec* out, int32_t* k, double* reg);

/* Reference initialisation, see above.  W is m x k and H n x k, row-major;
 * nthreads is the reference's --nthreads (number of row parts). */
int nomad_tools_init_factors(int32_t m, int32_t n, int32_t k, int32_t nthreads, uint64_t seed,
                             double* W, double* H);

#ifdef __cplusplus
}
#endif

#endif /* NOMAD_TOOLS_H_ */
