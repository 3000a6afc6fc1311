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
 *   nomad_tools_convert_tsv
 *       Scripts/convert.py: TSV rating files to train.dat / test.dat
 *   nomad_tools_synthgen_conf / nomad_tools_synthgen
 *       the reference's own synthetic generator (Code/nomad/synthgen-conf.cpp,
 *       synthgen.cpp, gen.sh)
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

/* Scripts/convert.py:1-147: "user<TAB>item<TAB>rating" lines of train_path
 * and test_path -> out_dir/train.dat and test.dat (class id 1211216, rows =
 * users, duplicate (user, item) lines summed, columns ascending in a row, as
 * scipy's csr_matrix builds them) plus out_dir/user_ids.txt and item_ids.txt
 * (the ids in index order).  Ids are indexed after a shuffle with Python's
 * random generator seeded `seed` (convert.py uses 12345), i.e. random.seed
 * then random.shuffle as Python 2 performs it; the ids enter the shuffle in
 * order of first appearance (train, then test), where convert.py starts from
 * a Python 2 set's iteration order.  *nusers, *nitems (may be NULL) receive
 * the matrix size. */
int nomad_tools_convert_tsv(const char* train_path, const char* test_path, const char* out_dir, uint64_t seed,
                            int64_t* nusers, int64_t* nitems);

/* Code/nomad/synthgen-conf.cpp:74-122: the degree-frequency files of the
 * reference's synthetic generator.  row_path gets "d<sep>f" for d = 0..ncol
 * with f = int(N(20, 5)) redrawn until in [0, ncol); col_path gets "d<sep>f"
 * for d = 0..nrow with f = int of a power law (xmin 1, exponent 2.2,
 * Code/nomad/pl-dis.hpp) redrawn while above nrow; std::mt19937(seed). */
int nomad_tools_synthgen_conf(int32_t nrow, int32_t ncol, int32_t seed, const char* row_path, const char* col_path,
                              const char* sep);

/* Code/nomad/synthgen.cpp:108-311 and gen.sh:28-55: out_dir/train.dat and
 * test.dat (class id 1015) of an nrow x ncol matrix: column weights drawn
 * from the column file, H (dim x ncol) and each row's factor ~ N(0, 1), row
 * degree from the row file x 5/4 (at most ncol), 80 % of it training
 * columns and the rest test columns, each drawn by weight without repeats
 * within its split, value = row . column + 0.1 N(0, 1).  The same libstdc++
 * distributions in the reference's order, std::mt19937(seed).  *train_nnz,
 * *test_nnz (may be NULL) receive the counts. */
int nomad_tools_synthgen(const char* row_path, const char* col_path, const char* out_dir, int32_t nrow,
                         int32_t ncol, int32_t dim, int32_t seed, int64_t* train_nnz, int64_t* test_nnz);

/* Checkpoint files in the reference's layout (Code/nomad/nb_checkpoint.cpp:
 * gen_cp_file_name :37-42, archive_local :45-56, _archive_msg_queue :58-85,
 * ColumnData::serialize pool.cpp:17-23):
 *   <dir>/cp-t<job_id>-e<epoch>-p<part>.state  int32 rows, int32 k, double W[rows*k]
 *   <dir>/cp-t<job_id>-e<epoch>-p<part>.msg    packed records {int32 column,
 *                                              int64 flag = 0, double h[k]}
 *   <dir>/cp-t<job_id>-e<epoch>-p0.smsg        empty (no send queue at one rank)
 * for parts 0..nparts-1 of the rows (ceil(m / nparts) rows each,
 * nb_data.cpp:83-89); column j goes to part j mod nparts.  W (m x k) and H
 * (n x k) are row-major double.  The reader (restore, nb_checkpoint.cpp:
 * 116-143) takes whole records only and requires every column to appear. */
int nomad_tools_write_checkpoint(const char* dir, int32_t job_id, int32_t epoch, int32_t nparts, int32_t m,
                                 int32_t n, int32_t k, const double* W, const double* H);
int nomad_tools_read_checkpoint(const char* dir, int32_t job_id, int32_t epoch, int32_t nparts, int32_t m, int32_t n,
                                int32_t k, double* W, double* H);

#ifdef __cplusplus
}
#endif

#endif /* NOMAD_TOOLS_H_ */
