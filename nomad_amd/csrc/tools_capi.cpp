// C ABI of the host-side tooling (include/nomad_tools.h).
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <string>

#include "convert.h"
#include "datfile.h"
#include "nomad_tools.h"
#include "synth.h"

struct nomad_matrix {
    nomad::Csc csc;
};

namespace {

thread_local std::string g_err;

int tfail(const std::string& msg) {
    g_err = msg;
    return -1;
}

nomad::SynthSpec to_spec(const nomad_synth_spec& s) {
    nomad::SynthSpec o;
    o.m = s.m;
    o.n = s.n;
    o.nnz_train = s.nnz_train;
    o.nnz_test = s.nnz_test;
    o.rank = s.rank;
    o.factor_var = s.factor_var;
    o.noise_std = s.noise_std;
    o.degree = s.degree;
    o.row_exp = s.row_exp;
    o.col_exp = s.col_exp;
    o.rating = s.rating;
    o.center = s.center;
    o.scale = s.scale;
    o.lo = s.lo;
    o.hi = s.hi;
    o.seed = s.seed;
    o.threads = s.threads;
    return o;
}

int32_t scaled(int64_t v, double s) {
    const int64_t r = int64_t(std::llround(double(v) * s));
    return int32_t(r < 1 ? 1 : r);
}

}  // namespace

extern "C" {

const char* nomad_tools_last_error(void) { return g_err.c_str(); }

int nomad_tools_read_dat(const char* path, nomad_matrix** out) {
    if (!path || !out) return tfail("null argument");
    *out = nullptr;
    nomad_matrix* a = new (std::nothrow) nomad_matrix();
    if (!a) return tfail("out of memory");
    std::string err;
    if (!nomad::read_dat(path, a->csc, err)) {
        delete a;
        return tfail(err);
    }
    *out = a;
    return 0;
}

int nomad_tools_read_dat_header(const char* path, int32_t* m, int32_t* n, int64_t* nnz) {
    if (!path || !m || !n || !nnz) return tfail("null argument");
    std::string err;
    if (!nomad::read_dat_header(path, *m, *n, *nnz, err)) return tfail(err);
    return 0;
}

int nomad_tools_write_dat(const char* path, const nomad_matrix* a, int32_t long_format) {
    if (!path || !a) return tfail("null argument");
    nomad::Csr r;
    nomad::csc_to_csr(a->csc, r);
    std::string err;
    if (!nomad::write_dat(path, r, long_format != 0, err)) return tfail(err);
    return 0;
}

int nomad_tools_matrix_from_csc(int32_t m, int32_t n, const int64_t* colptr, const int32_t* rowidx,
                                const double* val, nomad_matrix** out) {
    if (!out || !colptr || m < 0 || n < 0) return tfail("bad argument");
    const int64_t nnz = colptr[n];
    if (nnz > 0 && (!rowidx || !val)) return tfail("bad argument");
    nomad_matrix* a = new (std::nothrow) nomad_matrix();
    if (!a) return tfail("out of memory");
    a->csc.m = m;
    a->csc.n = n;
    a->csc.colptr.assign(colptr, colptr + n + 1);
    a->csc.rowidx.assign(rowidx, rowidx + nnz);
    a->csc.val.assign(val, val + nnz);
    *out = a;
    return 0;
}

int nomad_tools_matrix_info(const nomad_matrix* a, int32_t* m, int32_t* n, int64_t* nnz) {
    if (!a) return tfail("null matrix");
    if (m) *m = a->csc.m;
    if (n) *n = a->csc.n;
    if (nnz) *nnz = a->csc.nnz();
    return 0;
}

const int64_t* nomad_tools_matrix_colptr(const nomad_matrix* a) { return a ? a->csc.colptr.data() : nullptr; }
const int32_t* nomad_tools_matrix_rowidx(const nomad_matrix* a) { return a ? a->csc.rowidx.data() : nullptr; }
const double* nomad_tools_matrix_val(const nomad_matrix* a) { return a ? a->csc.val.data() : nullptr; }

void nomad_tools_matrix_free(nomad_matrix* a) { delete a; }

int nomad_tools_synth(const nomad_synth_spec* spec, nomad_matrix** train, nomad_matrix** test) {
    if (!spec || !train || !test) return tfail("null argument");
    *train = *test = nullptr;
    nomad::Csr tr, te;
    std::string err;
    if (!nomad::synth_generate(to_spec(*spec), tr, te, err)) return tfail(err);
    nomad_matrix* a = new (std::nothrow) nomad_matrix();
    nomad_matrix* b = new (std::nothrow) nomad_matrix();
    if (!a || !b) {
        delete a;
        delete b;
        return tfail("out of memory");
    }
    nomad::csr_to_csc(tr, a->csc);
    nomad::Csr().rowptr.swap(tr.rowptr);
    nomad::csr_to_csc(te, b->csc);
    *train = a;
    *test = b;
    return 0;
}

int nomad_tools_preset(int32_t config, double scale, nomad_synth_spec* o) {
    if (!o || !(scale > 0) || scale > 1) return tfail("bad argument");
    std::memset(o, 0, sizeof(*o));
    o->degree = 1;
    o->rating = 0;
    o->row_exp = 1.0;
    o->col_exp = 0.8;
    o->center = 3.0;
    o->scale = 1.0;
    o->lo = 1.0;
    o->hi = 5.0;
    switch (config) {
        case 1:  // 2,000 x 1,000, 100,000 uniform cells, 10% test, rank 16, N(0,1/4)
            o->m = scaled(2000, scale);
            o->n = scaled(1000, scale);
            o->nnz_train = scaled(90000, scale);
            o->nnz_test = scaled(10000, scale);
            o->rank = 16;
            o->factor_var = 0.25;
            o->noise_std = 0.5;
            o->degree = 0;
            o->seed = 12345;
            break;
        case 2:  // Netflix shape
            o->m = scaled(480189, scale);
            o->n = scaled(17770, scale);
            o->nnz_train = scaled(99072112, scale);
            o->nnz_test = scaled(1408395, scale);
            o->rank = 50;
            o->factor_var = 1.0 / std::sqrt(50.0);
            o->noise_std = 0.8;
            o->seed = 20250001;
            break;
        case 3:  // Yahoo! Music shape, ratings 0..100
            o->m = scaled(1000990, scale);
            o->n = scaled(624961, scale);
            o->nnz_train = scaled(252800275, scale);
            o->nnz_test = scaled(4003960, scale);
            o->rank = 64;
            o->factor_var = 1.0 / std::sqrt(64.0);
            o->center = 50.0;
            o->scale = 20.0;
            o->noise_std = 12.0 / o->scale;  // N(0, 12) in rating units (noise is added before scaling)
            o->lo = 0.0;
            o->hi = 100.0;
            o->seed = 20250002;
            break;
        case 4:  // gradient check: 100,000 x 20,000, 10M uniform cells, ratings affine to [1,5]
            o->m = scaled(100000, scale);
            o->n = scaled(20000, scale);
            o->nnz_train = scaled(10000000, scale);
            o->nnz_test = 0;
            o->rank = 32;
            o->factor_var = 1.0 / std::sqrt(32.0);
            o->noise_std = 0.0;
            o->degree = 0;
            o->rating = 1;
            o->seed = 20250003;
            break;
        default:
            return tfail("unknown preset");
    }
    // keep the request feasible at small scales
    const int64_t grid = int64_t(o->m) * o->n;
    const int64_t total = o->nnz_train + o->nnz_test;
    if (total > grid / 2) {
        const double f = double(grid / 2) / double(total);
        o->nnz_train = std::max<int64_t>(1, int64_t(double(o->nnz_train) * f));
        o->nnz_test = int64_t(double(o->nnz_test) * f);
    }
    if (o->degree == 1 && o->nnz_train + o->nnz_test < o->m) o->nnz_train = o->m - o->nnz_test;
    return 0;
}

int nomad_tools_preset_batch(int32_t p, double scale, nomad_synth_spec* o, int32_t* k, double* reg) {
    if (!o || !k || !reg || p < 0 || !(scale > 0) || scale > 1) return tfail("bad argument");
    std::memset(o, 0, sizeof(*o));
    std::mt19937_64 r(20250100ull + uint64_t(p));
    auto logu = [&](double lo, double hi) {
        std::uniform_real_distribution<double> u(std::log(lo), std::log(hi));
        return std::exp(u(r));
    };
    o->m = scaled(int64_t(logu(10000, 200000)), scale);
    o->n = scaled(int64_t(logu(2000, 50000)), scale);
    int64_t nnz = int64_t(logu(500000, 20000000) * scale);
    const int64_t cap = int64_t(o->m) * o->n / 4;
    nnz = std::max<int64_t>(std::min(nnz, cap), o->m);
    const int32_t ks[4] = {16, 32, 64, 128};
    const double regs[4] = {0.01, 0.05, 0.1, 0.5};
    *k = ks[r() % 4];
    *reg = regs[r() % 4];
    o->nnz_test = std::max<int64_t>(1, nnz / 50);
    o->nnz_train = nnz;
    o->rank = 16;
    o->factor_var = 0.25;
    o->noise_std = 0.5;
    o->degree = 1;
    o->row_exp = 1.0;
    o->col_exp = 0.8;
    o->rating = 0;
    o->center = 3.0;
    o->scale = 1.0;
    o->lo = 1.0;
    o->hi = 5.0;
    o->seed = 20250100ull + uint64_t(p);
    return 0;
}

int nomad_tools_init_factors(int32_t m, int32_t n, int32_t k, int32_t nthreads, uint64_t seed, double* W,
                             double* H) {
    if (m < 0 || n < 0 || k <= 0 || nthreads <= 0) return tfail("bad argument");
    const double hi = 1.0 / std::sqrt(double(k));
    if (W) {
        // part t: rows [t*rpp, min((t+1)*rpp, m)), rpp = ceil(m / nthreads)
        const int64_t rpp = m / nthreads + ((m % nthreads) > 0 ? 1 : 0);
        for (int32_t t = 0; t < nthreads; ++t) {
            const int64_t r0 = std::min<int64_t>(int64_t(t) * rpp, m);
            const int64_t r1 = std::min<int64_t>(int64_t(t + 1) * rpp, m);
            std::mt19937_64 rng(seed + uint64_t(t) + 1);
            std::uniform_real_distribution<double> dist(0, hi);
            for (int64_t e = r0 * k; e < r1 * k; ++e) W[e] = dist(rng);
        }
    }
    if (H) {
        std::mt19937_64 rng(seed + 139);
        std::uniform_real_distribution<double> dist(0, hi);
        for (int64_t e = 0; e < int64_t(n) * k; ++e) H[e] = dist(rng);
    }
    return 0;
}

int nomad_tools_convert_tsv(const char* train_path, const char* test_path, const char* out_dir, uint64_t seed,
                            int64_t* nusers, int64_t* nitems) {
    if (!train_path || !test_path || !out_dir) return tfail("nomad_tools_convert_tsv: null path");
    std::string err;
    int64_t nu = 0, ni = 0;
    try {
        if (!nomad::convert_tsv(train_path, test_path, out_dir, seed, nu, ni, err)) return tfail(err);
    } catch (const std::exception& e) {
        return tfail(std::string("nomad_tools_convert_tsv: ") + e.what());
    }
    if (nusers) *nusers = nu;
    if (nitems) *nitems = ni;
    return 0;
}

int nomad_tools_synthgen_conf(int32_t nrow, int32_t ncol, int32_t seed, const char* row_path, const char* col_path,
                              const char* sep) {
    if (!row_path || !col_path) return tfail("nomad_tools_synthgen_conf: null path");
    std::string err;
    if (!nomad::synthgen_conf(nrow, ncol, seed, row_path, col_path, sep ? sep : ",", err)) return tfail(err);
    return 0;
}

int nomad_tools_synthgen(const char* row_path, const char* col_path, const char* out_dir, int32_t nrow,
                         int32_t ncol, int32_t dim, int32_t seed, int64_t* train_nnz, int64_t* test_nnz) {
    if (!row_path || !col_path || !out_dir) return tfail("nomad_tools_synthgen: null path");
    std::string err;
    int64_t a = 0, b = 0;
    try {
        if (!nomad::synthgen(row_path, col_path, out_dir, nrow, ncol, dim, seed, a, b, err)) return tfail(err);
    } catch (const std::exception& e) {
        return tfail(std::string("nomad_tools_synthgen: ") + e.what());
    }
    if (train_nnz) *train_nnz = a;
    if (test_nnz) *test_nnz = b;
    return 0;
}

int nomad_tools_write_checkpoint(const char* dir, int32_t job_id, int32_t epoch, int32_t nparts, int32_t m,
                                 int32_t n, int32_t k, const double* W, const double* H) {
    if (!dir || !W || !H) return tfail("nomad_tools_write_checkpoint: null argument");
    std::string err;
    if (!nomad::write_checkpoint(dir, job_id, epoch, nparts, m, n, k, W, H, err)) return tfail(err);
    return 0;
}

int nomad_tools_read_checkpoint(const char* dir, int32_t job_id, int32_t epoch, int32_t nparts, int32_t m, int32_t n,
                                int32_t k, double* W, double* H) {
    if (!dir || !W || !H) return tfail("nomad_tools_read_checkpoint: null argument");
    std::string err;
    if (!nomad::read_checkpoint(dir, job_id, epoch, nparts, m, n, k, W, H, err)) return tfail(err);
    return 0;
}

}  // extern "C"
