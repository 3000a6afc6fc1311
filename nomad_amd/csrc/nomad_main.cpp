// nomad_float / nomad_double: the reference's one-rank driver on one MI355X.
//
// Same command line as the reference (Code/nomad/nomad_option.cpp:16-77), the
// same inputs (<input>/train.dat and <input>/test.dat, nb_data.cpp:182-220),
// the same timeout schedule (nb_main.cpp:109-134: train until the next
// cumulative timeout, stop, evaluate, optionally pause 3 s), the same
// "Statistics" block (nb_main.cpp:264-300) and the same --output dump: rows
// then columns, "row,<i>,v..." / "column,<j>,v...", into <output>0
// (nb_main.cpp:683-705, 350-366).  Training runs through the C ABI of
// libnomad_hip; the SGD threads of the reference are the GPU's waves.
//
// Checkpoints (--cptype other than none): after every timeout segment, with
// every column at rest, the factors go to the reference's checkpoint files
// (<cppath>/cp-t<jobid>-e<segment>-p<part>.state / .msg / .smsg,
// nb_checkpoint.cpp:37-143, --nthreads parts) plus cp-t<jobid>-e<segment>.nomad
// with the step-size state the reference does not save (the epoch count and,
// in async mode, the per-(column, workgroup) visit counts); --recover <e>
// starts from checkpoint e instead of the initial factors.
//
// Options the reference accepts for its multi-machine machinery (cpm, tpm,
// termination check, network control) are accepted and have no effect at one
// rank on one GPU.  GPU-specific options: --workers (row blocks, 0:
// automatic), --device, --storage, --rounding, --mode, --balance.
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include "nomad_hip.h"
#include "nomad_tools.h"

#ifndef NOMAD_DEFAULT_STORAGE
#define NOMAD_DEFAULT_STORAGE NOMAD_STORE_FP32
#endif

namespace {

struct Options {
    int nthreads = 4;
    double lrate = 0.001, drate = 0.1, reg = 1.0;
    unsigned long long seed = 12345;
    std::vector<double> timeouts{10.0};
    int dim = 100;
    bool pause = true;
    std::string input, output;
    int ptoken = 0;
    int verbose = 0;
    // GPU
    int workers = 0, device = 0;
    int storage = NOMAD_DEFAULT_STORAGE;
    int rounding = NOMAD_ROUND_NEAREST;
    int mode = NOMAD_MODE_ASYNC;
    bool balance = true;
    // checkpoints
    std::string cp_type = "none", cp_path;
    int job_id = 1, recover = -1;
};

struct Spec {
    const char* longname;
    char shortname;  // 0: none
    const char* help;
    enum Kind { kFlag, kInt, kDouble, kU64, kString, kBool, kDoubles, kIgnoredValue } kind;
    void* target;
};

bool parse_bool(const std::string& s, bool* out) {
    static const char* t[] = {"1", "true", "yes", "on"};
    static const char* f[] = {"0", "false", "no", "off"};
    for (const char* v : t)
        if (s == v) return *out = true, true;
    for (const char* v : f)
        if (s == v) return *out = false, true;
    return false;
}

bool parse_double(const std::string& s, double* out) {
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(s.c_str(), &end);
    if (errno || end == s.c_str() || *end) return false;
    *out = v;
    return true;
}

bool parse_int(const std::string& s, long long* out) {
    char* end = nullptr;
    errno = 0;
    const long long v = std::strtoll(s.c_str(), &end, 10);
    if (errno || end == s.c_str() || *end) return false;
    *out = v;
    return true;
}

bool looks_like_option(const std::string& s) {
    if (s.size() < 2 || s[0] != '-') return false;
    double d;
    return !parse_double(s, &d);  // "-1.5" is a value
}

void usage(const std::vector<Spec>& specs) {
    std::cerr << "nomad options:\n";
    for (const Spec& s : specs) {
        std::string name = std::string("  --") + s.longname;
        if (s.shortname) name += std::string(" [ -") + s.shortname + " ]";
        std::cerr << std::left << std::setw(34) << name << " " << s.help << "\n";
    }
}

bool parse_args(int argc, char** argv, Options& o) {
    std::string storage = o.storage == NOMAD_STORE_FP64 ? "fp64" : "fp32", rounding = "nearest", mode = "async";
    double ignored_d = 0;
    std::string ignored_s;
    bool help = false;
    std::vector<Spec> specs = {
        {"help", 'h', "produce help message", Spec::kFlag, &help},
        {"nthreads", 'n', "number of threads of the reference's W initialisation streams (default 4)",
         Spec::kInt, &o.nthreads},
        {"lrate", 'l', "learning rate (0.001)", Spec::kDouble, &o.lrate},
        {"drate", 'd', "decay rate (0.1)", Spec::kDouble, &o.drate},
        {"reg", 'r', "regularization parameter lambda (1.0)", Spec::kDouble, &o.reg},
        {"seed", 's', "seed value of random number generator (12345)", Spec::kU64, &o.seed},
        {"cpm", 0, "accepted, no effect at one rank", Spec::kIgnoredValue, &ignored_d},
        {"tpm", 0, "accepted, no effect at one rank", Spec::kIgnoredValue, &ignored_d},
        {"timeout", 't', "cumulative timeout seconds, one or more (10.0)", Spec::kDoubles, &o.timeouts},
        {"tm_value", 'u', "accepted, no effect", Spec::kIgnoredValue, &ignored_d},
        {"tm_error", 'e', "accepted, no effect", Spec::kIgnoredValue, &ignored_d},
        {"tm_interval", 0, "accepted, no effect", Spec::kIgnoredValue, &ignored_d},
        {"tm_portion", 0, "accepted, no effect", Spec::kIgnoredValue, &ignored_d},
        {"dim", 0, "dimension of latent space (100, at most 256)", Spec::kInt, &o.dim},
        {"reuse", 0, "accepted, no effect", Spec::kIgnoredValue, &ignored_d},
        {"pause", 0, "pause 3 s after each timeout (true)", Spec::kBool, &o.pause},
        {"random_send", 0, "accepted, no effect at one rank", Spec::kIgnoredValue, &ignored_d},
        {"r0delay", 0, "accepted, no effect at one rank", Spec::kIgnoredValue, &ignored_d},
        {"input", 'i', "path of data (directory with train.dat and test.dat)", Spec::kString, &o.input},
        {"path", 0, "the same as --input (README spelling)", Spec::kString, &o.input},
        {"ptoken", 0,
         "tokens in the pipeline: at most this many columns resident per GPU workgroup (0: the LDS slots, 128 "
         "for fp32/bf16 at 8 waves per workgroup, 64 at 4 or 16, 32 for fp64)",
         Spec::kInt, &o.ptoken},
        {"output", 'o', "path prefix of the factor dump (<output>0)", Spec::kString, &o.output},
        {"recover", 0, "recover from the checkpoint of this segment (-1: do not recover)", Spec::kInt, &o.recover},
        {"jobid", 'j', "job id in the checkpoint file names (1)", Spec::kInt, &o.job_id},
        {"cptype", 'c', "checkpoint type: none, or sync / async / vs (a snapshot after every timeout segment)",
         Spec::kString, &o.cp_type},
        {"cpinterval", 'k', "accepted (checkpoints are taken at the timeouts)", Spec::kIgnoredValue, &ignored_d},
        {"cppath", 'p', "directory of the checkpoint files", Spec::kString, &o.cp_path},
        {"net_delay", 0, "accepted, no effect", Spec::kIgnoredValue, &ignored_d},
        {"net_ratio", 0, "accepted, no effect", Spec::kIgnoredValue, &ignored_s},
        {"verbose", 'v', "verbose level", Spec::kInt, &o.verbose},
        {"workers", 0, "GPU row blocks / waves (0: automatic)", Spec::kInt, &o.workers},
        {"device", 0, "HIP device ordinal (0)", Spec::kInt, &o.device},
        {"storage", 0, "factor storage: fp32, bf16 or fp64", Spec::kString, &storage},
        {"rounding", 0, "bf16 rounding: nearest or stochastic", Spec::kString, &rounding},
        {"mode", 0, "async (token passing) or strat (deterministic)", Spec::kString, &mode},
        {"balance", 0, "row blocks balanced by ratings (true)", Spec::kBool, &o.balance},
    };
    bool bad = false;
    for (int i = 1; i < argc && !bad; ++i) {
        std::string a = argv[i];
        const Spec* spec = nullptr;
        std::string inline_val;
        bool has_inline = false;
        if (a.rfind("--", 0) == 0) {
            std::string name = a.substr(2);
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                inline_val = name.substr(eq + 1);
                name = name.substr(0, eq);
                has_inline = true;
            }
            for (const Spec& s : specs)
                if (name == s.longname) spec = &s;
        } else if (looks_like_option(a)) {
            for (const Spec& s : specs)
                if (s.shortname && a[1] == s.shortname) spec = &s;
            if (spec && a.size() > 2) {
                inline_val = a.substr(2);
                has_inline = true;
            }
        }
        if (!spec) {
            std::cerr << "error: unrecognised option '" << a << "'\n";
            bad = true;
            break;
        }
        if (spec->kind == Spec::kFlag) {
            *static_cast<bool*>(spec->target) = true;
            continue;
        }
        std::vector<std::string> vals;
        if (has_inline) vals.push_back(inline_val);
        if (spec->kind == Spec::kDoubles) {
            while (i + 1 < argc && !looks_like_option(argv[i + 1])) vals.push_back(argv[++i]);
        } else if (!has_inline) {
            if (i + 1 >= argc) {
                std::cerr << "error: the required argument for option '--" << spec->longname << "' is missing\n";
                bad = true;
                break;
            }
            vals.push_back(argv[++i]);
        }
        if (vals.empty()) {
            std::cerr << "error: the required argument for option '--" << spec->longname << "' is missing\n";
            bad = true;
            break;
        }
        const std::string& v = vals[0];
        long long iv;
        double dv;
        bool bv;
        switch (spec->kind) {
            case Spec::kInt:
                if (!parse_int(v, &iv)) bad = true;
                else *static_cast<int*>(spec->target) = int(iv);
                break;
            case Spec::kU64:
                if (!parse_int(v, &iv) || iv < 0) bad = true;
                else *static_cast<unsigned long long*>(spec->target) = (unsigned long long)iv;
                break;
            case Spec::kDouble:
                if (!parse_double(v, &dv)) bad = true;
                else *static_cast<double*>(spec->target) = dv;
                break;
            case Spec::kBool:
                if (!parse_bool(v, &bv)) bad = true;
                else *static_cast<bool*>(spec->target) = bv;
                break;
            case Spec::kString: *static_cast<std::string*>(spec->target) = v; break;
            case Spec::kDoubles: {
                auto* t = static_cast<std::vector<double>*>(spec->target);
                t->clear();
                for (const std::string& s : vals) {
                    if (!parse_double(s, &dv)) bad = true;
                    t->push_back(dv);
                }
                break;
            }
            default: break;
        }
        if (bad) std::cerr << "error: the argument ('" << v << "') for option '--" << spec->longname << "' is invalid\n";
    }
    if (!bad && !help) {
        if (o.input.empty()) {
            std::cerr << "input path has to be specified." << std::endl;
            bad = true;
        }
        if (storage == "fp32") o.storage = NOMAD_STORE_FP32;
        else if (storage == "bf16") o.storage = NOMAD_STORE_BF16;
        else if (storage == "fp64") o.storage = NOMAD_STORE_FP64;
        else bad = true;
        if (rounding == "nearest") o.rounding = NOMAD_ROUND_NEAREST;
        else if (rounding == "stochastic") o.rounding = NOMAD_ROUND_STOCHASTIC;
        else bad = true;
        if (mode == "async") o.mode = NOMAD_MODE_ASYNC;
        else if (mode == "strat") o.mode = NOMAD_MODE_STRAT;
        else bad = true;
        if (o.timeouts.empty()) bad = true;
        for (size_t t = 0; t < o.timeouts.size(); ++t)
            if (!(o.timeouts[t] > (t ? o.timeouts[t - 1] : 0.0))) {
                std::cerr << "timeouts must be positive and increasing." << std::endl;
                bad = true;
                break;
            }
        if (o.dim < 1 || o.dim > 256 || o.nthreads < 0 || o.ptoken < 0) bad = true;
        if (o.cp_type != "none" && o.cp_type != "sync" && o.cp_type != "async" && o.cp_type != "vs") {
            std::cerr << "checkpoint type must be none, sync, async or vs." << std::endl;
            bad = true;
        }
        if (o.cp_type != "none" && o.cp_path.empty()) {
            std::cerr << "checkpoint path has to be specified." << std::endl;
            bad = true;
        }
    }
    if (bad || help) {
        usage(specs);
        return false;
    }
    return true;
}

struct Matrix {
    nomad_matrix* p = nullptr;
    ~Matrix() {
        if (p) nomad_tools_matrix_free(p);
    }
};

int fail(const char* what, const char* msg) {
    std::cerr << "error: " << what << ": " << msg << std::endl;
    return 1;
}

// cp-t<job>-e<epoch>.nomad: "NOMADCP1", int32 mode, storage, m, n, k, 0,
// int64 epochs, int64 V, int32 visit counts[V] (V = 0 in strat mode)
std::string state_name(const Options& o, int epoch) {
    return o.cp_path + "/cp-t" + std::to_string(o.job_id) + "-e" + std::to_string(epoch) + ".nomad";
}

int save_checkpoint(const Options& o, nomad_ctx* ctx, int epoch, int32_t m, int32_t n, int k, std::vector<double>& W,
                    std::vector<double>& H) {
    if (nomad_get_factors(ctx, W.data(), H.data())) return fail("checkpoint", nomad_last_error());
    const int parts = o.nthreads > 0 ? o.nthreads : 1;
    if (nomad_tools_write_checkpoint(o.cp_path.c_str(), o.job_id, epoch, parts, m, n, k, W.data(), H.data()))
        return fail("checkpoint", nomad_tools_last_error());
    int64_t epochs = 0, V = 0;
    std::vector<int32_t> counts;
    if (nomad_get_progress(ctx, &epochs)) return fail("checkpoint", nomad_last_error());
    if (o.mode == NOMAD_MODE_ASYNC) {
        if (nomad_get_route_state(ctx, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &V))
            return fail("checkpoint", nomad_last_error());
        counts.resize(size_t(V));
        if (nomad_get_route_state(ctx, nullptr, nullptr, nullptr, counts.data(), nullptr, nullptr, nullptr, nullptr))
            return fail("checkpoint", nomad_last_error());
    }
    std::ofstream f(state_name(o, epoch), std::ios::binary);
    const int32_t head[6] = {o.mode, o.storage, m, n, k, 0};
    f.write("NOMADCP1", 8);
    f.write(reinterpret_cast<const char*>(head), sizeof(head));
    f.write(reinterpret_cast<const char*>(&epochs), 8);
    f.write(reinterpret_cast<const char*>(&V), 8);
    f.write(reinterpret_cast<const char*>(counts.data()), std::streamsize(4 * counts.size()));
    if (!f) return fail("checkpoint", "cannot write the .nomad state file");
    std::cout << "checkpoint " << epoch << " written to " << o.cp_path << std::endl;
    return 0;
}

// the step-size state of checkpoint `epoch`, if its .nomad file exists
int restore_state(const Options& o, nomad_ctx* ctx, int32_t m, int32_t n, int k) {
    std::ifstream f(state_name(o, o.recover), std::ios::binary);
    if (!f) {
        std::cout << "no " << state_name(o, o.recover) << ": step sizes restart from the first visit" << std::endl;
        return 0;
    }
    char magic[8];
    int32_t head[6];
    int64_t epochs = 0, V = 0;
    f.read(magic, 8);
    f.read(reinterpret_cast<char*>(head), sizeof(head));
    f.read(reinterpret_cast<char*>(&epochs), 8);
    f.read(reinterpret_cast<char*>(&V), 8);
    if (!f || std::memcmp(magic, "NOMADCP1", 8) != 0 || head[0] != o.mode || head[2] != m || head[3] != n ||
        head[4] != k)
        return fail("recover", "the .nomad state file does not match this run (mode or sizes)");
    if (nomad_set_progress(ctx, epochs)) return fail("recover", nomad_last_error());
    if (o.mode == NOMAD_MODE_ASYNC && V > 0) {
        std::vector<int32_t> counts(static_cast<size_t>(V));
        f.read(reinterpret_cast<char*>(counts.data()), std::streamsize(4 * V));
        if (!f) return fail("recover", "truncated .nomad state file");
        if (nomad_set_visit_counts(ctx, counts.data(), V))
            return fail("recover", "the visit counts do not fit this run's workers (same --workers and --balance)");
    }
    std::cout << "recovered checkpoint " << o.recover << " at epoch " << epochs << std::endl;
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (!parse_args(argc, argv, o)) return 1;

    Matrix tr, te;
    const std::string trp = o.input + "/train.dat", tep = o.input + "/test.dat";
    if (nomad_tools_read_dat(trp.c_str(), &tr.p)) return fail(trp.c_str(), nomad_tools_last_error());
    if (nomad_tools_read_dat(tep.c_str(), &te.p)) return fail(tep.c_str(), nomad_tools_last_error());
    int32_t m = 0, n = 0, tm = 0, tn = 0;
    int64_t nnz = 0, tnnz = 0;
    nomad_tools_matrix_info(tr.p, &m, &n, &nnz);
    nomad_tools_matrix_info(te.p, &tm, &tn, &tnnz);
    if (tm != m || tn != n) return fail("test.dat", "dimensions differ from train.dat");

    for (double t : o.timeouts) std::cout << "timeout: " << t << "\n";
    std::cout << "nrows: " << m << " ncols: " << n << " train nnz: " << nnz << " test nnz: " << tnnz << std::endl;

    const int k = o.dim;
    std::vector<double> W(size_t(m) * k), H(size_t(n) * k);
    const int init_threads = o.nthreads > 0 ? o.nthreads : 1;
    if (nomad_tools_init_factors(m, n, k, init_threads, o.seed, W.data(), H.data()))
        return fail("init", nomad_tools_last_error());

    nomad_params p{};
    p.m = m;
    p.n = n;
    p.k = k;
    p.storage = o.storage;
    p.mode = o.mode;
    p.rounding = o.rounding;
    p.nworkers = o.workers;
    p.ptoken = o.ptoken;
    p.device = o.device;
    p.flags = (o.balance && o.mode == NOMAD_MODE_ASYNC) ? NOMAD_FLAG_BALANCE : 0;
    p.lrate = o.lrate;
    p.drate = o.drate;
    p.reg = o.reg;
    p.seed = o.seed;
    nomad_ctx* ctx = nullptr;
    if (nomad_create(&p, &ctx)) return fail("nomad_create", nomad_last_error());
    auto done = [&](int rc) {
        nomad_destroy(ctx);
        return rc;
    };
    if (nomad_upload_train(ctx, nomad_tools_matrix_colptr(tr.p), nomad_tools_matrix_rowidx(tr.p),
                           nomad_tools_matrix_val(tr.p), nnz) ||
        nomad_upload_test(ctx, nomad_tools_matrix_colptr(te.p), nomad_tools_matrix_rowidx(te.p),
                          nomad_tools_matrix_val(te.p), tnnz) ||
        nomad_set_factors(ctx, W.data(), H.data()))
        return done(fail("upload", nomad_last_error()));
    if (o.recover >= 0) {
        // the reference's restore (nb_checkpoint.cpp:116-143): W parts from
        // .state, the columns from .msg, then this build's step-size state
        const int parts = o.nthreads > 0 ? o.nthreads : 1;
        if (nomad_tools_read_checkpoint(o.cp_path.c_str(), o.job_id, o.recover, parts, m, n, k, W.data(), H.data()))
            return done(fail("recover", nomad_tools_last_error()));
        if (nomad_set_factors(ctx, W.data(), H.data())) return done(fail("recover", nomad_last_error()));
        if (restore_state(o, ctx, m, n, k)) return done(1);
    }
    int32_t workers = 0;
    nomad_get_nworkers(ctx, &workers);
    std::cout << "workers: " << workers << std::endl;

    double train_time = 0.0, test_time = 0.0;
    int64_t updates = 0, failures = 0;
    for (size_t it = 0; it < o.timeouts.size(); ++it) {
        const double seg = it == 0 ? o.timeouts[0] : o.timeouts[it] - o.timeouts[it - 1];
        const auto t0 = std::chrono::steady_clock::now();
        nomad_train_stats ts{};
        int rc;
        if (o.mode == NOMAD_MODE_ASYNC) {
            rc = nomad_train(ctx, seg, 0, &ts);
        } else {
            // deterministic schedule: whole B x B sweeps until the segment's time is used
            rc = NOMAD_OK;
            const auto seg_end = t0 + std::chrono::duration<double>(seg);
            do {
                nomad_train_stats one{};
                rc = nomad_train(ctx, 0.0, 1, &one);
                ts.updates += one.updates;
                ts.failed_pops += one.failed_pops;
            } while (rc == NOMAD_OK && std::chrono::steady_clock::now() < seg_end);
        }
        if (rc) return done(fail("nomad_train", nomad_last_error()));
        const auto t1 = std::chrono::steady_clock::now();
        nomad_eval_stats es{};
        if (nomad_eval(ctx, &es)) return done(fail("nomad_eval", nomad_last_error()));
        const auto t2 = std::chrono::steady_clock::now();
        train_time += std::chrono::duration<double>(t1 - t0).count();
        test_time += std::chrono::duration<double>(t2 - t1).count();
        updates += ts.updates;
        failures += ts.failed_pops;

        std::ostringstream buf;
        buf << "Statistics:\n"
            << "=====================================================\n";
        buf << "elapsed time: " << o.timeouts[it] << " total training time: " << train_time
            << " total testing time: " << test_time;
        buf << " current training RMSE: " << std::fixed << std::setprecision(10)
            << std::sqrt(es.train_sse / double(es.train_count)) << "\n";
        buf << "current test RMSE: " << std::fixed << std::setprecision(10)
            << std::sqrt(es.test_sse / double(es.test_count)) << "\n";
        buf << std::fixed << std::setprecision(4) << "detail: "
            << "train: s=" << es.train_sse << ", c=" << es.train_count << "; test: s=" << es.test_sse
            << ", c=" << es.test_count << "; u=" << updates << ", f=" << failures << ", s=" << 0 << "\n";
        buf << "=====================================================";
        std::cout << buf.str() << std::endl;
        if (o.cp_type != "none" && save_checkpoint(o, ctx, int(it), m, n, k, W, H)) return done(1);
        if (o.pause) std::this_thread::sleep_for(std::chrono::duration<double>(3.0));
    }

    if (!o.output.empty()) {
        if (nomad_get_factors(ctx, W.data(), H.data())) return done(fail("nomad_get_factors", nomad_last_error()));
        std::ofstream ofs(o.output + "0", std::ofstream::out);
        if (!ofs) return done(fail("output", "cannot open the output file"));
        for (int32_t i = 0; i < m; ++i) {
            ofs << "row," << i;
            for (int t = 0; t < k; ++t) ofs << "," << W[size_t(i) * k + t];
            ofs << "\n";
        }
        for (int32_t j = 0; j < n; ++j) {
            ofs << "column," << j;
            for (int t = 0; t < k; ++t) ofs << "," << H[size_t(j) * k + t];
            ofs << "\n";
        }
    }
    return done(0);
}
