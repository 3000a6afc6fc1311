// Command-line front ends of the rating-file tooling (libnomad_tools.so), one
// executable per reference program, selected by the name it is run under:
//   nomad_convert train.tsv test.tsv out_dir [--seed 12345]   Scripts/convert.py
//   nomad_synthgen_conf [--seed --nrow --ncol --sep --rowfile --colfile]
//                                                    Code/nomad/synthgen-conf.cpp
//   nomad_synthgen [--seed --dim --nrow --ncol --rowfile --colfile --output]
//                                     Code/nomad/synthgen.cpp (+ gen.sh's concatenation)
//   nomad_gen out_dir nrow ncol [dim] [seed]          gen.sh (both steps)
// Option defaults are the reference programs' own.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <sys/stat.h>

#include "nomad_tools.h"

namespace {

// --name value / --name=value options over the given defaults
bool parse(int argc, char** argv, int first, std::map<std::string, std::string>& opt) {
    for (int i = first; i < argc; ++i) {
        std::string a = argv[i];
        if (a.rfind("--", 0) != 0) {
            std::fprintf(stderr, "unexpected argument: %s\n", argv[i]);
            return false;
        }
        a = a.substr(2);
        std::string v;
        const size_t eq = a.find('=');
        if (eq != std::string::npos) {
            v = a.substr(eq + 1);
            a = a.substr(0, eq);
        } else if (i + 1 < argc) {
            v = argv[++i];
        } else {
            std::fprintf(stderr, "option --%s needs a value\n", a.c_str());
            return false;
        }
        if (!opt.count(a)) {
            std::fprintf(stderr, "unknown option --%s\n", a.c_str());
            return false;
        }
        opt[a] = v;
    }
    return true;
}

int fail() {
    std::fprintf(stderr, "error: %s\n", nomad_tools_last_error());
    return 1;
}

int run_convert(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s [train_filename] [test_filename] [output_path] [--seed 12345]\n", argv[0]);
        return 1;
    }
    std::map<std::string, std::string> o{{"seed", "12345"}};
    if (!parse(argc, argv, 4, o)) return 1;
    int64_t nu = 0, ni = 0;
    if (nomad_tools_convert_tsv(argv[1], argv[2], argv[3], std::strtoull(o["seed"].c_str(), nullptr, 10), &nu, &ni))
        return fail();
    std::printf("users %lld items %lld\n", (long long)nu, (long long)ni);
    return 0;
}

int run_conf(int argc, char** argv) {
    std::map<std::string, std::string> o{{"seed", "12345"}, {"nrow", "480189"}, {"ncol", "17770"}, {"sep", ","},
                                         {"rowfile", "Results/synth/rowdegree.txt"},
                                         {"colfile", "Results/synth/coldegree.txt"}};
    if (!parse(argc, argv, 1, o)) return 1;
    if (nomad_tools_synthgen_conf(std::atoi(o["nrow"].c_str()), std::atoi(o["ncol"].c_str()),
                                  std::atoi(o["seed"].c_str()), o["rowfile"].c_str(), o["colfile"].c_str(),
                                  o["sep"].c_str()))
        return fail();
    std::printf("done\n");
    return 0;
}

int run_gen_step(int argc, char** argv) {
    std::map<std::string, std::string> o{{"seed", "12345"}, {"dim", "100"}, {"nrow", "480189"}, {"ncol", "17770"},
                                         {"rowfile", "./Results/synth/rowdegree.txt"},
                                         {"colfile", "./Results/synth/coldegree.txt"},
                                         {"output", "./Results/synth/"}};
    if (!parse(argc, argv, 1, o)) return 1;
    int64_t a = 0, b = 0;
    if (nomad_tools_synthgen(o["rowfile"].c_str(), o["colfile"].c_str(), o["output"].c_str(),
                             std::atoi(o["nrow"].c_str()), std::atoi(o["ncol"].c_str()), std::atoi(o["dim"].c_str()),
                             std::atoi(o["seed"].c_str()), &a, &b))
        return fail();
    std::printf("train_nnz: %lld\ntest_nnz: %lld\n", (long long)a, (long long)b);
    return 0;
}

int run_gen(int argc, char** argv) {
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s out_dir nrow ncol [dim=20] [seed=12345]\n", argv[0]);
        return 1;
    }
    const std::string pre = argv[1];
    const int nrow = std::atoi(argv[2]), ncol = std::atoi(argv[3]);
    const int dim = argc > 4 ? std::atoi(argv[4]) : 20;
    const int seed = argc > 5 ? std::atoi(argv[5]) : 12345;
    mkdir(pre.c_str(), 0755);
    const std::string rf = pre + "/rowdegree.txt", cf = pre + "/coldegree.txt";
    if (nomad_tools_synthgen_conf(nrow, ncol, seed, rf.c_str(), cf.c_str(), ",")) return fail();
    int64_t a = 0, b = 0;
    if (nomad_tools_synthgen(rf.c_str(), cf.c_str(), pre.c_str(), nrow, ncol, dim, seed, &a, &b)) return fail();
    std::remove(rf.c_str());
    std::remove(cf.c_str());
    std::printf("train_nnz: %lld\ntest_nnz: %lld\nsuccess\n", (long long)a, (long long)b);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const char* name = std::strrchr(argv[0], '/');
    name = name ? name + 1 : argv[0];
    if (std::strcmp(name, "nomad_convert") == 0) return run_convert(argc, argv);
    if (std::strcmp(name, "nomad_synthgen_conf") == 0) return run_conf(argc, argv);
    if (std::strcmp(name, "nomad_synthgen") == 0) return run_gen_step(argc, argv);
    if (std::strcmp(name, "nomad_gen") == 0) return run_gen(argc, argv);
    std::fprintf(stderr, "%s: run as nomad_convert, nomad_synthgen_conf, nomad_synthgen or nomad_gen\n", name);
    return 1;
}
