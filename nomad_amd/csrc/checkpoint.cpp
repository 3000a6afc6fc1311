// Checkpoint files in the reference's layout (Code/nomad/nb_checkpoint.cpp):
//   <dir>/cp-t<job>-e<epoch>-p<part>.state   int32 local_num_rows, int32 dim,
//                                             double rows[local_num_rows * dim]
//                                             (archive_local, :50-60)
//   <dir>/cp-t<job>-e<epoch>-p<part>.msg     column records, int32 col_index,
//                                             int64 flag, double values[dim],
//                                             packed (ColumnData::serialize,
//                                             pool.cpp:17-23; _archive_msg_queue)
//   <dir>/cp-t<job>-e<epoch>-p0.smsg          the send queue (empty at one rank)
// Part p holds rows [p * ceil(m / P), min((p + 1) * ceil(m / P), m)) (the
// reference's part split, nb_data.cpp:83-89).  A checkpoint here is taken
// between training segments, when every column is at rest (the reference's
// "sync" protocol at one rank); column j is written to the .msg file of part
// j mod P, whose thread it would be queued at on a restore.
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "convert.h"

namespace nomad {

namespace {

std::string cp_name(const std::string& dir, int job, int epoch, int part) {
    return dir + (dir.empty() ? "" : "/") + "cp-t" + std::to_string(job) + "-e" + std::to_string(epoch) + "-p" +
           std::to_string(part);
}

constexpr size_t kRecHead = sizeof(int32_t) + sizeof(int64_t);  // col_index, flag

}  // namespace

bool write_checkpoint(const std::string& dir, int job, int epoch, int nparts, int32_t m, int32_t n, int32_t k,
                      const double* W, const double* H, std::string& err) {
    if (nparts < 1 || m < 1 || n < 1 || k < 1) {
        err = "write_checkpoint: bad sizes";
        return false;
    }
    const int64_t per = m / nparts + ((m % nparts) > 0 ? 1 : 0);
    std::vector<char> rec(kRecHead + sizeof(double) * size_t(k));
    for (int p = 0; p < nparts; ++p) {
        const int64_t lo = std::min<int64_t>(per * p, m), hi = std::min<int64_t>(per * (p + 1), m);
        const int32_t rows = int32_t(hi - lo);
        const std::string base = cp_name(dir, job, epoch, p);
        std::FILE* f = std::fopen((base + ".state").c_str(), "wb");
        if (!f) {
            err = "cannot write " + base + ".state";
            return false;
        }
        bool ok = std::fwrite(&rows, 4, 1, f) == 1 && std::fwrite(&k, 4, 1, f) == 1 &&
                  std::fwrite(W + lo * k, sizeof(double), size_t(rows) * size_t(k), f) == size_t(rows) * size_t(k);
        ok = (std::fclose(f) == 0) && ok;
        f = std::fopen((base + ".msg").c_str(), "wb");
        if (!ok || !f) {
            if (f) std::fclose(f);
            err = "cannot write " + base;
            return false;
        }
        for (int32_t j = p; j < n && ok; j += nparts) {
            const int64_t flag = 0;
            std::memcpy(rec.data(), &j, 4);
            std::memcpy(rec.data() + 4, &flag, 8);
            std::memcpy(rec.data() + kRecHead, H + int64_t(j) * k, sizeof(double) * size_t(k));
            ok = std::fwrite(rec.data(), 1, rec.size(), f) == rec.size();
        }
        ok = (std::fclose(f) == 0) && ok;
        if (ok && p == 0) {
            f = std::fopen((base + ".smsg").c_str(), "wb");
            ok = f && std::fclose(f) == 0;
        }
        if (!ok) {
            err = "write error on " + base;
            return false;
        }
    }
    return true;
}

bool read_checkpoint(const std::string& dir, int job, int epoch, int nparts, int32_t m, int32_t n, int32_t k,
                     double* W, double* H, std::string& err) {
    if (nparts < 1 || m < 1 || n < 1 || k < 1) {
        err = "read_checkpoint: bad sizes";
        return false;
    }
    const int64_t per = m / nparts + ((m % nparts) > 0 ? 1 : 0);
    std::vector<char> seen(size_t(n), 0);
    std::vector<char> rec(kRecHead + sizeof(double) * size_t(k));
    for (int p = 0; p < nparts; ++p) {
        const int64_t lo = std::min<int64_t>(per * p, m), hi = std::min<int64_t>(per * (p + 1), m);
        const std::string base = cp_name(dir, job, epoch, p);
        std::FILE* f = std::fopen((base + ".state").c_str(), "rb");
        if (!f) {
            err = "cannot open " + base + ".state";
            return false;
        }
        int32_t rows = 0, dim = 0;
        bool ok = std::fread(&rows, 4, 1, f) == 1 && std::fread(&dim, 4, 1, f) == 1;
        if (ok && (rows != hi - lo || dim != k)) {
            std::fclose(f);
            err = base + ".state: " + std::to_string(rows) + " rows of dimension " + std::to_string(dim) +
                  ", expected " + std::to_string(hi - lo) + " of " + std::to_string(k);
            return false;
        }
        ok = ok && std::fread(W + lo * k, sizeof(double), size_t(rows) * size_t(k), f) == size_t(rows) * size_t(k);
        std::fclose(f);
        if (!ok) {
            err = "short read on " + base + ".state";
            return false;
        }
        for (const char* suffix : {".msg", ".smsg"}) {
            f = std::fopen((base + suffix).c_str(), "rb");
            if (!f) {
                if (std::strcmp(suffix, ".smsg") == 0) continue;  // written for part 0 only
                err = "cannot open " + base + suffix;
                return false;
            }
            // whole records only (a trailing partial record is an error)
            for (;;) {
                const size_t got = std::fread(rec.data(), 1, rec.size(), f);
                if (got == 0) break;
                if (got != rec.size()) {
                    std::fclose(f);
                    err = base + suffix + ": truncated column record";
                    return false;
                }
                int32_t j = 0;
                std::memcpy(&j, rec.data(), 4);
                if (j < 0 || j >= n) {
                    std::fclose(f);
                    err = base + suffix + ": column index out of range";
                    return false;
                }
                std::memcpy(H + int64_t(j) * k, rec.data() + kRecHead, sizeof(double) * size_t(k));
                seen[size_t(j)] = 1;
            }
            std::fclose(f);
        }
    }
    for (int32_t j = 0; j < n; ++j)
        if (!seen[size_t(j)]) {
            err = "checkpoint holds no record of column " + std::to_string(j);
            return false;
        }
    return true;
}

}  // namespace nomad
