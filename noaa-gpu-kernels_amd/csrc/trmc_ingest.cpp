// Forcing ingest: per-nexus lateral-inflow CSVs (nex-<id>_output.csv, rows
// "timestep, time, value") pivoted into the [timestep][reach] float tensor
// of forcings.bin, after data_parsers/parse_qlat_fast_cpp.cpp:76-224 or, with
// TRMC_INGEST_RUST, the parsing rules of rust_preprocessor/src/main.rs:28-91,
// 116-273.  Host C++: files are independent, so threads take contiguous chunks
// of the file list and write disjoint reach columns.
#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "trmc.h"
#include "trmc_util.h"

namespace fs = std::filesystem;

namespace {

// "nex-1328591_output.csv" -> 1328591 (digits between the first '-' and the
// next '_' or '.'), -1 when absent or not a number (parse_qlat_fast_cpp.cpp:43-58)
long long nexus_id(const std::string &name)
{
    const size_t dash = name.find('-');
    if (dash == std::string::npos) return -1;
    size_t end = name.find('_', dash + 1);
    if (end == std::string::npos) {
        end = name.find('.', dash + 1);
        if (end == std::string::npos) return -1;
    }
    const std::string num = name.substr(dash + 1, end - dash - 1);
    // std::stoi semantics: optional leading space/sign, digits, stop at the first
    // non-digit; nothing parsed or out of int range is -1
    const char *p = num.c_str();
    char *stop = nullptr;
    errno = 0;
    const long v = std::strtol(p, &stop, 10);
    if (stop == p || errno == ERANGE || v < INT32_MIN || v > INT32_MAX) return -1;
    return (long long)v;
}

// one row: the timestep before the first comma (atoi), the value after the
// last comma (atof, rounded to float) (parse_qlat_fast_cpp.cpp:62-74)
bool parse_row(const char *line, size_t len, int &ts, float &val)
{
    size_t i = 0;
    while (i < len && line[i] != ',') ++i;
    if (i == len) return false;
    ts = std::atoi(line);
    size_t j = len;
    while (j > 0 && line[j - 1] != ',') --j;
    if (j == 0 || j >= len) return false;
    val = (float)std::atof(line + j);
    return true;
}

// Rust rules, file name: "nex-" + a whole signed 64-bit integer up to the
// first '_' (or '.'); -1 marks no id (strip_wbid_from_filename, main.rs:28-33)
bool rust_nexus_id(const std::string &name, long long &id)
{
    if (name.rfind("nex-", 0) != 0) return false;
    const std::string rest = name.substr(4);
    size_t end = rest.find('_');
    if (end == std::string::npos) end = rest.find('.');
    if (end == std::string::npos) return false;
    const std::string num = rest.substr(0, end);
    size_t k = 0;
    if (k < num.size() && (num[k] == '+' || num[k] == '-')) ++k;
    if (k == num.size()) return false;
    for (size_t j = k; j < num.size(); ++j)
        if (num[j] < '0' || num[j] > '9') return false;
    errno = 0;
    const long long v = std::strtoll(num.c_str(), nullptr, 10);
    if (errno == ERANGE) return false;
    id = v;
    return true;
}

// Rust's f32 text grammar: [+-] then inf | infinity | nan (any case), or
// digits with an optional '.' and digits (at least one digit in all) and an
// optional exponent [eE][+-]digits; the value is the correctly rounded float
// (strtof on the accepted text)
bool rust_f32(const char *s, size_t n, float &out)
{
    size_t k = 0;
    if (k < n && (s[k] == '+' || s[k] == '-')) ++k;
    auto word = [&](const char *w) {
        const size_t m = std::strlen(w);
        if (n - k != m) return false;
        for (size_t j = 0; j < m; ++j)
            if ((s[k + j] | 0x20) != w[j]) return false;
        return true;
    };
    if (!(word("inf") || word("infinity") || word("nan"))) {
        size_t digits = 0;
        while (k < n && s[k] >= '0' && s[k] <= '9') { ++k; ++digits; }
        if (k < n && s[k] == '.') {
            ++k;
            while (k < n && s[k] >= '0' && s[k] <= '9') { ++k; ++digits; }
        }
        if (digits == 0) return false;
        if (k < n && (s[k] == 'e' || s[k] == 'E')) {
            ++k;
            if (k < n && (s[k] == '+' || s[k] == '-')) ++k;
            size_t ed = 0;
            while (k < n && s[k] >= '0' && s[k] <= '9') { ++k; ++ed; }
            if (ed == 0) return false;
        }
        if (k != n) return false;
    }
    char tmp[128];
    if (n >= sizeof(tmp)) {
        std::string big(s, n);
        out = std::strtof(big.c_str(), nullptr);
    } else {
        std::memcpy(tmp, s, n);
        tmp[n] = 0;
        out = std::strtof(tmp, nullptr);
    }
    return true;
}

// Rust rules, one row: spaces, the timestep's digits (no sign; none or too
// many for the integer type skip the row), then the text after the last comma
// with spaces skipped and ASCII whitespace trimmed, which must read as a float
// (parse_csv_bytes, main.rs:37-91)
bool parse_row_rust(const char *line, size_t len, unsigned long long &ts, float &val)
{
    size_t k = 0;
    while (k < len && line[k] == ' ') ++k;
    const size_t t0 = k;
    unsigned long long v = 0;
    for (; k < len && line[k] >= '0' && line[k] <= '9'; ++k) {
        const unsigned d = (unsigned)(line[k] - '0');
        if (v > (UINT64_MAX - d) / 10) return false;    // beyond a 64-bit usize
        v = v * 10 + d;
    }
    if (k == t0) return false;
    ts = v;
    size_t c = len;
    while (c > 0 && line[c - 1] != ',') --c;
    if (c == 0) return false;                      // no comma
    size_t b = c;                                  // first byte after the comma
    while (b < len && line[b] == ' ') ++b;
    auto ws = [](char x) { return x == ' ' || x == '\t' || x == '\n' || x == '\v' || x == '\f' || x == '\r'; };
    size_t e = len;
    while (e > b && ws(line[e - 1])) --e;
    while (b < e && ws(line[b])) ++b;
    return rust_f32(line + b, e - b, val);
}

}  // namespace

extern "C" {

int trmc_ingest_nex_csv(const char *csv_dir, const int64_t *topo_to_wbid, int n_reaches, int n_timesteps,
                        int n_threads, float *qlat, int64_t *counts)
{
    return trmc_ingest_nex_csv_rules(csv_dir, topo_to_wbid, n_reaches, n_timesteps, n_threads, TRMC_INGEST_CPP,
                                     qlat, counts);
}

int trmc_ingest_nex_csv_rules(const char *csv_dir, const int64_t *topo_to_wbid, int n_reaches, int n_timesteps,
                              int n_threads, int rules, float *qlat, int64_t *counts)
{
    if (rules != TRMC_INGEST_CPP && rules != TRMC_INGEST_RUST) return trmc::fail(TRMC_ERR_ARG, "unknown rules %d", rules);
    if (!csv_dir || (!topo_to_wbid && n_reaches > 0) || (!qlat && n_reaches > 0 && n_timesteps > 0))
        return trmc::fail(TRMC_ERR_ARG, "null argument");
    if (n_reaches < 0 || n_timesteps < 0) return trmc::fail(TRMC_ERR_ARG, "negative size");
    std::unordered_map<int64_t, int> topo_of;
    topo_of.reserve((size_t)n_reaches * 2);
    for (int i = 0; i < n_reaches; ++i) topo_of[topo_to_wbid[i]] = i;   // a repeated id keeps its last reach
    // file names only; paths are joined in the workers
    std::vector<std::string> names;
    std::error_code ec;
    for (fs::directory_iterator it(csv_dir, ec), end; !ec && it != end; it.increment(ec)) {
        if (!it->is_regular_file(ec)) continue;
        std::string name = it->path().filename().string();
        if (name.rfind("nex-", 0) != 0) continue;
        if (rules == TRMC_INGEST_RUST) {
            if (name.size() < 15 || name.compare(name.size() - 11, 11, "_output.csv") != 0) continue;
        } else if (name.find("_output.csv") == std::string::npos) {
            continue;
        }
        names.push_back(std::move(name));
    }
    if (ec) return trmc::fail(TRMC_ERR_IO, "cannot list %s: %s", csv_dir, ec.message().c_str());
    const std::string dir = std::string(csv_dir) + "/";
    const size_t R = (size_t)n_reaches;
    std::fill(qlat, qlat + (size_t)n_timesteps * R, 0.0f);
    if (n_threads <= 0) n_threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::atomic<long long> matched{0}, missing{0};
    const size_t per = (names.size() + n_threads - 1) / std::max(1, n_threads);
    // a worker that cannot allocate reports it; a range whose thread cannot
    // start is run on this thread
    std::atomic<bool> failed{false};
    auto run_ranges = [&](auto &&fn) {
        auto safe = [&](size_t b, size_t e) {
            try {
                fn(b, e);
            } catch (...) {
                failed.store(true);
            }
        };
        std::vector<std::thread> th;
        for (int t = 0; t < n_threads; ++t) {
            const size_t b = (size_t)t * per, e = std::min(names.size(), b + per);
            if (b >= e) break;
            try {
                th.emplace_back(safe, b, e);
            } catch (...) {
                safe(b, e);
            }
        }
        for (auto &t : th) t.join();
    };
    // each file's reach (-1: no number or unknown id), in parallel
    std::vector<int> reach_of(names.size(), -1);
    run_ranges([&](size_t b, size_t e) {
        for (size_t pi = b; pi < e; ++pi) {
            long long id = -1;
            if (rules == TRMC_INGEST_RUST) {
                if (!rust_nexus_id(names[pi], id)) continue;
            } else {
                id = nexus_id(names[pi]);
                if (id < 0) continue;
            }
            auto f = topo_of.find((int64_t)id);
            if (f != topo_of.end()) reach_of[pi] = f->second;
        }
    });
    if (failed.load()) return trmc::fail(TRMC_ERR_NOMEM, "out of host memory reading the forcing files");
    // One owner per reach column: of the files naming it, the last in name
    // order (a fixed choice, whatever order the directory lists them in).
    // C++ rules: every file writes its own rows, so a reach named by several
    // files takes, cell by cell, the row of the last file in name order that
    // has that timestep; the owner's thread applies the group in name order
    // (columns stay disjoint between threads).  Rust rules: the owner's rows
    // alone (the whole column of one file).
    std::vector<int> owner(R, -1);
    std::vector<char> several(R, 0);             // named by more than one file
    std::vector<int> group_of(R, -1);            // C++ rules: reach -> its group (several files only)
    std::vector<int> group_ptr, group_files;     // files of those reaches, in name order
    for (size_t pi = 0; pi < names.size(); ++pi) {
        const int r = reach_of[pi];
        if (r < 0) continue;
        if (owner[r] >= 0) several[r] = 1;
        if (owner[r] < 0 || names[pi] > names[owner[r]]) owner[r] = (int)pi;
    }
    if (rules == TRMC_INGEST_CPP) {
        std::vector<std::pair<int, int>> shared;    // (reach, file)
        for (size_t pi = 0; pi < names.size(); ++pi)
            if (reach_of[pi] >= 0 && several[reach_of[pi]]) shared.emplace_back(reach_of[pi], (int)pi);
        std::sort(shared.begin(), shared.end(), [&](const std::pair<int, int> &x, const std::pair<int, int> &y) {
            return x.first != y.first ? x.first < y.first : names[x.second] < names[y.second];
        });
        for (size_t i = 0; i < shared.size(); ++i) {
            if (i == 0 || shared[i].first != shared[i - 1].first) {
                group_of[shared[i].first] = (int)group_ptr.size();
                group_ptr.push_back((int)group_files.size());
            }
            group_files.push_back(shared[i].second);
        }
        group_ptr.push_back((int)group_files.size());
    }
    auto work = [&](size_t b, size_t e) {
        std::vector<char> buf(4096);
        std::string path;
        // read one file whole into buf; false when it cannot be read
        auto read_file = [&](size_t pi, size_t &got) {
            path.assign(dir).append(names[pi]);
            FILE *fp = std::fopen(path.c_str(), "rb");
            if (!fp) return false;
            std::fseek(fp, 0, SEEK_END);
            const long size = std::ftell(fp);
            std::fseek(fp, 0, SEEK_SET);
            if (size < 0) { std::fclose(fp); return false; }
            if ((long)buf.size() < size + 1) buf.resize((size_t)size + 1);
            got = std::fread(buf.data(), 1, (size_t)size, fp);
            buf[got] = 0;
            std::fclose(fp);
            return true;
        };
        auto apply_rows = [&](size_t got, int r) {
            size_t i = 0;
            while (i < got) {
                const size_t ls = i;
                while (i < got && buf[i] != '\n') ++i;
                const size_t ll = i - ls;
                if (i < got) ++i;
                if (ll < 5) continue;
                float val;
                if (rules == TRMC_INGEST_RUST) {
                    unsigned long long ts;
                    if (!parse_row_rust(buf.data() + ls, ll, ts, val)) continue;
                    if (ts < (unsigned long long)n_timesteps) qlat[(size_t)ts * R + r] = val;
                } else {
                    int ts;
                    if (!parse_row(buf.data() + ls, ll, ts, val)) continue;
                    if (ts >= 0 && ts < n_timesteps) qlat[(size_t)ts * R + r] = val;
                }
            }
        };
        for (size_t pi = b; pi < e; ++pi) {
            const int r = reach_of[pi];
            if (r < 0) { missing.fetch_add(1); continue; }
            if (group_of[r] >= 0) {
                // the group's files are read, counted and applied by the owner's thread
                if (owner[r] != (int)pi) continue;
                const int gi = group_of[r];
                for (int f = group_ptr[gi]; f < group_ptr[gi + 1]; ++f) {
                    size_t got = 0;
                    if (!read_file((size_t)group_files[f], got)) { missing.fetch_add(1); continue; }
                    matched.fetch_add(1);
                    apply_rows(got, r);
                }
                continue;
            }
            size_t got = 0;
            if (!read_file(pi, got)) { missing.fetch_add(1); continue; }
            matched.fetch_add(1);
            if (owner[r] != (int)pi) continue;   // Rust rules: a later file names the same reach
            apply_rows(got, r);
        }
    };
    run_ranges(work);
    if (failed.load()) return trmc::fail(TRMC_ERR_NOMEM, "out of host memory reading the forcing files");
    if (counts) {
        counts[0] = matched.load();
        counts[1] = missing.load();
    }
    return TRMC_OK;
}

int trmc_ingest_nex_csv_to_forcings(const char *csv_dir, const char *wbid_npy_path, const char *out_path,
                                    int n_timesteps, int n_threads, int64_t *counts)
{
    return trmc_ingest_nex_csv_to_forcings_rules(csv_dir, wbid_npy_path, out_path, n_timesteps, n_threads,
                                                 TRMC_INGEST_CPP, counts);
}

int trmc_ingest_nex_csv_to_forcings_rules(const char *csv_dir, const char *wbid_npy_path, const char *out_path,
                                          int n_timesteps, int n_threads, int rules, int64_t *counts)
{
    if (!csv_dir || !wbid_npy_path || !out_path) return trmc::fail(TRMC_ERR_ARG, "null path");
    // topo_to_wbid.npy: int64 array after the .npy header (magic 6, version 2,
    // little-endian header length 2, header) (parse_qlat_fast_cpp.cpp:90-110)
    FILE *f = std::fopen(wbid_npy_path, "rb");
    if (!f) return trmc::fail(TRMC_ERR_IO, "cannot open %s", wbid_npy_path);
    unsigned char pre[10];
    if (std::fread(pre, 1, 10, f) != 10 || std::memcmp(pre, "\x93NUMPY", 6) != 0) {
        std::fclose(f);
        return trmc::fail(TRMC_ERR_IO, "%s is not a .npy file", wbid_npy_path);
    }
    const long header_end = 10 + (long)(pre[8] | (pre[9] << 8));
    std::fseek(f, 0, SEEK_END);
    const long total = std::ftell(f);
    const size_t R = total > header_end ? (size_t)(total - header_end) / 8 : 0;
    std::vector<int64_t> wbid(R);
    std::fseek(f, header_end, SEEK_SET);
    const size_t got = R ? std::fread(wbid.data(), 8, R, f) : 0;
    std::fclose(f);
    if (got != R) return trmc::fail(TRMC_ERR_IO, "short read in %s", wbid_npy_path);
    std::vector<float> qlat((size_t)std::max(n_timesteps, 0) * R);
    int rc = trmc_ingest_nex_csv_rules(csv_dir, wbid.data(), (int)R, n_timesteps, n_threads, rules, qlat.data(),
                                       counts);
    if (rc) return rc;
    // forcings.bin: int32 T, qlat [T][R], qup0 [R] = 0, qdp0 [R] = 0, dp0 [R] = 0.5
    FILE *o = std::fopen(out_path, "wb");
    if (!o) return trmc::fail(TRMC_ERR_IO, "cannot write %s", out_path);
    const int32_t nt = n_timesteps;
    std::vector<float> zeros(R, 0.0f), half(R, 0.5f);
    bool ok = std::fwrite(&nt, 4, 1, o) == 1 && std::fwrite(qlat.data(), 4, qlat.size(), o) == qlat.size() &&
              std::fwrite(zeros.data(), 4, R, o) == R && std::fwrite(zeros.data(), 4, R, o) == R &&
              std::fwrite(half.data(), 4, R, o) == R;
    ok = (std::fclose(o) == 0) && ok;
    return ok ? TRMC_OK : trmc::fail(TRMC_ERR_IO, "write failed: %s", out_path);
}

}  // extern "C"
