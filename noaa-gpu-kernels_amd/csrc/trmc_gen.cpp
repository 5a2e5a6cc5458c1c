// Seeded synthetic workloads, binary file formats and error statistics.
//
// Generator design (builder choices for the benchmark workloads):
//  * counter-based random streams (SplitMix64 finaliser over seed/stream/index),
//    so values do not depend on thread count or call order;
//  * log/exp implemented here from +,-,*,/ only (no libm), so the same seed
//    gives the same bytes on every host whatever its libm build;
//  * every generated real is rounded to float before it is widened, so the
//    FP32-input and FP64 entry points see exactly equal values.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "trmc.h"
#include "trmc_rng.h"
#include "trmc_util.h"

namespace trmc {

using rng::det_exp;
using rng::det_log;
using rng::mix64;
using rng::normal_at;
using rng::stream_key;
using rng::uniform_at;

static inline double f32r(double x) { return (double)(float)x; }

// --------------------------------------------------------------- topology
namespace {
struct Tree {
    // c1 == -1 for single-upstream (chain) reaches, c2 == -1 below 3-way confluences
    std::vector<int> parent, c0, c1, c2;
    int root = 0;
    int add() {
        parent.push_back(-1); c0.push_back(-1); c1.push_back(-1); c2.push_back(-1);
        return (int)parent.size() - 1;
    }
    void replace_child(int p, int oldc, int newc) {
        if (p < 0) { root = newc; return; }
        if (c0[p] == oldc) c0[p] = newc;
        else if (c1[p] == oldc) c1[p] = newc;
        else c2[p] = newc;
    }
};
}  // namespace

// Uniform random binary topology (Shreve's topologically random networks)
// grown by random link attachment (Remy's construction): each new headwater
// splits a uniformly chosen existing link.  Then `n_chain` single-upstream
// reaches are inserted above uniformly chosen links.
static Tree grow_tree(int n_leaves, int n_chain, uint64_t seed)
{
    Tree t;
    uint64_t key = stream_key(seed, 101);
    uint64_t draw = 0;
    t.add();
    t.root = 0;
    for (int i = 1; i < n_leaves; ++i) {
        int existing = (int)t.parent.size();
        int x = std::min(existing - 1, (int)(uniform_at(key, draw++) * existing));
        int y = t.add(), z = t.add();
        int p = t.parent[x];
        t.parent[y] = p;
        t.replace_child(p, x, y);
        bool flip = uniform_at(key, draw++) < 0.5;
        t.c0[y] = flip ? z : x;
        t.c1[y] = flip ? x : z;
        t.parent[x] = y;
        t.parent[z] = y;
    }
    for (int i = 0; i < n_chain; ++i) {
        int existing = (int)t.parent.size();
        int x = std::min(existing - 1, (int)(uniform_at(key, draw++) * existing));
        int c = t.add();
        int p = t.parent[x];
        t.parent[c] = p;
        t.replace_child(p, x, c);
        t.c0[c] = x;
        t.parent[x] = c;
    }
    return t;
}

// Poisson(lambda) draw from uniforms of stream `key` (product of uniforms
// until it falls below e^-lambda; lambda is small here)
static int poisson_at(uint64_t key, uint64_t &draw, double lambda)
{
    const double lim = det_exp(-lambda);
    double prod = uniform_at(key, draw++);
    int k = 0;
    while (prod > lim) {
        prod *= uniform_at(key, draw++);
        ++k;
    }
    return k;
}

// Balanced random merge tree with chains, the construction of the
// reference's network generator (wavefront/gen_network.py:62-246): B
// headwaters are merged level by level in random order, three at a time with
// probability 0.34 and two otherwise (a lone leftover carries over); then
// every link gets Poisson(mean_chain - 1) single-upstream reaches inserted.
// Shallow (tens of levels) and wide, as the reference's benchmark networks.
static Tree merge_tree(int n_leaves, double mean_chain, uint64_t seed)
{
    Tree t;
    uint64_t key = stream_key(seed, 111);
    uint64_t draw = 0;
    std::vector<int> cur(n_leaves), next;
    for (int i = 0; i < n_leaves; ++i) cur[i] = t.add();
    while (cur.size() > 1) {
        for (size_t i = cur.size() - 1; i > 0; --i) {           // Fisher-Yates
            size_t j = std::min(i, (size_t)(uniform_at(key, draw++) * (double)(i + 1)));
            std::swap(cur[i], cur[j]);
        }
        next.clear();
        size_t i = 0;
        while (i < cur.size()) {
            const double r = uniform_at(key, draw++);
            size_t k = (r < 0.34 && i + 3 <= cur.size()) ? 3 : (i + 2 <= cur.size() ? 2 : 1);
            if (k == 1) { next.push_back(cur[i++]); continue; }
            const int p = t.add();
            t.c0[p] = cur[i];
            t.c1[p] = cur[i + 1];
            if (k == 3) t.c2[p] = cur[i + 2];
            for (size_t c = 0; c < k; ++c) t.parent[cur[i + c]] = p;
            next.push_back(p);
            i += k;
        }
        cur.swap(next);
    }
    t.root = cur.empty() ? 0 : cur[0];
    const int nodes = (int)t.parent.size();
    for (int u = 0; u < nodes; ++u) {
        if (t.parent[u] < 0) continue;
        const int extra = poisson_at(key, draw, std::max(0.0, mean_chain - 1.0));
        int below = u;
        for (int e = 0; e < extra; ++e) {                         // u -> c -> parent
            const int c = t.add();
            const int p = t.parent[below];
            t.parent[c] = p;
            t.replace_child(p, below, c);
            t.c0[c] = below;
            t.parent[below] = c;
            below = c;
        }
    }
    return t;
}

static int lay_out(const Tree &t, const std::vector<int> &lens, uint64_t seed, trmc_topo *out)
{
    const int R = (int)t.parent.size();
    // levels: headwater 0, else 1 + max(upstream levels) (reverse BFS order)
    std::vector<int> order;
    order.reserve(R);
    order.push_back(t.root);
    for (size_t i = 0; i < order.size(); ++i) {
        int n = order[i];
        if (t.c0[n] >= 0) order.push_back(t.c0[n]);
        if (t.c1[n] >= 0) order.push_back(t.c1[n]);
        if (t.c2[n] >= 0) order.push_back(t.c2[n]);
    }
    if ((int)order.size() != R) return TRMC_ERR_ARG;
    std::vector<int> level(R, 0);
    int n_levels = 0;
    for (int i = R - 1; i >= 0; --i) {
        int n = order[i], l = 0;
        if (t.c0[n] >= 0) l = std::max(l, level[t.c0[n]] + 1);
        if (t.c1[n] >= 0) l = std::max(l, level[t.c1[n]] + 1);
        if (t.c2[n] >= 0) l = std::max(l, level[t.c2[n]] + 1);
        level[n] = l;
        n_levels = std::max(n_levels, l + 1);
    }
    // relabel reaches in (level, random) order: ids are topological
    uint64_t key = stream_key(seed, 202);
    std::vector<std::pair<uint64_t, int>> keyed(R);
    for (int n = 0; n < R; ++n)
        keyed[n] = {((uint64_t)level[n] << 40) | (mix64(key ^ (uint64_t)n) >> 24), n};
    std::sort(keyed.begin(), keyed.end());
    std::vector<int> id_of(R);
    for (int i = 0; i < R; ++i) id_of[keyed[i].second] = i;

    int N = 0;
    for (int l : lens) N += l;
    std::memset(out, 0, sizeof(*out));
    out->n_reaches = R;
    out->n_segments = N;
    out->n_levels = n_levels;
    out->reach_seg_start = (int *)std::malloc(sizeof(int) * (R + 1));
    out->reach_seg_len = (int *)std::malloc(sizeof(int) * (R + 1));
    out->reach_level = (int *)std::malloc(sizeof(int) * (R + 1));
    out->reach_n_up = (int *)std::malloc(sizeof(int) * (R + 1));
    out->reach_up_ptr = (int *)std::malloc(sizeof(int) * (R + 1));
    out->reach_up_idx = (int *)std::malloc(sizeof(int) * (3 * R + 1));
    out->level_ptr = (int *)std::malloc(sizeof(int) * (n_levels + 1));
    out->level_reach = (int *)std::malloc(sizeof(int) * (R + 1));
    if (!out->reach_seg_start || !out->reach_seg_len || !out->reach_level || !out->reach_n_up ||
        !out->reach_up_ptr || !out->reach_up_idx || !out->level_ptr || !out->level_reach) {
        trmc_free_topo(out);
        return TRMC_ERR_NOMEM;
    }
    int seg = 0, nnz = 0, max_up = 0;
    out->reach_up_ptr[0] = 0;
    for (int i = 0; i < R; ++i) {
        int n = keyed[i].second;
        out->reach_seg_start[i] = seg;
        out->reach_seg_len[i] = lens[i];
        seg += lens[i];
        out->reach_level[i] = level[n];
        int ups[3], nu = 0;
        if (t.c0[n] >= 0) ups[nu++] = id_of[t.c0[n]];
        if (t.c1[n] >= 0) ups[nu++] = id_of[t.c1[n]];
        if (t.c2[n] >= 0) ups[nu++] = id_of[t.c2[n]];
        std::sort(ups, ups + nu);                                   // sorted upstream lists
        for (int k = 0; k < nu; ++k) out->reach_up_idx[nnz++] = ups[k];
        out->reach_n_up[i] = nu;
        max_up = std::max(max_up, nu);
        out->reach_up_ptr[i + 1] = nnz;
    }
    out->max_upstream = max_up;
    for (int L = 0; L <= n_levels; ++L) out->level_ptr[L] = 0;
    for (int i = 0; i < R; ++i) out->level_ptr[out->reach_level[i] + 1]++;
    for (int L = 0; L < n_levels; ++L) out->level_ptr[L + 1] += out->level_ptr[L];
    for (int i = 0; i < R; ++i) out->level_reach[i] = i;  // ids are sorted by level
    return TRMC_OK;
}

int generate_network(int kind, int n_segments, int n_reaches, uint64_t seed, trmc_topo *out)
{
    if (!out || n_segments <= 0) return TRMC_ERR_ARG;
    uint64_t key = stream_key(seed, 303);
    uint64_t draw = 0;
    std::vector<int> lens;
    if (kind == 1) {
        // reaches of U{1..40} segments until the total reaches n_segments
        int total = 0;
        while (total < n_segments) {
            int l = 1 + std::min(39, (int)(uniform_at(key, draw++) * 40.0));
            l = std::min(l, n_segments - total);
            lens.push_back(l);
            total += l;
        }
        int R = (int)lens.size();
        int leaves = (R + 1) / 2;
        Tree t = grow_tree(leaves, R - (2 * leaves - 1), seed);
        return lay_out(t, lens, seed, out);
    }
    if (kind == 2) {
        int R = n_reaches > 0 ? n_reaches : std::max(1, n_segments * 3 / 10);
        if (R > n_segments) R = n_segments;
        int leaves = std::max(1, R / 2);
        int binary = 2 * leaves - 1;
        if (binary > R) { leaves = (R + 1) / 2; binary = 2 * leaves - 1; }
        Tree t = grow_tree(leaves, R - binary, seed);
        // geometric lengths on {1..40} with mean n_segments / R, then nudge
        // random reaches by one segment until the total is exact
        double mean = (double)n_segments / (double)R;
        double p = std::min(1.0, 1.0 / mean);
        long long total = 0;
        lens.resize(R);
        for (int i = 0; i < R; ++i) {
            double u = uniform_at(key, draw++);
            int l = (p >= 1.0) ? 1 : 1 + (int)std::floor(det_log(u) / det_log(1.0 - p));
            l = std::max(1, std::min(40, l));
            lens[i] = l;
            total += l;
        }
        while (total != n_segments) {
            int i = std::min(R - 1, (int)(uniform_at(key, draw++) * R));
            if (total < n_segments && lens[i] < 40) { lens[i]++; total++; }
            else if (total > n_segments && lens[i] > 1) { lens[i]--; total--; }
        }
        return lay_out(t, lens, seed, out);
    }
    if (kind == 3) {
        // n_reaches is a target as in the reference's generator (B leaves,
        // B = n_reaches / (2 mean_chain)); each reach max(1, Poisson(12))
        // segments; n_segments is ignored
        const double mean_chain = 3.0;
        const int leaves = std::max(2, (int)std::lround((double)std::max(n_reaches, 1) / (mean_chain * 2.0)));
        Tree t = merge_tree(leaves, mean_chain, seed);
        const int R = (int)t.parent.size();
        lens.resize(R);
        long long total = 0;
        for (int i = 0; i < R; ++i) {
            lens[i] = std::max(1, poisson_at(key, draw, 12.0));
            total += lens[i];
        }
        if (total > 0x7fffffffLL) return TRMC_ERR_ARG;
        return lay_out(t, lens, seed, out);
    }
    return TRMC_ERR_ARG;
}

// Segment parameters: dx U[500,5000] m, n U[0.03,0.08], bw U[1,50] m,
// side slope cs U[1,3], tw = bw + U[2,20], twcc = tw + U[5,100],
// ncc = n * U[1.2,2.0], S0 lognormal (median 1e-3, sigma_ln 1) clipped to
// [1e-4, 1e-2].
int generate_params(int n, uint64_t seed, trmc_chparams_f64 *out)
{
    if (!out || n < 0) return TRMC_ERR_ARG;
    uint64_t k[9];
    for (int f = 0; f < 9; ++f) k[f] = stream_key(seed, 400 + f);
    for (int i = 0; i < n; ++i) {
        trmc_chparams_f64 p;
        p.dx = f32r(500.0 + 4500.0 * uniform_at(k[0], i));
        p.n_ch = f32r(0.03 + 0.05 * uniform_at(k[1], i));
        p.bw = f32r(1.0 + 49.0 * uniform_at(k[2], i));
        p.cs = f32r(1.0 + 2.0 * uniform_at(k[3], i));
        p.tw = f32r(p.bw + 2.0 + 18.0 * uniform_at(k[4], i));
        p.twcc = f32r(p.tw + 5.0 + 95.0 * uniform_at(k[5], i));
        p.ncc = f32r(p.n_ch * (1.2 + 0.8 * uniform_at(k[6], i)));
        double s0 = det_exp(det_log(1e-3) + normal_at(k[7], i));
        p.s0 = f32r(std::min(1e-2, std::max(1e-4, s0)));
        out[i] = p;
    }
    return TRMC_OK;
}

// Lateral inflow: lognormal, mean 0.1 m3/s, sigma_ln = 1 (mu = ln 0.1 - 1/2).
// Members [m0, m0 + M) of the stream family of `seed`.
int generate_forcing_members(int n, int T, int m0, int M, uint64_t seed, double *qlat)
{
    if (!qlat || n < 0 || T < 0 || M < 1 || m0 < 0) return TRMC_ERR_ARG;
    const size_t per_member = (size_t)n * (size_t)T;
    const size_t total = per_member * (size_t)M;
    unsigned nthreads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (total < (size_t)1 << 16) nthreads = 1;
    auto work = [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i) {
            const size_t m = i / per_member, j = i - m * per_member;
            qlat[i] = rng::forcing_at(seed, m0 + (int)m, j);
        }
    };
    std::vector<std::thread> pool;
    size_t chunk = (total + nthreads - 1) / nthreads;
    for (unsigned t = 0; t < nthreads; ++t) {
        size_t b = t * chunk, e = std::min(total, b + chunk);
        if (b < e) pool.emplace_back(work, b, e);
    }
    for (auto &th : pool) th.join();
    return TRMC_OK;
}

int generate_forcing(int n, int T, int M, uint64_t seed, double *qlat)
{
    return generate_forcing_members(n, T, 0, M, seed, qlat);
}

// qdp0 = q0 and dp0 = wide-channel Manning depth (q0 n / (bw sqrt S0))^(3/5).
int generate_state(int n, const trmc_chparams_f64 *p, double q0, double *qdp0, double *dp0)
{
    if (n < 0 || !p || !qdp0 || !dp0) return TRMC_ERR_ARG;
    for (int i = 0; i < n; ++i) {
        qdp0[i] = f32r(q0);
        double x = q0 * p[i].n_ch / (p[i].bw * std::sqrt(p[i].s0));
        dp0[i] = f32r(x > 0.0 ? det_exp(0.6 * det_log(x)) : 0.0);
    }
    return TRMC_OK;
}

// ---------------------------------------------------------- file formats
// Layouts of wavefront/gen_network.py:10-37, read as wavefront_mc.cu:599-672.

static bool read_exact(FILE *f, void *buf, size_t n)
{
    return n == 0 || std::fread(buf, 1, n, f) == n;
}

int load_topo(const char *path, trmc_topo *t, std::string &err)
{
    std::memset(t, 0, sizeof(*t));
    FILE *f = std::fopen(path, "rb");
    if (!f) { err = std::string("cannot open ") + path; return TRMC_ERR_IO; }
    int hdr[4];
    int rc = TRMC_OK;
    if (!read_exact(f, hdr, sizeof(hdr)) || hdr[0] < 0 || hdr[1] < 0 || hdr[2] < 0) {
        std::fclose(f);
        err = "short or invalid topo header";
        return TRMC_ERR_IO;
    }
    int R = hdr[0], L = hdr[2];
    t->n_reaches = R; t->n_segments = hdr[1]; t->n_levels = L; t->max_upstream = hdr[3];
    auto grab = [&](int **dst, size_t count) {
        *dst = (int *)std::malloc(sizeof(int) * (count + 1));
        if (!*dst) { rc = TRMC_ERR_NOMEM; return false; }
        if (!read_exact(f, *dst, sizeof(int) * count)) { rc = TRMC_ERR_IO; return false; }
        return true;
    };
    bool ok = grab(&t->reach_seg_start, R) && grab(&t->reach_seg_len, R) &&
              grab(&t->reach_level, R) && grab(&t->reach_n_up, R) &&
              grab(&t->reach_up_ptr, (size_t)R + 1);
    if (ok) {
        int nnz = t->reach_up_ptr[R];
        if (nnz < 0) { rc = TRMC_ERR_IO; ok = false; }
        else ok = grab(&t->reach_up_idx, nnz) && grab(&t->level_ptr, (size_t)L + 1) &&
                  grab(&t->level_reach, R);
    }
    std::fclose(f);
    if (!ok) {
        trmc_free_topo(t);
        err = std::string("short read in ") + path;
        return rc == TRMC_OK ? TRMC_ERR_IO : rc;
    }
    return TRMC_OK;
}

// Disjoint union: network i's segments and reaches follow network i-1's;
// level L of the union lists level L of every network, in network order.
int concat_topo(const trmc_topo *nets, int n, trmc_topo *out, int *seg_offset, int *reach_offset)
{
    long long R = 0, N = 0, nnz = 0;
    int L = 0, max_up = 0;
    for (int i = 0; i < n; ++i) {
        const trmc_topo &t = nets[i];
        if (t.n_reaches < 0 || t.n_segments < 0 || t.n_levels < 0) return TRMC_ERR_ARG;
        if (t.n_reaches > 0 && (!t.reach_seg_start || !t.reach_seg_len || !t.reach_up_ptr || !t.level_ptr ||
                                !t.level_reach))
            return TRMC_ERR_ARG;
        R += t.n_reaches;
        N += t.n_segments;
        nnz += t.n_reaches > 0 ? t.reach_up_ptr[t.n_reaches] : 0;
        L = std::max(L, t.n_levels);
        max_up = std::max(max_up, t.max_upstream);
    }
    if (R > INT32_MAX / 2 || N > INT32_MAX / 2 || nnz > INT32_MAX / 2) return TRMC_ERR_ARG;
    *out = trmc_topo{};
    out->n_reaches = (int)R;
    out->n_segments = (int)N;
    out->n_levels = L;
    out->max_upstream = max_up;
    auto alloc = [](size_t k) { return (int *)std::malloc(sizeof(int) * (k + 1)); };
    out->reach_seg_start = alloc(R);
    out->reach_seg_len = alloc(R);
    out->reach_level = alloc(R);
    out->reach_n_up = alloc(R);
    out->reach_up_ptr = alloc(R + 1);
    out->reach_up_idx = alloc(nnz);
    out->level_ptr = alloc(L + 1);
    out->level_reach = alloc(R);
    if (!out->reach_seg_start || !out->reach_seg_len || !out->reach_level || !out->reach_n_up ||
        !out->reach_up_ptr || !out->reach_up_idx || !out->level_ptr || !out->level_reach)
        return TRMC_ERR_NOMEM;
    int r0 = 0, s0 = 0, k0 = 0;
    out->reach_up_ptr[0] = 0;
    for (int i = 0; i < n; ++i) {
        const trmc_topo &t = nets[i];
        if (seg_offset) seg_offset[i] = s0;
        if (reach_offset) reach_offset[i] = r0;
        for (int r = 0; r < t.n_reaches; ++r) {
            out->reach_seg_start[r0 + r] = t.reach_seg_start[r] + s0;
            out->reach_seg_len[r0 + r] = t.reach_seg_len[r];
            out->reach_level[r0 + r] = t.reach_level ? t.reach_level[r] : 0;
            out->reach_n_up[r0 + r] = t.reach_up_ptr[r + 1] - t.reach_up_ptr[r];
            for (int k = t.reach_up_ptr[r]; k < t.reach_up_ptr[r + 1]; ++k)
                out->reach_up_idx[k0 + k] = t.reach_up_idx[k] + r0;
            out->reach_up_ptr[r0 + r + 1] = k0 + t.reach_up_ptr[r + 1];
        }
        k0 += t.n_reaches > 0 ? t.reach_up_ptr[t.n_reaches] : 0;
        r0 += t.n_reaches;
        s0 += t.n_segments;
    }
    if (seg_offset) seg_offset[n] = s0;
    if (reach_offset) reach_offset[n] = r0;
    int w = 0;
    out->level_ptr[0] = 0;
    for (int lv = 0; lv < L; ++lv) {
        int rb = 0;
        for (int i = 0; i < n; ++i) {
            const trmc_topo &t = nets[i];
            if (lv < t.n_levels)
                for (int k = t.level_ptr[lv]; k < t.level_ptr[lv + 1]; ++k) out->level_reach[w++] = t.level_reach[k] + rb;
            rb += t.n_reaches;
        }
        out->level_ptr[lv + 1] = w;
    }
    return TRMC_OK;
}

int save_topo(const char *path, const trmc_topo *t)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return TRMC_ERR_IO;
    int hdr[4] = {t->n_reaches, t->n_segments, t->n_levels, t->max_upstream};
    int R = t->n_reaches;
    bool ok = std::fwrite(hdr, sizeof(int), 4, f) == 4;
    auto put = [&](const int *p, size_t n) { if (ok && n) ok = std::fwrite(p, sizeof(int), n, f) == n; };
    put(t->reach_seg_start, R); put(t->reach_seg_len, R); put(t->reach_level, R);
    put(t->reach_n_up, R); put(t->reach_up_ptr, (size_t)R + 1);
    put(t->reach_up_idx, (size_t)t->reach_up_ptr[R]);
    put(t->level_ptr, (size_t)t->n_levels + 1); put(t->level_reach, R);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? TRMC_OK : TRMC_ERR_IO;
}

// Statistics of wavefront_mc.cu:822-858 (pct :567-572 takes the element at
// floor(p/100 * (n-1)) of the sorted errors).
int compute_stats(const double *pred, const double *truth, int n, double *out)
{
    if (n < 0 || !out || (n > 0 && (!pred || !truth))) return TRMC_ERR_ARG;
    std::vector<double> rel;
    rel.reserve((size_t)n);
    double max_abs = 0.0;
    int nans = 0;
    for (int i = 0; i < n; ++i) {
        double g = pred[i], t = truth[i];
        if (std::isnan(g) || std::isinf(g)) { ++nans; continue; }
        double ae = std::fabs(g - t);
        max_abs = std::fmax(max_abs, ae);
        if (std::fabs(t) > 1e-3) rel.push_back(ae / std::fabs(t));
    }
    std::sort(rel.begin(), rel.end());
    auto pct = [&](double p) {
        if (rel.empty()) return 0.0;
        return rel[(size_t)((p / 100.0) * (double)(rel.size() - 1))];
    };
    size_t w1 = 0, w10 = 0;
    for (double r : rel) { w1 += r < 0.01; w10 += r < 0.1; }
    out[0] = max_abs;
    out[1] = pct(50.0); out[2] = pct(90.0); out[3] = pct(99.0); out[4] = pct(99.9);
    out[5] = rel.empty() ? 0.0 : rel.back();
    out[6] = rel.empty() ? 1.0 : (double)w1 / (double)rel.size();
    out[7] = rel.empty() ? 1.0 : (double)w10 / (double)rel.size();
    out[8] = (double)nans;
    return TRMC_OK;
}

}  // namespace trmc
