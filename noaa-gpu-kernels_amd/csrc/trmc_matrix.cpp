// Propagation matrices of the matrix-form linearized MC, after
// wavefront/linear_mc_matrix.cu:118-229 (build_propagation_matrices): for each
// reach r (every upstream reach has a smaller index)
//   row_Q(r) = prune(C1[r] * sum_u row_Q(u)) + C2[r] at each upstream column
//              + C3[r] at column r, pruned again
//   row_q(r) = prune(C1[r] * sum_u row_q(u)) + C4/ql[r] at column r, pruned again
// with prune dropping |value| < tol.  A row's upstream rows are merged one
// after another, in CSR order, into a sorted working row (a column new to it
// takes the product, a present one adds it), as the reference does, so every
// value is the reference's float result.  A row needs only its upstream rows,
// so the rows of one level (longest upstream path) are built in parallel by
// host threads, level after level; each thread keeps two scratch rows and
// appends finished rows to chunks of its own that never move (rows of earlier
// levels are read through pointers while later ones are written), and the rows
// are copied into CSR order at the end.  One-time host work.
#include <algorithm>
#include <atomic>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "trmc.h"
#include "trmc_util.h"

namespace {

struct RowRef {
    const int *col = nullptr;
    const float *val = nullptr;
    int len = 0;
};

// append-only storage of finished rows in fixed chunks
struct Arena {
    struct Chunk {
        std::unique_ptr<int[]> col;
        std::unique_ptr<float[]> val;
        size_t cap = 0, used = 0;
    };
    std::vector<Chunk> chunks;
    RowRef put(const std::vector<int> &c, const std::vector<float> &v)
    {
        const size_t len = c.size();
        if (chunks.empty() || chunks.back().cap - chunks.back().used < len) {
            Chunk k;
            k.cap = std::max<size_t>(len, (size_t)1 << 18);
            k.col.reset(new int[k.cap]);
            k.val.reset(new float[k.cap]);
            chunks.push_back(std::move(k));
        }
        Chunk &k = chunks.back();
        std::copy(c.begin(), c.end(), k.col.get() + k.used);
        std::copy(v.begin(), v.end(), k.val.get() + k.used);
        RowRef r{k.col.get() + k.used, k.val.get() + k.used, (int)len};
        k.used += len;
        return r;
    }
};

// a sorted working row
struct Work {
    std::vector<int> col, tcol;
    std::vector<float> val, tval;
    void clear() { col.clear(); val.clear(); }
    // this = merge(this, scale * row)
    void merge_scaled(const RowRef &m, float scale)
    {
        if (col.empty()) {
            for (int j = 0; j < m.len; ++j) {
                col.push_back(m.col[j]);
                val.push_back(m.val[j] * scale);
            }
            return;
        }
        tcol.clear();
        tval.clear();
        size_t i = 0;
        int j = 0;
        while (i < col.size() && j < m.len) {
            if (col[i] < m.col[j]) {
                tcol.push_back(col[i]); tval.push_back(val[i]); ++i;
            } else if (col[i] > m.col[j]) {
                tcol.push_back(m.col[j]); tval.push_back(m.val[j] * scale); ++j;
            } else {
                tcol.push_back(col[i]); tval.push_back(val[i] + m.val[j] * scale); ++i; ++j;
            }
        }
        for (; i < col.size(); ++i) { tcol.push_back(col[i]); tval.push_back(val[i]); }
        for (; j < m.len; ++j) { tcol.push_back(m.col[j]); tval.push_back(m.val[j] * scale); }
        col.swap(tcol);
        val.swap(tval);
    }
    // value v onto column c (inserted in order when absent)
    void add(int c, float v)
    {
        auto it = std::lower_bound(col.begin(), col.end(), c);
        const size_t at = (size_t)(it - col.begin());
        if (it != col.end() && *it == c) {
            val[at] = val[at] + v;
        } else {
            col.insert(it, c);
            val.insert(val.begin() + (std::ptrdiff_t)at, v);
        }
    }
    // drop |value| < tol (NaN stays, as remove_if over fabs(v) < tol keeps it)
    void prune(float tol)
    {
        size_t k = 0;
        for (size_t i = 0; i < col.size(); ++i)
            if (!(std::fabs(val[i]) < tol)) { col[k] = col[i]; val[k] = val[i]; ++k; }
        col.resize(k);
        val.resize(k);
    }
};

// spinning barrier for a fixed set of threads
class Barrier {
public:
    explicit Barrier(int n) : n_(n) {}
    void set_count(int n) { n_ = n; }   // before any thread waits
    void wait()
    {
        const unsigned gen = gen_.load(std::memory_order_acquire);
        if (count_.fetch_add(1, std::memory_order_acq_rel) + 1 == n_) {
            count_.store(0, std::memory_order_relaxed);
            gen_.fetch_add(1, std::memory_order_acq_rel);
        } else {
            while (gen_.load(std::memory_order_acquire) == gen) std::this_thread::yield();
        }
    }

private:
    int n_;
    std::atomic<int> count_{0};
    std::atomic<unsigned> gen_{0};
};

int to_csr(const std::vector<RowRef> &rows, int n, trmc_csr *out)
{
    size_t nnz = 0;
    for (int r = 0; r < n; ++r) nnz += (size_t)rows[r].len;
    if (nnz > (size_t)INT_MAX) return trmc::fail(TRMC_ERR_NOMEM, "propagation matrix exceeds 2^31 entries (raise tol)");
    out->n = n;
    out->nnz = (int)nnz;
    out->rowptr = (int *)std::malloc(sizeof(int) * ((size_t)n + 1));
    out->col = (int *)std::malloc(sizeof(int) * (nnz ? nnz : 1));
    out->val = (float *)std::malloc(sizeof(float) * (nnz ? nnz : 1));
    if (!out->rowptr || !out->col || !out->val) return trmc::fail(TRMC_ERR_NOMEM, "out of host memory");
    size_t p = 0;
    out->rowptr[0] = 0;
    for (int r = 0; r < n; ++r) {
        std::memcpy(out->col + p, rows[r].col, sizeof(int) * (size_t)rows[r].len);
        std::memcpy(out->val + p, rows[r].val, sizeof(float) * (size_t)rows[r].len);
        p += (size_t)rows[r].len;
        out->rowptr[r + 1] = (int)p;
    }
    return TRMC_OK;
}

}  // namespace

extern "C" {

void trmc_free_csr(trmc_csr *m)
{
    if (!m) return;
    std::free(m->rowptr);
    std::free(m->col);
    std::free(m->val);
    std::memset(m, 0, sizeof(*m));
}

int trmc_matrix_build(int n, const int *up_ptr, const int *up_idx, const float *C1, const float *C2,
                      const float *C3, const float *C4oq, float tol, trmc_csr *P_Q, trmc_csr *P_q)
{
    if (!P_Q || !P_q) return trmc::fail(TRMC_ERR_ARG, "null output");
    std::memset(P_Q, 0, sizeof(*P_Q));
    std::memset(P_q, 0, sizeof(*P_q));
    if (n < 0) return trmc::fail(TRMC_ERR_ARG, "negative n_reaches");
    if (!up_ptr || (n > 0 && (!C1 || !C2 || !C3 || !C4oq))) return trmc::fail(TRMC_ERR_ARG, "null array");
    if (up_ptr[0] != 0) return trmc::fail(TRMC_ERR_TOPOLOGY, "reach_up_ptr[0] = %d", up_ptr[0]);
    // levels: 0 for a headwater, else 1 + the highest upstream level
    std::vector<int> level((size_t)n, 0);
    int n_levels = n > 0 ? 1 : 0;
    for (int r = 0; r < n; ++r) {
        if (up_ptr[r + 1] < up_ptr[r]) return trmc::fail(TRMC_ERR_TOPOLOGY, "reach_up_ptr decreases at %d", r);
        for (int k = up_ptr[r]; k < up_ptr[r + 1]; ++k) {
            if (!up_idx || up_idx[k] < 0 || up_idx[k] >= r)
                return trmc::fail(TRMC_ERR_TOPOLOGY, "reach %d: upstream %d is not an earlier reach", r,
                                  up_idx ? up_idx[k] : -1);
            level[r] = std::max(level[r], level[up_idx[k]] + 1);
        }
        n_levels = std::max(n_levels, level[r] + 1);
    }
    std::vector<int> lptr((size_t)n_levels + 1, 0), lrows((size_t)n);
    for (int r = 0; r < n; ++r) lptr[level[r] + 1]++;
    for (int L = 0; L < n_levels; ++L) lptr[L + 1] += lptr[L];
    {
        std::vector<int> fill(lptr.begin(), lptr.end() - 1);
        for (int r = 0; r < n; ++r) lrows[fill[level[r]]++] = r;
    }
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<unsigned>(std::min(hw, 16u), (unsigned)std::max(1, n / 20000));
    std::vector<RowRef> rQ((size_t)n), rq((size_t)n);
    std::vector<Arena> aQ((size_t)nt), aq((size_t)nt);
    std::atomic<int> next{0};
    std::atomic<bool> failed{false};
    Barrier bar(nt);
    // a thread whose allocation fails stops building rows but still passes
    // every barrier, so the others finish; the call then reports the failure
    std::atomic<bool> go{false};
    auto worker = [&](int tid) {
        while (!go.load(std::memory_order_acquire)) std::this_thread::yield();
        Work wQ, wq;
        for (int L = 0; L < n_levels; ++L) {
            const int b = lptr[L], e = lptr[L + 1];
            for (;;) {
                const int i0 = b + next.fetch_add(16, std::memory_order_relaxed);
                if (i0 >= e) break;
                if (failed.load(std::memory_order_relaxed)) continue;
                try {
                for (int i = i0; i < std::min(i0 + 16, e); ++i) {
                    const int r = lrows[i];
                    const float c1 = C1[r];
                    const int ub = up_ptr[r], ue = up_ptr[r + 1];
                    wQ.clear();
                    wq.clear();
                    for (int k = ub; k < ue; ++k) {
                        wQ.merge_scaled(rQ[up_idx[k]], c1);
                        wq.merge_scaled(rq[up_idx[k]], c1);
                    }
                    wQ.prune(tol);
                    wq.prune(tol);
                    for (int k = ub; k < ue; ++k) wQ.add(up_idx[k], C2[r]);
                    wQ.add(r, C3[r]);
                    wq.add(r, C4oq[r]);
                    wQ.prune(tol);
                    wq.prune(tol);
                    rQ[r] = aQ[tid].put(wQ.col, wQ.val);
                    rq[r] = aq[tid].put(wq.col, wq.val);
                }
                } catch (...) {
                    failed.store(true, std::memory_order_relaxed);
                }
            }
            bar.wait();                       // every row of level L is written
            if (tid == 0) next.store(0, std::memory_order_relaxed);
            bar.wait();                       // the counter is reset for level L + 1
        }
    };
    {
        // the barrier counts the threads that did start
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) {
            try {
                pool.emplace_back(worker, t);
            } catch (...) {
                break;
            }
        }
        bar.set_count((int)pool.size() + 1);
        go.store(true, std::memory_order_release);
        worker(0);
        for (auto &th : pool) th.join();
    }
    if (failed.load()) return trmc::fail(TRMC_ERR_NOMEM, "out of host memory building the matrices");
    int rc = to_csr(rQ, n, P_Q);
    if (!rc) rc = to_csr(rq, n, P_q);
    if (rc) {
        trmc_free_csr(P_Q);
        trmc_free_csr(P_q);
    }
    return rc;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Secant-fitted coefficients of the matrix form, after
// wavefront/linear_mc_matrix_secant.cu: per reach, the float MC secant on
// depth at a steady reference flow (mc_secant_solve_cpu, :57-195), its last
// iterate's C1..C3 and C4/ql; the reference flows follow the reach DAG in
// level order (main, :358-418).  Float operations in the reference's order
// (the coefficients are the matrices' input and are checked bit for bit);
// host threads over reaches, one-time set-up.
namespace {

struct SecGeo {
    float twl, R, A, AC, WP, WPC;
};

SecGeo sec_geometry(float h, float bfd, float bw, float twcc, float z)
{
    SecGeo g;
    g.twl = bw + 2.0f * z * h;
    float hgb = std::fmax(h - bfd, 0.0f), hlb = std::fmin(bfd, h);
    if (hgb > 0.0f && twcc <= 0.0f) {
        hgb = 0.0f;
        hlb = h;
    }
    g.A = (bw + hlb * z) * hlb;
    g.WP = bw + 2.0f * hlb * std::sqrt(1.0f + z * z);
    g.AC = twcc * hgb;
    g.WPC = (hgb > 0.0f) ? (twcc + 2.0f * hgb) : 0.0f;
    g.R = (g.WP + g.WPC > 0.0f) ? (g.A + g.AC) / (g.WP + g.WPC) : 0.0f;
    return g;
}

float sec_celerity(float h, const SecGeo &g, const trmc_chparams_f32 &p, float z)
{
    if (!(h > 0.0f)) return 0.0f;
    const float r23 = std::pow(g.R, 2.0f / 3.0f), r53 = std::pow(g.R, 5.0f / 3.0f);
    return std::fmax(0.0f, (std::sqrt(p.s0) / p.n_ch) *
                               ((5.0f / 3.0f) * r23 -
                                (2.0f / 3.0f) * r53 * (2.0f * std::sqrt(1.0f + z * z) / (p.bw + 2.0f * h * z))));
}

float sec_manning(const SecGeo &g, const trmc_chparams_f32 &p)
{
    if (!(g.WP + g.WPC > 0.0f)) return 0.0f;
    return (1.0f / (((g.WP * p.n_ch) + (g.WPC * p.ncc)) / (g.WP + g.WPC))) * (g.A + g.AC) *
           std::pow(g.R, 2.0f / 3.0f) * std::sqrt(p.s0);
}

// half-bankfull coefficients with X = 0.25 (the reference's fallback)
void sec_fallback(const trmc_chparams_f32 &p, float dt, float z, float bfd, float c[4])
{
    const float h = std::fmax(0.5f * bfd, 0.1f);
    const float A = (p.bw + h * z) * h;
    const float WP = p.bw + 2.0f * h * std::sqrt(1.0f + z * z);
    const float R = (WP > 0.0f) ? A / WP : 0.0f;
    const float r23 = std::pow(R, 2.0f / 3.0f), r53 = std::pow(R, 5.0f / 3.0f);
    const float Ck = std::fmax(1e-6f, (std::sqrt(p.s0) / p.n_ch) *
                                          ((5.0f / 3.0f) * r23 - (2.0f / 3.0f) * r53 *
                                                                     (2.0f * std::sqrt(1.0f + z * z) /
                                                                      (p.bw + 2.0f * h * z))));
    const float Km = std::fmax(dt, p.dx / Ck);
    const float X = 0.25f;
    float D = Km * (1.0f - X) + dt / 2.0f;
    if (D == 0.0f) D = 1.0f;
    c[0] = (Km * X + dt / 2.0f) / D;
    c[1] = (dt / 2.0f - Km * X) / D;
    c[2] = (Km * (1.0f - X) - dt / 2.0f) / D;
    c[3] = dt / D;
}

// one reach: c = {C1, C2, C3, C4/ql}
void sec_solve(float dt, float qup, float quc, float qdp, float ql, const trmc_chparams_f32 &p, float depthp,
               float c[4])
{
    const float z = (p.cs == 0.0f) ? 1.0f : 1.0f / p.cs;
    float bfd;
    if (p.bw > p.tw) bfd = p.bw / 0.00001f;
    else if (p.bw == p.tw) bfd = p.bw / (2.0f * z);
    else bfd = (p.tw - p.bw) / (2.0f * z);
    if (p.n_ch <= 0.0f || p.s0 <= 0.0f || z <= 0.0f || p.bw <= 0.0f ||
        (ql <= 0.0f && qup <= 0.0f && quc <= 0.0f && qdp <= 0.0f)) {
        sec_fallback(p, dt, z, bfd, c);
        return;
    }
    const float mindepth = 0.01f;
    const float depthc = std::fmax(depthp, 0.0f);
    float h = depthc * 1.33f + mindepth, h_0 = depthc * 0.67f;
    float C1 = 0, C2 = 0, C3 = 0, C4 = 0, X = 0.25f, Qj_0 = 0.0f, Qj = 0.0f;
    int maxiter = 100;
    for (int att = 0; att < 5; att++) {
        float aerror = 0.01f, rerror = 1.0f;
        int iter = 0;
        while (rerror > 0.01f && aerror >= mindepth && iter <= maxiter) {
            // first point h_0: its X uses the previous iteration's Qj_0
            const SecGeo g0 = sec_geometry(h_0, bfd, p.bw, p.twcc, z);
            const float Ck0 = sec_celerity(h_0, g0, p, z);
            const float Km0 = (Ck0 > 0.0f) ? std::fmax(dt, p.dx / Ck0) : dt;
            const float twu0 = (h_0 > bfd && p.twcc > 0.0f) ? p.twcc : g0.twl;
            const float X0 = (Ck0 > 0.0f && twu0 * p.s0 * Ck0 * p.dx > 0.0f)
                                 ? std::fmin(0.5f, std::fmax(0.0f, 0.5f * (1.0f - Qj_0 / (2.0f * twu0 * p.s0 * Ck0 *
                                                                                          p.dx))))
                                 : 0.5f;
            float D0 = Km0 * (1.0f - X0) + dt / 2.0f;
            if (D0 == 0.0f) D0 = 1.0f;
            const float c1_0 = (Km0 * X0 + dt / 2.0f) / D0;
            const float c2_0 = (dt / 2.0f - Km0 * X0) / D0;
            const float c3_0 = (Km0 * (1.0f - X0) - dt / 2.0f) / D0;
            const float c4_0 = (ql * dt) / D0;
            const float q0 = c1_0 * qup + c2_0 * quc + c3_0 * qdp + c4_0;
            Qj_0 = q0 - sec_manning(g0, p);
            // second point h
            const SecGeo g1 = sec_geometry(h, bfd, p.bw, p.twcc, z);
            const float Ck1 = sec_celerity(h, g1, p, z);
            const float Km1 = (Ck1 > 0.0f) ? std::fmax(dt, p.dx / Ck1) : dt;
            const float twu1 = (h > bfd && p.twcc > 0.0f) ? p.twcc : g1.twl;
            X = (Ck1 > 0.0f && twu1 * p.s0 * Ck1 * p.dx > 0.0f)
                    ? std::fmin(0.5f, std::fmax(0.25f, 0.5f * (1.0f - q0 / (2.0f * twu1 * p.s0 * Ck1 * p.dx))))
                    : 0.5f;
            float D1 = Km1 * (1.0f - X) + dt / 2.0f;
            if (D1 == 0.0f) D1 = 1.0f;
            C1 = (Km1 * X + dt / 2.0f) / D1;
            C2 = (dt / 2.0f - Km1 * X) / D1;
            C3 = (Km1 * (1.0f - X) - dt / 2.0f) / D1;
            C4 = (ql * dt) / D1;
            if (C4 < 0.0f && std::fabs(C4) > C1 * qup + C2 * quc + C3 * qdp) C4 = -(C1 * qup + C2 * quc + C3 * qdp);
            Qj = (C1 * qup + C2 * quc + C3 * qdp + C4) - sec_manning(g1, p);
            float h_1 = (Qj_0 - Qj != 0.0f) ? h - (Qj * (h_0 - h)) / (Qj_0 - Qj) : h;
            if (h_1 < 0.0f) h_1 = h;
            if (h > 0.0f) {
                rerror = std::fabs((h_1 - h) / h);
                aerror = std::fabs(h_1 - h);
            } else {
                rerror = 0.0f;
                aerror = 0.9f;
            }
            h_0 = std::fmax(0.0f, h);
            h = std::fmax(0.0f, h_1);
            iter++;
            if (h < mindepth) break;
        }
        if (iter < maxiter) break;
        h *= 1.33f;
        h_0 *= 0.67f;
        maxiter += 25;
    }
    c[0] = C1;
    c[1] = C2;
    c[2] = C3;
    if (std::fabs(ql) > 1e-12f) {
        c[3] = C4 / ql;
    } else {
        float fb[4];
        sec_fallback(p, dt, z, bfd, fb);
        c[3] = fb[3];
    }
}

}  // namespace

extern "C" int trmc_matrix_secant_coeffs_f32(const trmc_chparams_f32 *params, int n_reaches, const int *up_ptr,
                                             const int *up_idx, const int *level_ptr, const int *level_reach,
                                             int n_levels, const float *qlat0, int n_segments, float dt, float *C1,
                                             float *C2, float *C3, float *C4oq)
{
    if (n_reaches < 0 || n_levels < 0 || n_segments < 0) return trmc::fail(TRMC_ERR_ARG, "negative size");
    if (n_reaches == 0) return TRMC_OK;
    if (!params || !up_ptr || !level_ptr || !level_reach || !qlat0 || !C1 || !C2 || !C3 || !C4oq)
        return trmc::fail(TRMC_ERR_ARG, "null array");
    if (n_segments < n_reaches) return trmc::fail(TRMC_ERR_ARG, "qlat0 needs a value per reach (n_segments >= n_reaches)");
    if (level_ptr[0] != 0 || level_ptr[n_levels] != n_reaches)
        return trmc::fail(TRMC_ERR_TOPOLOGY, "level_ptr must run from 0 to n_reaches");
    // every reach once, its upstream reaches in earlier levels
    std::vector<int> lev((size_t)n_reaches, -1);
    for (int L = 0; L < n_levels; ++L) {
        if (level_ptr[L + 1] < level_ptr[L]) return trmc::fail(TRMC_ERR_TOPOLOGY, "level_ptr decreases at %d", L);
        for (int i = level_ptr[L]; i < level_ptr[L + 1]; ++i) {
            const int r = level_reach[i];
            if (r < 0 || r >= n_reaches || lev[r] >= 0) return trmc::fail(TRMC_ERR_TOPOLOGY, "level_reach entry %d", i);
            lev[r] = L;
        }
    }
    for (int r = 0; r < n_reaches; ++r)
        if (lev[r] < 0) return trmc::fail(TRMC_ERR_TOPOLOGY, "reach %d is in no level", r);
    if (up_ptr[0] != 0) return trmc::fail(TRMC_ERR_TOPOLOGY, "reach_up_ptr[0] = %d", up_ptr[0]);
    for (int r = 0; r < n_reaches; ++r) {
        if (up_ptr[r + 1] < up_ptr[r]) return trmc::fail(TRMC_ERR_TOPOLOGY, "reach_up_ptr decreases at %d", r);
        for (int k = up_ptr[r]; k < up_ptr[r + 1]; ++k)
            if (!up_idx || up_idx[k] < 0 || up_idx[k] >= n_reaches || lev[up_idx[k]] >= lev[r])
                return trmc::fail(TRMC_ERR_TOPOLOGY, "reach %d: upstream not in an earlier level", r);
    }
    // mean first-step lateral inflow over all segments, in double
    double mean = 0.0;
    for (int i = 0; i < n_segments; ++i) mean += qlat0[i];
    mean /= n_segments;
    const float fmean = (float)mean;
    // steady reference flows: upstream sum + own inflow + the mean, in level order
    std::vector<float> refq((size_t)n_reaches, 0.0f);
    for (int L = 0; L < n_levels; ++L)
        for (int i = level_ptr[L]; i < level_ptr[L + 1]; ++i) {
            const int r = level_reach[i];
            float sum_up = 0.0f;
            for (int k = up_ptr[r]; k < up_ptr[r + 1]; ++k) sum_up += refq[up_idx[k]];
            refq[r] = sum_up + qlat0[r] + fmean;
        }
    auto work = [&](int b, int e) {
        for (int r = b; r < e; ++r) {
            const trmc_chparams_f32 &p = params[r];
            const float qdp = std::fmax(refq[r], 0.1f);
            float qup = 0.0f;
            for (int k = up_ptr[r]; k < up_ptr[r + 1]; ++k) qup += refq[up_idx[k]];
            qup = std::fmax(qup, 0.1f);
            const float quc = qup;
            const float ql = std::fmax(qlat0[r], 0.01f);
            const float depthp = std::fmax(0.1f, std::pow(qdp / (p.tw + 1.0f), 0.6f));
            float c[4];
            sec_solve(dt, qup, quc, qdp, ql, p, depthp, c);
            C1[r] = c[0];
            C2[r] = c[1];
            C3[r] = c[2];
            C4oq[r] = c[3];
        }
    };
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<unsigned>(std::min(hw, 16u), (unsigned)std::max(1, n_reaches / 4096));
    const int per = (n_reaches + nt - 1) / nt;
    std::vector<std::thread> pool;
    int started_to = 0;
    for (int t = 1; t < nt; ++t) {
        const int b = t * per, e = std::min(n_reaches, b + per);
        if (b >= e) break;
        try {
            pool.emplace_back(work, b, e);
            started_to = e;
        } catch (...) {
            break;
        }
    }
    work(0, std::min(per, n_reaches));
    if (started_to < n_reaches && started_to > 0) work(started_to, n_reaches);
    else if (started_to == 0 && per < n_reaches) work(per, n_reaches);
    for (auto &th : pool) th.join();
    return TRMC_OK;
}
