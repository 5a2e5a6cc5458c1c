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
