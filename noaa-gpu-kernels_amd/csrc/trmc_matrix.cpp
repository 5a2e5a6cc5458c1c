// Propagation matrices of the matrix-form linearized MC, after
// wavefront/linear_mc_matrix.cu:118-229 (build_propagation_matrices): for each
// reach r in index order (every upstream reach has a smaller index)
//   row_Q(r) = prune(C1[r] * sum_u row_Q(u)) + C2[r] at each upstream column
//              + C3[r] at column r, pruned again
//   row_q(r) = prune(C1[r] * sum_u row_q(u)) + C4/ql[r] at column r, pruned again
// with prune dropping |value| < tol.  The upstream rows are merged one after
// another, in CSR order, into a sorted working row (a column new to it takes
// the product, a present one adds it), as the reference does, but through two
// scratch rows reused for every merge and written once into the output, so
// no allocation happens per merge.  One-time host work.
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "trmc.h"
#include "trmc_util.h"

namespace {

struct Rows {
    std::vector<int> rowptr, col;
    std::vector<float> val;
};

// a sorted working row
struct Work {
    std::vector<int> col, tcol;
    std::vector<float> val, tval;
    void clear() { col.clear(); val.clear(); }
    // this = merge(this, scale * row k of m)
    void merge_scaled(const Rows &m, int k, float scale)
    {
        const int b = m.rowptr[k], e = m.rowptr[k + 1];
        if (col.empty()) {
            for (int j = b; j < e; ++j) {
                col.push_back(m.col[j]);
                val.push_back(m.val[j] * scale);
            }
            return;
        }
        tcol.clear();
        tval.clear();
        size_t i = 0;
        int j = b;
        while (i < col.size() && j < e) {
            if (col[i] < m.col[j]) {
                tcol.push_back(col[i]); tval.push_back(val[i]); ++i;
            } else if (col[i] > m.col[j]) {
                tcol.push_back(m.col[j]); tval.push_back(m.val[j] * scale); ++j;
            } else {
                tcol.push_back(col[i]); tval.push_back(val[i] + m.val[j] * scale); ++i; ++j;
            }
        }
        for (; i < col.size(); ++i) { tcol.push_back(col[i]); tval.push_back(val[i]); }
        for (; j < e; ++j) { tcol.push_back(m.col[j]); tval.push_back(m.val[j] * scale); }
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
    bool emit(Rows &m) const
    {
        if (m.col.size() + col.size() > (size_t)INT_MAX) return false;
        m.col.insert(m.col.end(), col.begin(), col.end());
        m.val.insert(m.val.end(), val.begin(), val.end());
        m.rowptr.push_back((int)m.col.size());
        return true;
    }
};

int to_csr(const Rows &m, int n, trmc_csr *out)
{
    const size_t nnz = m.col.size();
    out->n = n;
    out->nnz = (int)nnz;
    out->rowptr = (int *)std::malloc(sizeof(int) * ((size_t)n + 1));
    out->col = (int *)std::malloc(sizeof(int) * (nnz ? nnz : 1));
    out->val = (float *)std::malloc(sizeof(float) * (nnz ? nnz : 1));
    if (!out->rowptr || !out->col || !out->val) return trmc::fail(TRMC_ERR_NOMEM, "out of host memory");
    std::memcpy(out->rowptr, m.rowptr.data(), sizeof(int) * ((size_t)n + 1));
    if (nnz) {
        std::memcpy(out->col, m.col.data(), sizeof(int) * nnz);
        std::memcpy(out->val, m.val.data(), sizeof(float) * nnz);
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
    for (int r = 0; r < n; ++r) {
        if (up_ptr[r + 1] < up_ptr[r]) return trmc::fail(TRMC_ERR_TOPOLOGY, "reach_up_ptr decreases at %d", r);
        for (int k = up_ptr[r]; k < up_ptr[r + 1]; ++k)
            if (!up_idx || up_idx[k] < 0 || up_idx[k] >= r)
                return trmc::fail(TRMC_ERR_TOPOLOGY, "reach %d: upstream %d is not an earlier reach", r,
                                  up_idx ? up_idx[k] : -1);
    }
    Rows mQ, mq;
    mQ.rowptr.reserve((size_t)n + 1);
    mq.rowptr.reserve((size_t)n + 1);
    mQ.rowptr.push_back(0);
    mq.rowptr.push_back(0);
    Work wQ, wq;
    for (int r = 0; r < n; ++r) {
        const float c1 = C1[r];
        const int ub = up_ptr[r], ue = up_ptr[r + 1];
        wQ.clear();
        wq.clear();
        for (int k = ub; k < ue; ++k) {
            wQ.merge_scaled(mQ, up_idx[k], c1);
            wq.merge_scaled(mq, up_idx[k], c1);
        }
        wQ.prune(tol);
        wq.prune(tol);
        for (int k = ub; k < ue; ++k) wQ.add(up_idx[k], C2[r]);
        wQ.add(r, C3[r]);
        wq.add(r, C4oq[r]);
        wQ.prune(tol);
        wq.prune(tol);
        if (!wQ.emit(mQ) || !wq.emit(mq))
            return trmc::fail(TRMC_ERR_NOMEM, "propagation matrix exceeds 2^31 entries (raise tol)");
    }
    int rc = to_csr(mQ, n, P_Q);
    if (!rc) rc = to_csr(mq, n, P_q);
    if (rc) {
        trmc_free_csr(P_Q);
        trmc_free_csr(P_q);
    }
    return rc;
}

}  // extern "C"
