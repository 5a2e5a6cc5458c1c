// Propagation matrices of the matrix-form linearized MC, after
// wavefront/linear_mc_matrix.cu:118-229 (build_propagation_matrices): for each
// reach r in index order (every upstream reach has a smaller index)
//   row_Q(r) = prune(C1[r] * sum_u row_Q(u)) + C2[r] at each upstream column
//              + C3[r] at column r, pruned again
//   row_q(r) = prune(C1[r] * sum_u row_q(u)) + C4/ql[r] at column r, pruned again
// with prune dropping |value| < tol.  The reference merges the upstream rows
// one after another into a sorted list; here one dense accumulator per matrix
// takes them in the same upstream order, so each column sees the same float
// operations in the same order (first product alone, later ones added), and a
// touched list restores the sorted rows.  One-time host work.
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

// one matrix's accumulator: acc[c] valid while mark[c] == current row
struct Acc {
    std::vector<float> acc;
    std::vector<int> mark, touched;
    explicit Acc(int n) : acc((size_t)n), mark((size_t)n, -1) {}
    // value v (a product or a coefficient) onto column c of row r
    void add(int r, int c, float v)
    {
        if (mark[c] == r) {
            acc[c] = acc[c] + v;
        } else {
            mark[c] = r;
            acc[c] = v;
            touched.push_back(c);
        }
    }
    // drop |value| < tol (NaN stays, as remove_if over fabs(v) < tol keeps it)
    void prune(float tol)
    {
        size_t k = 0;
        for (int c : touched) {
            if (std::fabs(acc[c]) < tol) mark[c] = -1;
            else touched[k++] = c;
        }
        touched.resize(k);
    }
    bool emit(Rows &m)
    {
        std::sort(touched.begin(), touched.end());
        if (m.col.size() + touched.size() > (size_t)INT_MAX) return false;
        for (int c : touched) {
            m.col.push_back(c);
            m.val.push_back(acc[c]);
        }
        m.rowptr.push_back((int)m.col.size());
        touched.clear();
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
    Acc aQ(n), aq(n);
    for (int r = 0; r < n; ++r) {
        const float c1 = C1[r];
        const int ub = up_ptr[r], ue = up_ptr[r + 1];
        for (int k = ub; k < ue; ++k) {
            const int u = up_idx[k];
            for (int e = mQ.rowptr[u]; e < mQ.rowptr[u + 1]; ++e) aQ.add(r, mQ.col[e], mQ.val[e] * c1);
            for (int e = mq.rowptr[u]; e < mq.rowptr[u + 1]; ++e) aq.add(r, mq.col[e], mq.val[e] * c1);
        }
        aQ.prune(tol);
        aq.prune(tol);
        for (int k = ub; k < ue; ++k) aQ.add(r, up_idx[k], C2[r]);
        aQ.add(r, r, C3[r]);
        aq.add(r, r, C4oq[r]);
        aQ.prune(tol);
        aq.prune(tol);
        if (!aQ.emit(mQ) || !aq.emit(mq))
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
