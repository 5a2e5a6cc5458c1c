// Hydrofabric network builder: waterbody flowpaths "wb-<id>" draining into
// nexus "nex-<m>" (nexus m feeds waterbody m) and their channel attributes,
// to the reference's topology / parameter layout with one segment per reach,
// after data_parsers/parse_gpkg.py:43-270.  The caller reads the GeoPackage
// tables and passes the numbers of the rows whose id parses (a toid negative
// for NULL, unparsable or negative, which the reference treats alike) and the
// attribute values (NaN for NULL); everything from duplicate removal to the
// level-sorted CSR is done here.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "trmc.h"
#include "trmc_util.h"

namespace {

// value > 0 with a NULL (NaN) or zero failing, as `v and v > 0` does
inline bool positive(double v) { return v > 0.0; }
// a double attribute against a float parameter: numpy compares a Python float
// with a float32 scalar after rounding the former to float32
inline bool above(double v, float p) { return (float)v > p; }

}  // namespace

extern "C" {

int trmc_build_flowpath_network(const int64_t *fp_id, const int64_t *fp_toid, int n_rows,
                                const int64_t *at_link, const double *at_vals, int n_attr, trmc_topo *out,
                                trmc_chparams_f32 **params_out, int64_t **topo_to_wbid_out)
{
    if (!out || !params_out || !topo_to_wbid_out) return trmc::fail(TRMC_ERR_ARG, "null output");
    if (n_rows < 0 || n_attr < 0) return trmc::fail(TRMC_ERR_ARG, "negative size");
    if ((n_rows > 0 && (!fp_id || !fp_toid)) || (n_attr > 0 && (!at_link || !at_vals)))
        return trmc::fail(TRMC_ERR_ARG, "null input");
    std::memset(out, 0, sizeof(*out));
    *params_out = nullptr;
    *topo_to_wbid_out = nullptr;
    // waterbodies with a number, the first row of a repeated id kept (:64-94)
    std::vector<int64_t> ids, toids;
    std::unordered_map<int64_t, int> idx_of;
    idx_of.reserve((size_t)n_rows * 2);
    for (int r = 0; r < n_rows; ++r) {
        if (idx_of.count(fp_id[r])) continue;
        idx_of[fp_id[r]] = (int)ids.size();
        ids.push_back(fp_id[r]);
        toids.push_back(fp_toid[r]);
    }
    const int n = (int)ids.size();
    // upstream lists in row order; a nexus number that is no waterbody ends
    // the path (an outlet) (:99-115)
    std::vector<std::vector<int>> ups((size_t)n), downs((size_t)n);
    for (int i = 0; i < n; ++i) {
        if (toids[i] < 0) continue;
        auto d = idx_of.find(toids[i]);
        if (d == idx_of.end()) continue;
        ups[d->second].push_back(i);
    }
    for (int v = 0; v < n; ++v)
        for (int u : ups[v]) downs[u].push_back(v);
    // Kahn's order with a stack of ready waterbodies, headwaters first in index
    // order, the most recently freed one taken next (:124-145)
    std::vector<int> in_deg((size_t)n), ready, order;
    order.reserve((size_t)n);
    for (int i = 0; i < n; ++i) {
        in_deg[i] = (int)ups[i].size();
        if (in_deg[i] == 0) ready.push_back(i);
    }
    while (!ready.empty()) {
        const int u = ready.back();
        ready.pop_back();
        order.push_back(u);
        for (int d : downs[u])
            if (--in_deg[d] == 0) ready.push_back(d);
    }
    if ((int)order.size() != n)
        return trmc::fail(TRMC_ERR_TOPOLOGY, "flowpaths contain a cycle (%d of %d ordered)", (int)order.size(), n);
    std::vector<int> new_of((size_t)n);
    for (int k = 0; k < n; ++k) new_of[order[k]] = k;
    // relabelled upstream lists, sorted; level = 1 + the highest upstream's (:147-160)
    std::vector<std::vector<int>> nup((size_t)n);
    for (int old = 0; old < n; ++old) {
        auto &l = nup[new_of[old]];
        for (int u : ups[old]) l.push_back(new_of[u]);
        std::sort(l.begin(), l.end());
    }
    std::vector<int> level((size_t)n, 0);
    int n_levels = n > 0 ? 1 : 0, max_up = 0, nnz = 0;
    for (int r = 0; r < n; ++r) {
        for (int u : nup[r]) level[r] = std::max(level[r], level[u] + 1);
        n_levels = std::max(n_levels, level[r] + 1);
        max_up = std::max(max_up, (int)nup[r].size());
        nnz += (int)nup[r].size();
    }
    auto ialloc = [](size_t k) { return (int *)std::malloc(sizeof(int) * (k ? k : 1)); };
    out->n_reaches = n;
    out->n_segments = n;
    out->n_levels = n_levels;
    out->max_upstream = max_up;
    out->reach_seg_start = ialloc(n);
    out->reach_seg_len = ialloc(n);
    out->reach_level = ialloc(n);
    out->reach_n_up = ialloc(n);
    out->reach_up_ptr = ialloc((size_t)n + 1);
    out->reach_up_idx = ialloc(nnz);
    out->level_ptr = ialloc((size_t)n_levels + 1);
    out->level_reach = ialloc(n);
    auto *params = (trmc_chparams_f32 *)std::malloc(sizeof(trmc_chparams_f32) * (n ? n : 1));
    auto *wbid = (int64_t *)std::malloc(sizeof(int64_t) * (n ? n : 1));
    if (!out->reach_seg_start || !out->reach_seg_len || !out->reach_level || !out->reach_n_up ||
        !out->reach_up_ptr || !out->reach_up_idx || !out->level_ptr || !out->level_reach || !params || !wbid) {
        trmc_free_topo(out);
        std::free(params);
        std::free(wbid);
        return trmc::fail(TRMC_ERR_NOMEM, "out of host memory");
    }
    int p = 0;
    out->reach_up_ptr[0] = 0;
    for (int r = 0; r < n; ++r) {
        out->reach_seg_start[r] = r;
        out->reach_seg_len[r] = 1;
        out->reach_level[r] = level[r];
        out->reach_n_up[r] = (int)nup[r].size();
        for (int u : nup[r]) out->reach_up_idx[p++] = u;
        out->reach_up_ptr[r + 1] = p;
    }
    // level_ptr by counts; level_reach = a stable sort by level (:162-164)
    for (int L = 0; L <= n_levels; ++L) out->level_ptr[L] = 0;
    for (int r = 0; r < n; ++r) out->level_ptr[level[r] + 1]++;
    for (int L = 0; L < n_levels; ++L) out->level_ptr[L + 1] += out->level_ptr[L];
    {
        std::vector<int> fill(out->level_ptr, out->level_ptr + n_levels);
        for (int r = 0; r < n; ++r) out->level_reach[fill[level[r]]++] = r;
    }
    // channel parameters: defaults, then every attribute row of a known
    // waterbody in table order onto its reach (:188-231)
    for (int r = 0; r < n; ++r) params[r] = {1000.0f, 20.0f, 25.0f, 60.0f, 0.06f, 0.12f, 2.0f, 0.001f};
    for (int a = 0; a < n_attr; ++a) {
        auto f = idx_of.find(at_link[a]);
        if (f == idx_of.end()) continue;
        trmc_chparams_f32 &q = params[new_of[f->second]];
        const double *v = at_vals + (size_t)a * 8;   // Length_m, n, nCC, BtmWdth, TopWdth, TopWdthCC, ChSlp, So
        if (positive(v[0])) q.dx = (float)std::max(100.0, v[0]);
        if (positive(v[3])) q.bw = (float)v[3];
        if (positive(v[4]) && above(v[4], q.bw)) q.tw = (float)v[4];
        else q.tw = q.bw * 1.2f;
        if (positive(v[5]) && above(v[5], q.tw)) q.twcc = (float)v[5];
        else q.twcc = q.tw * 3.0f;
        if (positive(v[1])) q.n_ch = (float)std::max(0.01, v[1]);
        if (positive(v[2])) q.ncc = (float)std::max(0.02, v[2]);
        if (positive(v[6])) q.cs = (float)std::max(0.1, v[6]);
        if (positive(v[7])) q.s0 = (float)std::max(1e-5, v[7]);
    }
    for (int old = 0; old < n; ++old) wbid[new_of[old]] = ids[old];
    *params_out = params;
    *topo_to_wbid_out = wbid;
    return TRMC_OK;
}

}  // extern "C"
