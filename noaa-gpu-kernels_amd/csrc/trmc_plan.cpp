// Schedule planning: validation, unit cutting, ALAP levels, device ordering.
#include "trmc_plan.h"

#include <algorithm>
#include <cstdio>

#include "trmc.h"

namespace trmc {

static int fail(std::string &err, int code, const char *fmt, long long a = 0, long long b = 0)
{
    char buf[256];
    std::snprintf(buf, sizeof(buf), fmt, a, b);
    err = buf;
    return code;
}

int build_plan(const int *reach_seg_start, const int *reach_seg_len,
               const int *reach_up_ptr, const int *reach_up_idx,
               const int *level_ptr, const int *level_reach,
               int n_reaches, int n_levels, int n_segments, int unit_segments,
               Plan &P, std::string &err)
{
    P = Plan();
    if (n_reaches < 0 || n_levels < 0 || n_segments < 0)
        return fail(err, TRMC_ERR_ARG, "negative size");
    if (unit_segments <= 0) unit_segments = 1;
    P.n_reaches = n_reaches;
    P.n_segments = n_segments;
    P.unit_segments = unit_segments;
    P.dev_of_seg.assign((size_t)n_segments, -1);
    if (n_reaches == 0) {
        P.level_ptr.assign(1, 0);
        P.unit_seg.assign(1, 0);
        P.up_ptr.assign(1, 0);
        P.n_levels = 0;
        return 0;
    }
    if (!reach_seg_start || !reach_seg_len || !reach_up_ptr || !level_ptr || !level_reach)
        return fail(err, TRMC_ERR_ARG, "null topology array");

    // --- level structure: every reach exactly once, in an earlier level than
    //     everything downstream of it (the reference's sweep order).
    if (level_ptr[0] != 0 || level_ptr[n_levels] != n_reaches)
        return fail(err, TRMC_ERR_TOPOLOGY, "level_ptr must span [0, n_reaches] (got end %lld)",
                    (long long)level_ptr[n_levels]);
    std::vector<int> level_of(n_reaches, -1);
    for (int L = 0; L < n_levels; ++L) {
        if (level_ptr[L + 1] < level_ptr[L])
            return fail(err, TRMC_ERR_TOPOLOGY, "level_ptr decreases at level %lld", L);
        for (int i = level_ptr[L]; i < level_ptr[L + 1]; ++i) {
            int r = level_reach[i];
            if (r < 0 || r >= n_reaches)
                return fail(err, TRMC_ERR_TOPOLOGY, "level_reach[%lld] = %lld out of range", i, r);
            if (level_of[r] >= 0)
                return fail(err, TRMC_ERR_TOPOLOGY, "reach %lld listed twice in level_reach", r);
            level_of[r] = L;
        }
    }
    if (reach_up_ptr[0] != 0)
        return fail(err, TRMC_ERR_TOPOLOGY, "reach_up_ptr[0] != 0");
    int nnz = reach_up_ptr[n_reaches];
    if (nnz > 0 && !reach_up_idx) return fail(err, TRMC_ERR_ARG, "null reach_up_idx");
    for (int r = 0; r < n_reaches; ++r) {
        if (reach_up_ptr[r + 1] < reach_up_ptr[r])
            return fail(err, TRMC_ERR_TOPOLOGY, "reach_up_ptr decreases at reach %lld", r);
        P.max_upstream = std::max(P.max_upstream, reach_up_ptr[r + 1] - reach_up_ptr[r]);
        for (int k = reach_up_ptr[r]; k < reach_up_ptr[r + 1]; ++k) {
            int u = reach_up_idx[k];
            if (u < 0 || u >= n_reaches)
                return fail(err, TRMC_ERR_TOPOLOGY, "reach_up_idx[%lld] = %lld out of range", k, u);
            if (level_of[u] >= level_of[r])
                return fail(err, TRMC_ERR_TOPOLOGY,
                            "upstream reach %lld is not in an earlier level than reach %lld", u, r);
        }
    }
    // --- segments: disjoint ranges inside [0, n_segments)
    std::vector<unsigned char> seen((size_t)n_segments, 0);
    for (int r = 0; r < n_reaches; ++r) {
        long long s = reach_seg_start[r], n = reach_seg_len[r];
        if (n < 0 || (n > 0 && (s < 0 || s + n > n_segments)))
            return fail(err, TRMC_ERR_TOPOLOGY, "reach %lld segment range out of bounds", r);
        for (long long i = s; i < s + n; ++i) {
            if (seen[(size_t)i])
                return fail(err, TRMC_ERR_TOPOLOGY, "segment %lld belongs to two reaches", i);
            seen[(size_t)i] = 1;
        }
    }

    // --- cut reaches into units, in the reference's topological sweep order
    std::vector<int> first_unit(n_reaches), last_unit(n_reaches);
    std::vector<int> useg_begin, useg_len;          // original segment ranges
    std::vector<int> uptr(1, 0), uidx, uflags, ureach, uinit;
    useg_begin.reserve((size_t)n_segments + n_reaches);
    for (int ii = 0; ii < n_reaches; ++ii) {
        int r = level_reach[ii];
        int S = reach_seg_len[r];
        int nch = std::max(1, (S + unit_segments - 1) / unit_segments);
        int base = S / nch, rem = S % nch, off = reach_seg_start[r];
        first_unit[r] = (int)useg_begin.size();
        for (int j = 0; j < nch; ++j) {
            int len = base + (j < rem ? 1 : 0);
            int u = (int)useg_begin.size();
            useg_begin.push_back(off);
            useg_len.push_back(len);
            if (j == 0) {
                for (int k = reach_up_ptr[r]; k < reach_up_ptr[r + 1]; ++k)
                    uidx.push_back(last_unit[reach_up_idx[k]]);
                uflags.push_back(0);
            } else {
                uidx.push_back(u - 1);
                uflags.push_back(1);
            }
            uptr.push_back((int)uidx.size());
            bool last = (j == nch - 1);
            ureach.push_back(last ? r : -1);
            // outflow at t = -1: reach_last0 for the reach's last unit, else the
            // initial qdc of the unit's last segment (qdp0), wavefront_mc.cu:676-680
            uinit.push_back(last ? -(r + 1) : off + len - 1);
            off += len;
        }
        last_unit[r] = (int)useg_begin.size() - 1;
    }
    const int U = (int)useg_begin.size();

    // --- ASAP (longest path) then ALAP levels on the unit DAG
    std::vector<int> asap(U, 0);
    int Lu = 0;
    for (int u = 0; u < U; ++u) {
        int a = 0;
        for (int k = uptr[u]; k < uptr[u + 1]; ++k) a = std::max(a, asap[uidx[k]] + 1);
        asap[u] = a;
        Lu = std::max(Lu, a + 1);
    }
    std::vector<int> alap(U, Lu - 1);
    // units are in topological order, so a reverse sweep sees every
    // downstream unit before its upstreams
    for (int u = U - 1; u >= 0; --u)
        for (int k = uptr[u]; k < uptr[u + 1]; ++k) {
            int v = uidx[k];
            alap[v] = std::min(alap[v], alap[u] - 1);
        }
    int max_span = 1;
    for (int u = 0; u < U; ++u)
        for (int k = uptr[u]; k < uptr[u + 1]; ++k)
            max_span = std::max(max_span, alap[u] - alap[uidx[k]]);

    // --- device order: stable counting sort by ALAP level
    P.n_units = U;
    P.n_levels = Lu;
    P.max_span = max_span;
    P.ring = max_span + 2;
    P.level_ptr.assign((size_t)Lu + 1, 0);
    for (int u = 0; u < U; ++u) P.level_ptr[(size_t)alap[u] + 1]++;
    for (int L = 0; L < Lu; ++L) P.level_ptr[(size_t)L + 1] += P.level_ptr[L];
    std::vector<int> dev_of_unit(U), fill(P.level_ptr.begin(), P.level_ptr.end() - 1);
    for (int u = 0; u < U; ++u) dev_of_unit[u] = fill[(size_t)alap[u]]++;
    std::vector<int> unit_of_dev(U);
    for (int u = 0; u < U; ++u) unit_of_dev[dev_of_unit[u]] = u;

    P.unit_level.resize(U);
    P.unit_seg.assign((size_t)U + 1, 0);
    P.up_ptr.assign((size_t)U + 1, 0);
    P.up_idx.clear();
    P.up_idx.reserve(uidx.size());
    P.unit_flags.resize(U);
    P.unit_reach.resize(U);
    P.unit_init.resize(U);
    P.seg_of_dev.clear();
    P.seg_of_dev.reserve((size_t)n_segments);
    for (int du = 0; du < U; ++du) {
        int u = unit_of_dev[du];
        P.unit_level[du] = alap[u];
        for (int k = uptr[u]; k < uptr[u + 1]; ++k) P.up_idx.push_back(dev_of_unit[uidx[k]]);
        P.up_ptr[(size_t)du + 1] = (int)P.up_idx.size();
        P.unit_flags[du] = uflags[u];
        P.unit_reach[du] = ureach[u];
        P.unit_init[du] = uinit[u];
        for (int i = 0; i < useg_len[u]; ++i) {
            int s = useg_begin[u] + i;
            P.dev_of_seg[(size_t)s] = (int)P.seg_of_dev.size();
            P.seg_of_dev.push_back(s);
        }
        P.unit_seg[(size_t)du + 1] = (int)P.seg_of_dev.size();
    }
    P.n_dev_segments = (int)P.seg_of_dev.size();
    if (P.max_upstream <= 4) {
        P.unit_rec.assign((size_t)U * 8, 0);
        for (int du = 0; du < U; ++du) {
            int *r = &P.unit_rec[(size_t)du * 8];
            const int nu = P.up_ptr[(size_t)du + 1] - P.up_ptr[du];
            r[0] = P.unit_level[du];
            r[1] = P.unit_seg[du];
            r[2] = P.unit_seg[(size_t)du + 1];
            // bit 1: the unit lies on a longest dependency chain (ASAP level ==
            // ALAP level, no slack); the dataflow kernel raises the issue
            // priority of a wave while it solves such a unit
            const int u = unit_of_dev[du];
            const bool crit = alap[u] == asap[u];
            r[3] = (P.unit_flags[du] & 0xFF) | (crit ? 2 : 0) | (nu << 8);
            for (int q = 0; q < nu; ++q) r[4 + q] = P.up_idx[(size_t)P.up_ptr[du] + q];
        }
    }
    P.reach_last_unit.resize(n_reaches);
    for (int r = 0; r < n_reaches; ++r) P.reach_last_unit[r] = dev_of_unit[last_unit[r]];
    return 0;
}

int64_t max_window_items(const Plan &p, int n_steps, int n_members)
{
    if (p.n_units == 0 || n_steps <= 0) return 0;
    int64_t best = 0;
    const int L = p.n_levels;
    for (int k = 0; k < L + n_steps - 1; ++k) {
        int lo = std::max(0, k - n_steps + 1), hi = std::min(k, L - 1);
        if (lo > hi) continue;
        int64_t w = (int64_t)p.level_ptr[(size_t)hi + 1] - p.level_ptr[(size_t)lo];
        best = std::max(best, w);
    }
    return best * std::max(1, n_members);
}

}  // namespace trmc
