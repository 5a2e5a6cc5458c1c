// gfx950 kernels for Muskingum-Cunge routing over a river DAG, FP64.
//
// Work item = (unit, member, step).  A unit is a run of consecutive segments
// of one reach (see trmc_plan.h); its outflow for step t goes to ring slot
// (t + 1 + ring_shift) mod ring.  Segment state (qdc, depth, velocity) is
// updated in place: a unit is the only writer of its segments, and it carries
// the previous-step qdc of segment i-1 in a register for segment i's qup
// (wavefront_mc.cu:388 reads qdp_prev_seg[idx-1]).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "trmc_device.h"
#include "trmc_mlp.h"
#include "trmc_rng.h"
#include "trmc_solver.h"

namespace trmc {

struct Acc {
    unsigned long long iters, capped, fallback, noncanon;
    unsigned long long mlp_calls, mlp_ticks;   // wave-uniform
    unsigned long long trips, step_trips;      // wave-uniform
    unsigned long long step_lanes;
    unsigned long long sched_ticks;            // wave-uniform
};

// with_vel: the velocity is wanted (only the launch's last step stores one:
// no step reads the velocity of the one before, so the earlier ones would be
// overwritten unread)
template <int MODE>
__device__ __forceinline__ SolveOut solve_one(const RouteArgs &a, double qup, double quc, double qdp,
                                              double ql, const Seg &c, double dp, bool with_vel)
{
    if (MODE == 0) return solve_reference(a.dt, qup, quc, qdp, ql, c, dp, with_vel);
    if (MODE == 1) return solve_converged(a.dt, qup, quc, qdp, ql, c, dp, a.tol, a.max_eval, with_vel);
    const ConvCfg k = {a.dt, a.tol, a.max_eval, a.learned_budget};
    SecantConv s;                                // LEARNED: scalar MLP guess
    if (!s.start(qup, quc, qdp, ql, c, dp, true))
        while (!s.step(c, k))
            if (s.need_guess)
                s.set_guess(a.guess_model != 0
                                ? mlp_guess_from(s.guess_anchor(), s.guess_flow_new(), s.guess_flow_old(), 0.0f, 0.0f)
                                : mlp_guess_scalar(a.mlp, s.guess_anchor(), s.guess_flow_new(), s.guess_flow_old(), c));
    return s.finish(c, with_vel);
}

// One unit, one member, one step.  Upstream reach outflows are summed in the
// reference's CSR order starting from 0.0 (wavefront_mc.cu:523-529); an
// internal unit boundary passes its single upstream value directly.
template <int MODE>
__device__ __forceinline__ void route_item(const RouteArgs &a, int u, int m, int t, Acc &acc,
                                          unsigned *max_it = nullptr)
{
    const int M = a.n_members;
    const size_t UM = (size_t)a.n_units * (size_t)M;
    const int D = a.ring;
    const double *prev = a.ring_buf + (size_t)((t + a.ring_shift) % D) * UM;    // H(t-1)
    double *cur = a.ring_buf + (size_t)((t + 1 + a.ring_shift) % D) * UM;      // H(t)
    const int b = a.up_ptr[u], e = a.up_ptr[u + 1];
    double qup_in, quc_in;
    if (a.unit_flags[u] & 1) {
        const size_t v = (size_t)a.up_idx[b] * M + m;
        qup_in = prev[v];
        quc_in = cur[v];
    } else {
        qup_in = 0.0;
        quc_in = 0.0;
        for (int k = b; k < e; ++k) {
            const size_t v = (size_t)a.up_idx[k] * M + m;
            qup_in += prev[v];
            quc_in += cur[v];
        }
    }
    const int s0 = a.unit_seg[u], s1 = a.unit_seg[u + 1];
    const double *ql_row = a.qlat + (size_t)(a.t_base + t) * (size_t)a.n_dev_segs * M;
    const bool last_step = t == a.n_steps - 1;
    double last = 0.0, qup_next = 0.0, quc_next = 0.0;
    for (int s = s0; s < s1; ++s) {
        const size_t sm = (size_t)s * M + m;
        const double qup = (s == s0) ? qup_in : qup_next;
        const double quc = (s == s0) ? quc_in : quc_next;
        const double qdp = a.q[sm];
        const double dp = a.d[sm];
        const double ql = ql_row[sm];
        const Seg c = load_seg(a.segc, (size_t)a.n_dev_segs, s);
        const SolveOut o = solve_one<MODE>(a, qup, quc, qdp, ql, c, dp, last_step);
        a.q[sm] = o.q;
        a.d[sm] = o.depth;
        if (last_step) a.v[sm] = o.vel;
        qup_next = qdp;      // previous-step qdc of this segment feeds the next one's qup
        quc_next = o.q;
        last = o.q;
        acc.iters += (unsigned long long)o.iters;
        if (max_it && (unsigned)o.iters > *max_it) *max_it = (unsigned)o.iters;
        acc.capped += (unsigned long long)(o.flags & 1);
        acc.fallback += (unsigned long long)((o.flags >> 1) & 1);
        acc.noncanon += (unsigned long long)((o.flags >> 2) & 1);
    }
    cur[(size_t)u * M + m] = last;
    if (a.hist_out) {
        const int r = a.unit_reach[u];
        if (r >= 0) a.hist_out[((size_t)m * a.n_steps + t) * (size_t)a.n_reaches + r] = last;
    }
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long x)
{
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

__device__ __forceinline__ void flush_acc(const RouteArgs &a, const Acc &acc, long long wait,
                                          long long span)
{
    unsigned long long it = wave_sum(acc.iters), cp = wave_sum(acc.capped), fb = wave_sum(acc.fallback);
    const unsigned long long nc = wave_sum(acc.noncanon);
    const unsigned long long sl = wave_sum(acc.step_lanes);
    if ((threadIdx.x & 63) == 0) {
        if (acc.trips) atomicAdd(&a.counters[CNT_TRIPS], acc.trips);
        if (acc.step_trips) atomicAdd(&a.counters[CNT_STEP_TRIPS], acc.step_trips);
        if (sl) atomicAdd(&a.counters[CNT_STEP_LANES], sl);
        if (it) atomicAdd(&a.counters[CNT_ITERS], it);
        if (cp) atomicAdd(&a.counters[CNT_CAPPED], cp);
        if (fb) atomicAdd(&a.counters[CNT_FALLBACK], fb);
        if (nc) atomicAdd(&a.counters[CNT_NONCANON], nc);
        if (acc.mlp_calls) atomicAdd(&a.counters[CNT_MLP_CALLS], acc.mlp_calls);
        if (acc.mlp_ticks) atomicAdd(&a.counters[CNT_MLP_TICKS], acc.mlp_ticks);
        if (acc.sched_ticks) atomicAdd(&a.counters[CNT_SCHED_TICKS], acc.sched_ticks);
    }
    if (threadIdx.x == 0) {
        atomicAdd(&a.counters[CNT_WAIT_TICKS], (unsigned long long)wait);
        atomicAdd(&a.counters[CNT_BLOCK_TICKS], (unsigned long long)span);
    }
}

// Two-level device-scope barrier.  Counters only grow within a launch, so the
// last arriver of barrier `epoch` in a group sees (epoch+1)*group_size - 1.
// Word layout (unsigned, 128-byte spaced): [0] release epoch, [32] top
// counter, [64 + 32 g] group g.  Every wave fences at agent scope before and
// after, so stores of the super-level are visible across XCDs.
__device__ __forceinline__ bool grid_sync(const RouteArgs &a, unsigned epoch, long long &wait)
{
    __shared__ int s_ok;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const long long t0 = (long long)wall_clock64();
        const unsigned nb = gridDim.x;
        const unsigned G = (unsigned)a.groups < nb ? (unsigned)a.groups : nb;
        const unsigned g = blockIdx.x % G;
        const unsigned gsize = nb / G + (g < nb % G ? 1u : 0u);
        unsigned *flag = a.bar;
        unsigned *top = a.bar + 32;
        unsigned *grp = a.bar + 64 + 32 * g;
        const unsigned target = epoch + 1;
        int ok = 1;
        unsigned old = __hip_atomic_fetch_add(grp, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (old == target * gsize - 1) {
            unsigned old2 = __hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (old2 == target * G - 1)
                __hip_atomic_store(flag, target, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        while (__hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != KSTATUS_OK) {
                ok = 0;
                break;
            }
            if ((long long)wall_clock64() - t0 > a.timeout_ticks) {
                __hip_atomic_store(a.status, (int)KSTATUS_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        wait += (long long)wall_clock64() - t0;
        s_ok = ok;
    }
    __syncthreads();
    __threadfence();
    return s_ok != 0;
}

// Persistent time-skewed sweep: super-level k routes every (unit, member) with
// level(unit) + t == k, 0 <= t < n_steps.  Units are level-sorted, so the
// items of k are one contiguous range; item i always goes to thread
// i mod nthreads, keeping a unit on one CU (and XCD) across super-levels.
template <int MODE>
__global__ void __launch_bounds__(256) k_route_skewed(RouteArgs a)
{
    const long long t_start = (long long)wall_clock64();
    if (a.trace_t && blockIdx.x == 0 && threadIdx.x == 0 && a.k_end > a.k_begin)
        a.trace_t[a.k_end - a.k_begin] = (unsigned long long)t_start;   // start stamp after the last slot
    const long long nthreads = (long long)gridDim.x * blockDim.x;
    const long long gtid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int M = a.n_members, T = a.n_steps, L = a.n_levels;
    Acc acc = {};
    long long wait = 0;
    unsigned epoch = 0;
    for (int k = a.k_begin; k < a.k_end; ++k) {
        const int lo_l = max(0, k - T + 1), hi_l = min(k, L - 1);
        if (lo_l <= hi_l) {
            const long long lo = (long long)a.level_ptr[lo_l] * M;
            const long long hi = (long long)a.level_ptr[hi_l + 1] * M;
            long long off = gtid - lo % nthreads;
            if (off < 0) off += nthreads;
            unsigned max_it = 0;
            for (long long i = lo + off; i < hi; i += nthreads) {
                const int u = (int)(i / M);
                const int m = (int)(i - (long long)u * M);
                route_item<MODE>(a, u, m, k - a.unit_level[u], acc, a.trace_it ? &max_it : nullptr);
            }
            if (a.trace_it && max_it) atomicMax(&a.trace_it[k - a.k_begin], max_it);
        }
        if (k + 1 < a.k_end) {
            if (!grid_sync(a, epoch, wait)) break;
            ++epoch;
        }
        if (a.trace_t && blockIdx.x == 0 && threadIdx.x == 0)
            a.trace_t[k - a.k_begin] = (unsigned long long)wall_clock64();
    }
    flush_acc(a, acc, wait, (long long)wall_clock64() - t_start);
}

// One unit, one member, one step of the dataflow schedule: upstream outflows
// come from the full history (slot t+1 holds step t, slot 0 the initial
// value), already checked present by the caller.
template <int MODE>
__device__ __forceinline__ void route_item_hist(const RouteArgs &a, int u, int m, int t,
                                               const double *up_prev, const double *up_cur, int n_up,
                                               bool internal, Acc &acc)
{
    const int M = a.n_members;
    double qup_in, quc_in;
    if (internal) {
        qup_in = up_prev[0];
        quc_in = up_cur[0];
    } else {
        qup_in = 0.0;
        quc_in = 0.0;
#pragma unroll
        for (int k = 0; k < kDataflowMaxUp; ++k) {
            if (k < n_up) {
                qup_in += up_prev[k];
                quc_in += up_cur[k];
            }
        }
    }
    const int s0 = a.unit_seg[u], s1 = a.unit_seg[u + 1];
    const double *ql_row = a.qlat + (size_t)(a.t_base + t) * (size_t)a.n_dev_segs * M;
    const bool last_step = t == a.n_steps - 1;
    double last = 0.0, qup_next = 0.0, quc_next = 0.0;
    for (int s = s0; s < s1; ++s) {
        const size_t sm = (size_t)s * M + m;
        const double qup = (s == s0) ? qup_in : qup_next;
        const double quc = (s == s0) ? quc_in : quc_next;
        const double qdp = a.q[sm];
        const double dp = a.d[sm];
        const double ql = ql_row[sm];
        const Seg c = load_seg(a.segc, (size_t)a.n_dev_segs, s);
        const SolveOut o = solve_one<MODE>(a, qup, quc, qdp, ql, c, dp, last_step);
        a.q[sm] = o.q;
        a.d[sm] = o.depth;
        if (last_step) a.v[sm] = o.vel;
        qup_next = qdp;
        quc_next = o.q;
        last = o.q;
        acc.iters += (unsigned long long)o.iters;
        acc.capped += (unsigned long long)(o.flags & 1);
        acc.fallback += (unsigned long long)((o.flags >> 1) & 1);
        acc.noncanon += (unsigned long long)((o.flags >> 2) & 1);
    }
    const size_t slot = (size_t)(t + 1) * (size_t)a.n_units * M + (size_t)u * M + m;
    __hip_atomic_store(&a.flow_hist[slot], (unsigned long long)__double_as_longlong(last),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Dataflow sweep without grid barriers.  Lane g owns member g % M of units
// u_j = g / M + j S (S = nthreads / M, level-sorted) and walks its items in
// super-level order k = level(u) + t.  An item starts when the outflows of
// its upstream units for steps t and t-1 are present in the history (not
// kPending); the value is its own ready flag, so no fences are needed: the
// item reads nothing else another lane wrote.  Every awaited item has a
// smaller k, so by induction on k every wait ends while all lanes are
// resident.  Lanes advance independently; a slow solve delays only the
// items downstream of it.
static constexpr int kMaxUp = kDataflowMaxUp;   // upstream links per unit kept in registers

template <int MODE>
__global__ void __launch_bounds__(256) k_route_dataflow(RouteArgs a)
{
    const long long t_start = (long long)wall_clock64();
    const int M = a.n_members, T = a.n_steps, U = a.n_units;
    const long long nthreads = (long long)gridDim.x * blockDim.x;
    const long long gtid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long S = nthreads / M;                  // unit stride (nthreads % M == 0)
    const int m = (int)(gtid % M);
    const long long base = gtid / M;
    const size_t UM = (size_t)U * M;
    Acc acc = {};
    long long wait_ticks = 0;
    const long long n_own = (base < U && gtid < S * M) ? (U - 1 - base) / S + 1 : 0;
    auto level_of = [&](long long jj) -> int {
        return jj < n_own ? a.unit_rec[2 * (size_t)(base + jj * S)].x : 0x7FFFFFFF;
    };
    // The lane's window of units [j_lo, j_hi] at super-level k.  Its items at
    // one super-level are independent of one another (an item waits only on
    // items of smaller k), so the lane takes them in any order: a queue of
    // the window's positions, 64 at a time (bit i of done_mask: position
    // bo + i routed at this k).  An item whose upstream outflows are not
    // there yet is passed over and polled again after the others, so one
    // late upstream unit does not hold the lane's other ready units.
    long long j_lo = 0, j_hi = -1;
    int lev_lo = level_of(0), lev_hi1 = lev_lo;
    int k = 0;
    bool done = (n_own == 0 || T <= 0);
    bool new_k = true;                 // the window of k is to be formed
    long long bo = 0;                  // first window position of the queue's 64
    unsigned long long done_mask = 0ull;
    int probe = 0;                     // next position to try, relative to bo
    int u = 0, t = 0, n_up = 0, pos = 0;
    bool internal = false;
    int ups[kMaxUp];
    long long spin_start = (long long)wall_clock64();
    while (true) {
        bool have_item = false;
        if (!done) {
            for (;;) {
                if (new_k) {
                    while (j_hi + 1 < n_own && lev_hi1 <= k) {
                        ++j_hi;
                        lev_hi1 = level_of(j_hi + 1);
                    }
                    while (j_lo <= j_hi && lev_lo + T - 1 < k) {
                        ++j_lo;
                        lev_lo = level_of(j_lo);
                    }
                    if (j_lo > j_hi) {                  // nothing active: jump to the next unit's level
                        if (j_hi + 1 >= n_own) { done = true; break; }
                        k = lev_hi1;
                        continue;
                    }
                    new_k = false;
                    bo = 0;
                    done_mask = 0ull;
                    probe = 0;
                }
                const long long W = j_hi - j_lo + 1;
                const int n = (int)min(64LL, W - bo);
                const unsigned long long full = (n == 64) ? ~0ull : ((1ull << n) - 1ull);
                if (done_mask == full) {                // these 64 are routed: the next 64, or k + 1
                    bo += 64;
                    done_mask = 0ull;
                    probe = 0;
                    if (bo >= W) {
                        ++k;
                        new_k = true;
                    }
                    continue;
                }
                // the first position at or after probe (cyclically) not yet routed
                const unsigned long long open = ~done_mask & full;
                const unsigned long long from = open & ~((probe >= 64 ? 0ull : (1ull << probe)) - 1ull);
                pos = (int)__ffsll((long long)(from ? from : open)) - 1;
                have_item = true;
                break;
            }
            if (have_item) {
                u = (int)(base + (j_lo + bo + pos) * S);
                const int4 r0 = a.unit_rec[2 * (size_t)u];
                const int4 r1 = a.unit_rec[2 * (size_t)u + 1];
                t = k - r0.x;
                n_up = r0.w >> 8;
                internal = (r0.w & 1) != 0;
                ups[0] = r1.x;
                ups[1] = r1.y;
                ups[2] = r1.z;
                ups[3] = r1.w;
            }
        }
        bool ready = false;
        double up_prev[kMaxUp], up_cur[kMaxUp];
        if (have_item) {
            ready = true;
#pragma unroll
            for (int q = 0; q < kMaxUp; ++q) {
                up_prev[q] = 0.0;
                up_cur[q] = 0.0;
                if (q < n_up) {
                    const size_t v = (size_t)ups[q] * M + m;
                    const unsigned long long bp = __hip_atomic_load(&a.flow_hist[(size_t)t * UM + v],
                                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const unsigned long long bc = __hip_atomic_load(&a.flow_hist[(size_t)(t + 1) * UM + v],
                                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ready = ready && bp != kPending && bc != kPending;
                    up_prev[q] = __longlong_as_double((long long)bp);
                    up_cur[q] = __longlong_as_double((long long)bc);
                }
            }
            probe = pos + 1;                             // after this one, try the next position
        }
        if (ready) {
            wait_ticks += (long long)wall_clock64() - spin_start;
            route_item_hist<MODE>(a, u, m, t, up_prev, up_cur, n_up, internal, acc);
            done_mask |= 1ull << pos;
            spin_start = (long long)wall_clock64();
        }
        ++acc.trips;
        if (__any(ready)) ++acc.step_trips;
        if (ready) ++acc.step_lanes;
        const bool busy = !done;
        if (!__any(busy)) break;
        if (!__any(ready)) {
            if (have_item && (long long)wall_clock64() - spin_start > a.timeout_ticks)
                __hip_atomic_store(a.status, (int)KSTATUS_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != KSTATUS_OK) break;
            __builtin_amdgcn_s_sleep(1);
        }
    }
    flush_acc(a, acc, wait_ticks, (long long)wall_clock64() - t_start);
}

// Dataflow sweep with secant iterations interleaved across the lanes of a
// wave.  Same ownership, ordering and readiness test as k_route_dataflow, but
// a lane's work is a state machine, and each trip of the lane loop is
//   (1) a scheduling phase: lanes without a running solve look for their next
//       item, poll its upstream outflows, and load the next segment (all the
//       memory reads of the loop happen here), then
//   (2) a solve phase: up to batch_iters secant iterations for every lane
//       with a running solve, touching registers only.  A lane whose solve
//       ends stores its segment results (and, at the unit's last segment,
//       publishes the outflow) at once and idles until the next phase (1).
// A slow solve therefore holds its own lane only, not the other 63 lanes of
// its wave.  REFERENCE solves are SecantRef, identical to solve_reference;
// CONVERGED and LEARNED solves are SecantConv, one residual evaluation per
// iteration.  LEARNED: a lane that opens a segment evaluates F at the
// previous depth there; after the scheduling phase MFMA forwards of the
// wave, each packing the two samples (the routed flow and the anchor's
// Manning flow) of up to 32 such lanes, give every such lane its guess
// (weights staged in LDS).
// LEARNED: a lane whose routed and anchor Manning flows differ by less than
// this in |ln(q_new / q_old)| (about 20 %) skips the network: its correction
// e^{y(q_new) - y(q_old)} is then near 1, and the wave's forward would cost
// more than the evaluations it saves
constexpr float kNetMinLogRatio = 0.2f;


// ------------------------------------------------- certified-solve servers
// A CONVERGED/LEARNED solve whose F(mindepth) <= 0 needs the certified scan
// (trmc_solver.h), whose interval arithmetic would add its registers to the
// routing loop if it ran there.  Routing lanes hand such solves to server
// waves (the last blocks of the same cooperative launch) and wait, as they
// wait for upstream outflows:
//   request:  the lane writes its inputs to its own slot, takes a ticket
//             (scan_ctl[0]) and publishes (ticket + 1) << 20 | lane in ring
//             entry ticket mod scan_cap with release order, once the entry's
//             previous occupant has been read (entry 0 or 1);
//   serve:    a server lane claims the next ticket (scan_ctl[32]), waits for
//             its entry (acquire), reads the request, marks the entry read
//             (1), solves, writes the answer and then scan_resp_seq[lane] =
//             ticket + 1 (release);
//   end:      every routing wave counts itself in scan_ctl[64] when it leaves
//             its loop (a routing lane leaves only with its answer), and a
//             server lane whose ticket has no request ends once all have.
// Every wait also ends when the launch's status leaves OK.
__device__ __forceinline__ void scan_post(const RouteArgs &a, long long lane, int seg, const SecantConv &sr,
                                          unsigned long long &ticket)
{
    double *rq = a.scan_req + (size_t)lane * 5;
    rq[0] = sr.qup;
    rq[1] = sr.quc;
    rq[2] = sr.qdp;
    rq[3] = sr.ql;
    rq[4] = (double)seg;
    ticket = __hip_atomic_fetch_add(&a.scan_ctl[0], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    unsigned long long *e = &a.scan_ring[ticket & (unsigned long long)(a.scan_cap - 1)];
    // the entry's previous occupant (ticket - scan_cap) must have been read
    const long long t0 = (long long)wall_clock64();
    while (__hip_atomic_load(e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > 1ull) {
        if ((long long)wall_clock64() - t0 > a.timeout_ticks) {
            __hip_atomic_store(a.status, (int)KSTATUS_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __hip_atomic_store(e, ((ticket + 1ull) << kScanLaneBits) | (unsigned long long)lane, __ATOMIC_RELEASE,
                       __HIP_MEMORY_SCOPE_AGENT);
}

// One certified solve by the whole wave, the device form of
// certified_solve_lanes (trmc_solver.h): each round the 64 lanes take one
// piece of the span each, side by side; the first unproven piece comes from a
// ballot and its values from the lanes that hold them.  Every lane ends with
// the same root, outflow and count.
__device__ __forceinline__ int certified_solve_wave(double qup, double quc, double qdp, double ql, const Seg &c,
                                                    const ConvCfg &k, double &h, double &q)
{
    const int lane = (int)(threadIdx.x & 63u);
    Coef C0;
    double qmn0;
    const double f001 = residual(0.01, k.dt, qup, quc, qdp, ql, c, C0, qmn0);   // <= 0: why it was sent
    const double rmax = routed_max(k.dt, qup, quc, qdp, ql);
    ScanWave w;
    scan_wave_begin(w, f001, outflow(C0, qup, quc, qdp));
    int evals = 1;
    double b_end = 0.0;
    bool geo = false;
    while (w.state == 0 && w.rounds < kScanCap && scan_wave_round(w, c, rmax, b_end, geo)) {
        const double x0 = scan_piece_x(w, b_end, geo, lane);
        const double x1 = scan_piece_x(w, b_end, geo, lane + 1);
        ScanPiece pc = scan_piece_eval(x1, w.spos, k.dt, qup, quc, qdp, ql, c);
        double f0 = __shfl_up(pc.f, 1u);                  // F(x0): the previous lane's right end
        if (lane == 0) f0 = w.fa;
        const bool proven = scan_piece_proof(x0, x1, f0, pc, w.spos, k.dt, qup, quc, qdp, ql, c, rmax);
        const unsigned long long open = __ballot(!proven);
        const int first = open ? (int)__ffsll((long long)open) - 1 : 64;
        const int f = first < 64 ? first : 63;
        const int left = (first == 0 || first == 64) ? 0 : first - 1;
        ScanPiece P;
        P.f = __shfl(pc.f, f);
        P.q = __shfl(pc.q, f);
        P.change = __shfl((int)pc.change, f) != 0;
        P.proven = false;
        double fl = __shfl(pc.f, left), qlv = __shfl(pc.q, left);
        if (first == 0 || first == 64) {
            fl = w.fa;
            qlv = w.qa;
        }
        const double xf = __shfl(x0, f), xn = __shfl(x1, f);
        const double f_end = __shfl(pc.f, 63), q_end = __shfl(pc.q, 63);
        scan_wave_combine(w, first, xf, xn, fl, qlv, P, b_end, f_end, q_end);
        ++evals;
    }
    if (w.state == 0) w.state = w.spos ? 4 : 1;           // the round cap
    return evals + scan_wave_finish(w, qup, quc, qdp, ql, c, k, h, q);
}

template <int MODE>
__device__ __forceinline__ void scan_server(const RouteArgs &a)
{
    // One request at a time per wave, solved by all of its lanes: the wave
    // takes the next ticket, waits for that request (every lane reads the
    // same words, lane 0's values decide), and answers it from lane 0.
    const unsigned long long route_waves = (unsigned long long)a.n_route_blocks * ((blockDim.x + 63) / 64);
    const ConvCfg kc = {a.dt, a.tol, a.max_eval, a.learned_budget};
    const unsigned long long mask = (unsigned long long)(a.scan_cap - 1);
    const int lane = (int)(threadIdx.x & 63u);
    // a routing lane waits on every answer: serve ahead of the routing waves
    // that share the CU (idle servers sleep, so they take little from them)
    __builtin_amdgcn_s_setprio(3);
    for (;;) {
        unsigned long long my = 0;
        if (lane == 0) my = __hip_atomic_fetch_add(&a.scan_ctl[32], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        my = __shfl(my, 0);
        unsigned long long *e = &a.scan_ring[my & mask];
        unsigned long long v = 0;
        for (;;) {
            v = __shfl(__hip_atomic_load(e, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT), 0);
            if ((v >> kScanLaneBits) == my + 1ull) break;
            // no request for this ticket yet: none will come once every routing wave is done
            const unsigned long long done = __shfl(
                __hip_atomic_load(&a.scan_ctl[64], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0);
            const int st = __shfl(__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT), 0);
            if (done >= route_waves || st != KSTATUS_OK) return;
            // requests are rare: poll every few microseconds, so that idle
            // servers take no memory bandwidth or issue slots from the routing
            __builtin_amdgcn_s_sleep(127);
        }
        const long long req = (long long)(v & ((1ull << kScanLaneBits) - 1ull));
        const double *rq = a.scan_req + (size_t)req * 5;
        const double qup = rq[0], quc = rq[1], qdp = rq[2], ql = rq[3];
        const int seg = (int)rq[4];
        // the entry is free for the ticket cap later (the request itself
        // stays in the requesting lane's slot until it has its answer)
        if (lane == 0) __hip_atomic_store(e, 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const long long ts0 = (long long)wall_clock64();
        const Seg c = load_seg(a.segc, (size_t)a.n_dev_segs, seg);
        double h = 0.01, q = 0.0;
        const int n = certified_solve_wave(qup, quc, qdp, ql, c, kc, h, q);
        if (lane == 0) {
            double *rs = a.scan_resp + (size_t)req * 3;
            rs[0] = h;
            rs[1] = q;
            rs[2] = (double)n;
            __hip_atomic_store(&a.scan_resp_seq[req], my + 1ull, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            atomicAdd(&a.counters[CNT_SCAN_SOLVE_TICKS], (unsigned long long)((long long)wall_clock64() - ts0));
        }
    }
}

// the solver and the segment constants of each rule (MODE 4: the FP32
// REFERENCE rule, float arithmetic on float values held in the double state)
template <int MODE> struct IlSolver { typedef SecantConv type; typedef Seg seg; };
template <> struct IlSolver<0> { typedef SecantRef type; typedef Seg seg; };
template <> struct IlSolver<3> { typedef SecantRef type; typedef Seg seg; };   // LINEAR: no iteration (unused)
template <> struct IlSolver<4> { typedef SecantRef32 type; typedef Seg32 seg; };

template <int MODE>
__device__ __forceinline__ void route_dataflow_il(const RouteArgs &a)
{
    if constexpr (MODE == 1 || MODE == 2) {
        if ((int)blockIdx.x >= a.n_route_blocks) {   // a server block (whole block: no barrier below is split)
            scan_server<MODE>(a);
            return;
        }
    }
    __shared__ float sW[MODE == 2 ? kMlpParams : 1];
    if constexpr (MODE == 2) {
        // the weights exist only for the network guess (guess_model 0)
        if (a.guess_model == 0)
            for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) sW[i] = a.mlp[i];
        __syncthreads();
    }
    const long long t_start = (long long)wall_clock64();
    const int M = a.n_members, T = a.n_steps, U = a.n_units;
    const long long nthreads = (long long)a.n_route_blocks * blockDim.x;
    const long long gtid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long S = nthreads / M;
    const int m = (int)(gtid % M);
    const long long base = gtid / M;
    const size_t UM = (size_t)U * M;
    const size_t nd = (size_t)a.n_dev_segs;
    // FP64 secant rules: the constants and state (flow, depth, velocity) of
    // the first kIlCacheSlots segments of the lane's current unit live in LDS,
    // [kIlCacheSlots][blockDim] Seg records, then [kIlCacheSlots][3][blockDim]
    constexpr bool kCache = (MODE == 1 || MODE == 2);
    extern __shared__ double il_lds[];
    const int bd = (int)blockDim.x, tid = (int)threadIdx.x;
    Seg *const segs = reinterpret_cast<Seg *>(il_lds);
    double *const sst = il_lds + (size_t)kIlCacheSlots * bd * (sizeof(Seg) / sizeof(double));
    auto lds = [&](int jj, int f) -> double & { return sst[(size_t)(jj * 3 + f) * bd + tid]; };
    unsigned long long scan_ticket = 0;
    bool scan_posted = false;
    long long scan_t0 = 0;
    Acc acc = {};
    long long wait_ticks = 0;
    long long j_lo = 0, j_hi = -1, j = 0;
    int k = 0;
    const long long n_own = (base < U && gtid < S * M) ? (U - 1 - base) / S + 1 : 0;
    // the window of owned units [j_lo, j_hi] at super-level k keeps the levels
    // of units j_lo and j_hi + 1 (loaded when the window moves)
    auto level_of = [&](long long jj) -> int {
        return jj < n_own ? a.unit_rec[2 * (size_t)(base + jj * S)].x : 0x7FFFFFFF;
    };
    int lev_lo = level_of(0), lev_hi1 = lev_lo;
    // The lane's current unit: its record (and, kCache, its segments in LDS).
    // A lane walks one unit's items over T consecutive super-levels and its
    // units lie far apart in level order (on the 1M-segment network 0.35 % of
    // a lane's items change unit), so the record and segment constants come
    // from HBM once per unit, not once per item.
    int cu = -1, cu_lev = 0;
    // the outflows at t - 1 of item (u, t) are those summed for (u, t - 1):
    // carried over (the same values, the same order), they are not reloaded
    int carry_u = -1, carry_t = -1;
    double carry_sp = 0.0, carry_p0 = 0.0;
    double ql_pre0 = 0.0, ql_pre1 = 0.0;             // kCache: lateral inflow of the first two segments
    // PH_NEXT: the unit's next segment needs opening (loads) in phase (1)
    enum { PH_DONE = 0, PH_NEED, PH_WAIT, PH_NEXT, PH_SOLVE };
    int phase = (n_own == 0 || T <= 0) ? PH_DONE : PH_NEED;
    int u = 0, t = 0, n_up = 0, s = 0, s0 = 0, s1 = 0;
    bool internal = false, crit = false;
    int4 ups = make_int4(0, 0, 0, 0);
    double qup_in = 0.0, quc_in = 0.0, qup_next = 0.0, quc_next = 0.0, last = 0.0, cur_qdp = 0.0;
    const double *ql_row = a.qlat;
    typename IlSolver<MODE>::seg c;
    typename IlSolver<MODE>::type sr;
    const ConvCfg kc = {a.dt, a.tol, a.max_eval, a.learned_budget};
    long long spin_start = 0;

    // true when the segment needs no iteration (guard)
    auto start_solve = [&](double qup, double quc, double qdp, double ql, double dp) -> bool {
        if constexpr (MODE == 0 || MODE == 3) {
            return sr.start(a.dt, qup, quc, qdp, ql, c, dp);
        } else if constexpr (MODE == 4) {   // every value is a float (exact conversions)
            return sr.start((float)a.dt, (float)qup, (float)quc, (float)qdp, (float)ql, c, (float)dp);
        } else {
            return sr.start(qup, quc, qdp, ql, c, dp, MODE == 2);
        }
    };

    // LEARNED: the guesses of the lanes whose solve has just evaluated F at
    // the previous depth (the anchor) and asks for one
    auto make_guesses = [&]() {
        if constexpr (MODE == 2) {
            const bool want = phase == PH_SOLVE && sr.need_guess;
            if (!__any(want)) return;
            if (a.guess_model != 0) {                // power law alone: no network
                if (want)
                    sr.set_guess(mlp_guess_from(sr.guess_anchor(), sr.guess_flow_new(), sr.guess_flow_old(), 0.0f,
                                                0.0f));
                return;
            }
            // the network's factor e^{y(q_new) - y(q_old)} is near 1 when
            // the flow barely changes: such lanes take the power law alone
            bool net = false;
            if (want) {
                const float r = (float)(fmax(sr.guess_flow_new(), 1e-6) / fmax(sr.guess_flow_old(), 1e-6));
                net = !(fabsf(__logf(r)) < kNetMinLogRatio);
                if (!net)
                    sr.set_guess(mlp_guess_from(sr.guess_anchor(), sr.guess_flow_new(), sr.guess_flow_old(), 0.0f,
                                                0.0f));
            }
            if (__any(net)) {                        // wave-uniform: all 64 lanes run the forward
                const long long m0 = (long long)wall_clock64();
                float xn[kMlpIn] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f}, xo[kMlpIn] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
                if (net) {
                    mlp_segment_inputs(sr.guess_flow_new(), c, xn);
                    mlp_segment_inputs(sr.guess_flow_old(), c, xo);
                }
                float yn = 0.0f, yo = 0.0f;
                const int calls = mlp_wave_pairs(sW, net, xn, xo, yn, yo);
                if (net)
                    sr.set_guess(mlp_guess_from(sr.guess_anchor(), sr.guess_flow_new(), sr.guess_flow_old(), yn, yo));
                acc.mlp_calls += (unsigned long long)calls;
                acc.mlp_ticks += (unsigned long long)((long long)wall_clock64() - m0);
            }
        }
    };

    auto publish = [&]() {
        const size_t slot = (size_t)(t + 1) * UM + (size_t)u * M + m;
        __hip_atomic_store(&a.flow_hist[slot], (unsigned long long)__double_as_longlong(last),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    // store a finished segment and move to the next one (stores only)
    auto close_segment = [&](const SolveOut &o, double qdp) {
        const int jj = s - s0;
        const bool last_step = t == T - 1;       // the only step whose velocity is kept
        if (kCache && jj < kIlCacheSlots) {
            lds(jj, 0) = o.q;
            lds(jj, 1) = o.depth;
            if (last_step) lds(jj, 2) = o.vel;
        } else {
            const size_t sm = (size_t)s * M + m;
            a.q[sm] = o.q;
            if constexpr (MODE != 3) {            // LINEAR carries flow only
                a.d[sm] = o.depth;
                if (last_step) a.v[sm] = o.vel;
            }
        }
        acc.iters += (unsigned long long)o.iters;
        acc.capped += (unsigned long long)(o.flags & 1);
        acc.fallback += (unsigned long long)((o.flags >> 1) & 1);
        acc.noncanon += (unsigned long long)((o.flags >> 2) & 1);
        qup_next = qdp;
        quc_next = o.q;
        last = o.q;
        ++s;
        if (s >= s1) {
            publish();
            ++j;
            phase = PH_NEED;
        } else {
            phase = PH_NEXT;
        }
    };
    // load segment s and start its solve; guard-trivial segments close at once
    auto open_segments = [&]() {
        while (phase == PH_NEXT) {
            const size_t sm = (size_t)s * M + m;
            const double qup = (s == s0) ? qup_in : qup_next;
            const double quc = (s == s0) ? quc_in : quc_next;
            if constexpr (MODE == 3) {
                // linearized MC, linear_mc.cu:171-180: fixed coefficients, terms
                // in the reference's order C1 new + C2 old + C3 Q_old + C4/ql ql
                const double qdp = a.q[sm];
                SolveOut o;
                o.q = a.lin[s] * quc + a.lin[nd + s] * qup + a.lin[2 * nd + s] * qdp + a.lin[3 * nd + s] * ql_row[sm];
                o.depth = 0.0;
                o.vel = 0.0;
                o.iters = 0;
                o.flags = 0;
                close_segment(o, qdp);
                continue;
            }
            const int jj = s - s0;
            double qdp, dp, ql;
            if (kCache && jj < kIlCacheSlots) {
                if constexpr (kCache) c = segs[(size_t)jj * bd + tid];
                qdp = lds(jj, 0);
                dp = lds(jj, 1);
                ql = (jj == 0) ? ql_pre0 : ql_pre1;
            } else {
                if constexpr (MODE == 4) c = load_seg32(a.segc32, nd, s);
                else c = load_seg(a.segc, nd, s);
                qdp = a.q[sm];
                dp = a.d[sm];
                ql = ql_row[sm];
            }
            if (!start_solve(qup, quc, qdp, ql, dp)) {
                cur_qdp = qdp;
                phase = PH_SOLVE;
            } else {
                close_segment(sr.finish(c, t == T - 1), qdp);
            }
        }
    };
    // kCache: the state of the unit in LDS back to HBM
    auto write_back = [&]() {
        if constexpr (kCache) {
            if (cu < 0) return;
            const int n = min(kIlCacheSlots, s1 - s0);
            for (int jj = 0; jj < n; ++jj) {
                const size_t sm = (size_t)(s0 + jj) * M + m;
                a.q[sm] = lds(jj, 0);
                a.d[sm] = lds(jj, 1);
                a.v[sm] = lds(jj, 2);
            }
        }
    };
    // unit uu becomes the lane's current unit: its record, and (kCache) the
    // constants and state of its first segments into LDS
    auto fill = [&](int uu) {
        const int4 r0 = a.unit_rec[2 * (size_t)uu];
        ups = a.unit_rec[2 * (size_t)uu + 1];
        cu = uu;
        cu_lev = r0.x;
        s0 = r0.y;
        s1 = r0.z;
        internal = (r0.w & 1) != 0;
        crit = (r0.w & 2) != 0;
        n_up = r0.w >> 8;
        if constexpr (kCache) {
            const int n = min(kIlCacheSlots, s1 - s0);
            for (int jj = 0; jj < n; ++jj) {
                const int sg = s0 + jj;
                segs[(size_t)jj * bd + tid] = load_seg(a.segc, nd, sg);
                const size_t sm = (size_t)sg * M + m;
                lds(jj, 0) = a.q[sm];
                lds(jj, 1) = a.d[sm];
                lds(jj, 2) = a.v[sm];
            }
        }
    };

    while (true) {
        bool progressed = false;
        const long long sched_t0 = (long long)wall_clock64();
        // ---------------------------------------------- (1) scheduling phase
        if (phase == PH_NEXT) {
            open_segments();
            progressed = true;
        }
        if (phase == PH_NEED) {
            for (;;) {
                while (j_hi + 1 < n_own && lev_hi1 <= k) {
                    ++j_hi;
                    lev_hi1 = level_of(j_hi + 1);
                }
                while (j_lo <= j_hi && lev_lo + T - 1 < k) {
                    ++j_lo;
                    lev_lo = level_of(j_lo);
                }
                if (j < j_lo) j = j_lo;
                if (j_lo <= j_hi && j <= j_hi) break;
                if (j_lo > j_hi) {
                    if (j_hi + 1 >= n_own) { phase = PH_DONE; break; }
                    k = lev_hi1;
                } else {
                    ++k;
                }
                j = j_lo;
            }
            if (phase == PH_NEED) {
                u = (int)(base + j * S);
                if (u != cu) {                       // another unit: its record (and segments)
                    write_back();
                    fill(u);
                }
                t = k - cu_lev;
                ql_row = a.qlat + (size_t)(a.t_base + t) * nd * M;
                if constexpr (kCache) {             // loaded with the outflows below
                    if (s0 < s1) ql_pre0 = ql_row[(size_t)s0 * M + m];
                    if (s0 + 1 < s1) ql_pre1 = ql_row[(size_t)(s0 + 1) * M + m];
                }
                phase = PH_WAIT;
                spin_start = (long long)wall_clock64();
            }
        }
        if (phase == PH_WAIT) {
            bool ready = true;
            double sp = 0.0, sc = 0.0, first_p = 0.0, first_c = 0.0;
            float spf = 0.0f, scf = 0.0f;          // MODE 4: float sums (wavefront_mc.cu:444-451)
            const bool carried = (MODE != 4) && cu == carry_u && t == carry_t;
            const int upv[4] = {ups.x, ups.y, ups.z, ups.w};
#pragma unroll
            for (int q = 0; q < kMaxUp; ++q) {
                if (q < n_up) {
                    const size_t v = (size_t)upv[q] * M + m;
                    const unsigned long long bc = __hip_atomic_load(&a.flow_hist[(size_t)(t + 1) * UM + v],
                                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const unsigned long long bp =
                        carried ? 0ull
                                : __hip_atomic_load(&a.flow_hist[(size_t)t * UM + v], __ATOMIC_RELAXED,
                                                    __HIP_MEMORY_SCOPE_AGENT);
                    ready = ready && (carried || bp != kPending) && bc != kPending;
                    const double vp = __longlong_as_double((long long)bp);
                    const double vc = __longlong_as_double((long long)bc);
                    // fixed-order sums starting from 0.0 (wavefront_mc.cu:523-529)
                    sp += vp;
                    sc += vc;
                    if constexpr (MODE == 4) {
                        spf += (float)vp;
                        scf += (float)vc;
                    }
                    if (q == 0) { first_p = vp; first_c = vc; }
                }
            }
            if (carried) {
                sp = carry_sp;
                first_p = carry_p0;
            }
            if (ready) {
                wait_ticks += (long long)wall_clock64() - spin_start;
                if constexpr (MODE == 4) {
                    sp = (double)spf;
                    sc = (double)scf;
                }
                carry_u = cu;
                carry_t = t + 1;
                carry_sp = sc;
                carry_p0 = first_c;
                qup_in = internal ? first_p : sp;
                quc_in = internal ? first_c : sc;
                s = s0;
                last = 0.0;
                progressed = true;
                if (s >= s1) {                       // unit without segments: outflow 0
                    publish();
                    ++j;
                    phase = PH_NEED;
                } else {
                    phase = PH_NEXT;
                    open_segments();
                }
            }
        }
        // ------------- CONVERGED/LEARNED: certified solves handed to the servers
        if constexpr (MODE == 1 || MODE == 2) {
            if (phase == PH_SOLVE && sr.awaiting_scan()) {
                if (!scan_posted) {
                    scan_post(a, gtid, s, sr, scan_ticket);
                    scan_posted = true;
                    scan_t0 = (long long)wall_clock64();
                } else if (__hip_atomic_load(&a.scan_resp_seq[gtid], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) ==
                           scan_ticket + 1ull) {
                    const double *rs = a.scan_resp + (size_t)gtid * 3;
                    sr.take_scan_result(rs[0], rs[1], (int)rs[2]);
                    atomicAdd(&a.counters[CNT_SCAN_WAIT_TICKS],
                              (unsigned long long)((long long)wall_clock64() - scan_t0));
                    scan_posted = false;
                    progressed = true;
                    close_segment(sr.finish(c, t == T - 1), cur_qdp);
                } else if ((long long)wall_clock64() - scan_t0 > a.timeout_ticks) {
                    __hip_atomic_store(a.status, (int)KSTATUS_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
        // -------------------------------------- (2) solve phase, registers only
        // a wave solving a unit on the longest dependency chain issues first
        auto solving = [&]() -> bool {
            if constexpr (MODE == 1 || MODE == 2) return phase == PH_SOLVE && !sr.awaiting_scan();
            else return phase == PH_SOLVE;
        };
        if (__any(crit && solving())) __builtin_amdgcn_s_setprio(2);
        else __builtin_amdgcn_s_setprio(0);
        ++acc.trips;
        acc.sched_ticks += (unsigned long long)((long long)wall_clock64() - sched_t0);
        for (int r = 0; r < a.batch_iters; ++r) {
            if (!__any(solving())) break;
            ++acc.step_trips;
            if (solving()) {
                ++acc.step_lanes;
                progressed = true;
                bool fin;
                if constexpr (MODE == 0 || MODE == 3 || MODE == 4) fin = sr.step(c);
                else fin = sr.template step<true>(c, kc);
                if (fin) close_segment(sr.finish(c, t == T - 1), cur_qdp);
            }
            make_guesses();
        }
        if (!__any(phase != PH_DONE)) break;
        if (!__any(progressed)) {
            if (phase == PH_WAIT && (long long)wall_clock64() - spin_start > a.timeout_ticks)
                __hip_atomic_store(a.status, (int)KSTATUS_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (__hip_atomic_load(a.status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != KSTATUS_OK) break;
            __builtin_amdgcn_s_sleep(1);
        }
    }
    write_back();
    if constexpr (MODE == 1 || MODE == 2) {      // this routing wave is done: servers may end
        if ((threadIdx.x & 63) == 0)
            __hip_atomic_fetch_add(&a.scan_ctl[64], 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    flush_acc(a, acc, wait_ticks, (long long)wall_clock64() - t_start);
}

// Two waves per SIMD at least for the converged rules (the certified-solve
// server code in the same kernel needs the registers); the REFERENCE rule's
// kernel fits three (at most 168 VGPRs), which its short solves use.
template <int MODE>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) k_route_dataflow_il(RouteArgs a)
{
    route_dataflow_il<MODE>(a);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) k_route_dataflow_il_ref(RouteArgs a)
{
    route_dataflow_il<0>(a);
}

// The element kernels below loop over their range with the grid's stride: a
// dispatch counts its work-items in 32 bits, so one lane per element would
// truncate ranges beyond 2^32 (a 1M-unit, 16-member, 288-step history has
// 4.6e9 entries).
__global__ void k_hist_init(unsigned long long *hist, const double *ring, int src_slot, long long um, int T)
{
    const long long total = um * (long long)(T + 1);
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        if (i < um) hist[i] = (unsigned long long)__double_as_longlong(ring[(size_t)src_slot * um + i]);
        else hist[i] = kPending;
    }
}

__global__ void k_hist_finish(const unsigned long long *hist, double *ring, int dst_slot, long long um, int T)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < um; i += stride)
        ring[(size_t)dst_slot * um + i] = __longlong_as_double((long long)hist[(size_t)T * um + i]);
}

__global__ void k_hist_extract(const unsigned long long *hist, const int *reach_last_unit, int R, int U,
                               int M, int T, double *out)
{
    const long long total = (long long)M * T * R;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int r = (int)(i % R);
        const long long mt = i / R;
        const int t = (int)(mt % T), m = (int)(mt / T);
        out[i] = __longlong_as_double((long long)hist[((size_t)(t + 1) * U + reach_last_unit[r]) * M + m]);
    }
}

// One level of one step (reference-style schedule, graph-captured per step).
template <int MODE>
__global__ void __launch_bounds__(256) k_route_level(RouteArgs a, int level)
{
    const int M = a.n_members;
    const long long lo = (long long)a.level_ptr[level] * M;
    const long long hi = (long long)a.level_ptr[level + 1] * M;
    const long long i = lo + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    Acc acc = {};
    if (i < hi) {
        const int u = (int)(i / M);
        const int m = (int)(i - (long long)u * M);
        route_item<MODE>(a, u, m, *a.t_dev, acc);
    }
    unsigned long long it = wave_sum(acc.iters), cp = wave_sum(acc.capped);
    if ((threadIdx.x & 63) == 0) {
        if (it) atomicAdd(&a.counters[CNT_ITERS], it);
        if (cp) atomicAdd(&a.counters[CNT_CAPPED], cp);
    }
}

__global__ void k_step_advance(int *t_dev)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *t_dev += 1;
}

__global__ void k_reset(unsigned *w, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = 0u;
}

// Independent reaches, one step (troute_mc_final.cu:142-164): one thread per
// reach, sequential segments, quc of segment 0 = uq.
__global__ void __launch_bounds__(256) k_reach_parallel(const double *__restrict__ segc, int n_seg,
                                                        const double *__restrict__ ql,
                                                        const int *__restrict__ rs,
                                                        const int *__restrict__ rl,
                                                        const double *__restrict__ qu,
                                                        const double *__restrict__ qd,
                                                        const double *__restrict__ dp,
                                                        double *__restrict__ qo,
                                                        double *__restrict__ vo,
                                                        double *__restrict__ dpo, double uq,
                                                        double dt, int nr)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nr) return;
    double quc = uq;
    const int start = rs[r], ns = rl[r];
    for (int i = 0; i < ns; ++i) {
        const int idx = start + i;
        const Seg c = load_seg(segc, (size_t)n_seg, idx);
        const SolveOut o = solve_reference(dt, qu[idx], quc, qd[idx], ql[idx], c, dp[idx]);
        qo[idx] = o.q;
        vo[idx] = o.vel;
        dpo[idx] = o.depth;
        quc = o.q;
    }
}

// Independent segment solves: in [n][5] {qup, quc, qdp, ql, depthp}.
__global__ void __launch_bounds__(256) k_solve_batch(const double *__restrict__ segc, int n,
                                                     const double *__restrict__ in,
                                                     double *__restrict__ out,
                                                     int *__restrict__ iters, double dt, int mode,
                                                     double tol, int max_eval)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double *x = in + (size_t)i * 5;
    const Seg c = load_seg(segc, (size_t)n, i);
    SolveOut o;
    if (mode == 0) {
        o = solve_reference(dt, x[0], x[1], x[2], x[3], c, x[4]);
    } else {
        o = solve_converged(dt, x[0], x[1], x[2], x[3], c, x[4], tol, max_eval);
    }
    out[(size_t)i * 3 + 0] = o.q;
    out[(size_t)i * 3 + 1] = o.depth;
    out[(size_t)i * 3 + 2] = o.vel;
    if (iters) iters[i] = o.iters;
}

template <typename TSrc>
__global__ void k_scatter_forcing(const TSrc *__restrict__ src, int n_src, long long total,
                                  const int *__restrict__ seg_of_dev, int nd, int M, int m,
                                  double *__restrict__ dst)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const long long t = i / nd;
        const int ds = (int)(i - t * nd);
        dst[(size_t)i * M + m] = (double)src[(size_t)t * n_src + seg_of_dev[ds]];
    }
}

// Synthetic lateral inflow straight into the forcing layout
// [t][device segment][member]: the values trmc_generate_forcing gives for
// draw t * N + segment of member m.
__global__ void k_gen_forcing(double *__restrict__ dst, const int *__restrict__ seg_of_dev, int nd, int N, int M,
                              long long total, unsigned long long seed)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const int m = (int)(i % M);
        const long long r = i / M;
        const int ds = (int)(r % nd);
        const long long t = r / nd;
        dst[i] = rng::forcing_at(seed, m, (unsigned long long)t * (unsigned long long)N + seg_of_dev[ds]);
    }
}

// Plain streaming copy (16-byte vectors, grid-stride): the reference point
// for HBM bandwidth and for scaling the memory-traffic counters.
__global__ void __launch_bounds__(256) k_stream_copy(const double2 *__restrict__ src, double2 *__restrict__ dst,
                                                    long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

// ----------------------------------------------------------------- launchers

hipError_t launch_stream_copy(const void *src, void *dst, size_t bytes, hipStream_t stream)
{
    const long long n = (long long)(bytes / sizeof(double2));
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_stream_copy, dim3(4096), dim3(256), 0, stream, (const double2 *)src, (double2 *)dst, n);
    return hipGetLastError();
}

hipError_t launch_gen_forcing(double *dst, const int *seg_of_dev, int nd, int N, int M, int T,
                              unsigned long long seed, hipStream_t stream)
{
    const long long total = (long long)T * nd * M;
    if (total <= 0) return hipSuccess;
    const long long blocks = std::min<long long>((total + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_gen_forcing, dim3((unsigned)blocks), dim3(256), 0, stream, dst, seg_of_dev, nd, N, M, total,
                       seed);
    return hipGetLastError();
}

hipError_t launch_reset(unsigned *bar, int nwords, hipStream_t stream)
{
    const int blocks = std::max(1, std::min(1024, (nwords + 255) / 256));
    hipLaunchKernelGGL(k_reset, dim3(blocks), dim3(256), 0, stream, bar, nwords);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_skewed_t(const RouteArgs &a, int grid, int block, bool cooperative,
                                  hipStream_t stream)
{
    if (cooperative) {
        RouteArgs copy = a;
        void *args[] = {&copy};
        return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_route_skewed<MODE>),
                                          dim3(grid), dim3(block), args, 0, stream);
    }
    hipLaunchKernelGGL(k_route_skewed<MODE>, dim3(grid), dim3(block), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_route_skewed(const RouteArgs &a, int mode, int grid, int block, bool cooperative,
                               hipStream_t stream)
{
    return mode == 0   ? launch_skewed_t<0>(a, grid, block, cooperative, stream)
           : mode == 1 ? launch_skewed_t<1>(a, grid, block, cooperative, stream)
                       : launch_skewed_t<2>(a, grid, block, cooperative, stream);
}

hipError_t route_skewed_occupancy(int mode, int block, int *blocks_per_cu)
{
    return mode == 0   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_skewed<0>, block, 0)
           : mode == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_skewed<1>, block, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_skewed<2>, block, 0);
}

hipError_t launch_route_level(const RouteArgs &a, int mode, int level, long long items, int block,
                              hipStream_t stream)
{
    if (items <= 0) return hipSuccess;
    const dim3 grid((unsigned)((items + block - 1) / block));
    if (mode == 0)
        hipLaunchKernelGGL(k_route_level<0>, grid, dim3(block), 0, stream, a, level);
    else if (mode == 1)
        hipLaunchKernelGGL(k_route_level<1>, grid, dim3(block), 0, stream, a, level);
    else
        hipLaunchKernelGGL(k_route_level<2>, grid, dim3(block), 0, stream, a, level);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_dataflow_t(const RouteArgs &a, int grid, int block, hipStream_t stream)
{
    RouteArgs copy = a;
    void *args[] = {&copy};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&k_route_dataflow<MODE>), dim3(grid),
                                      dim3(block), args, 0, stream);
}

size_t dataflow_il_lds_bytes(int mode, int block)
{
    return (mode == 1 || mode == 2) ? (size_t)kIlCacheSlots * (sizeof(Seg) + 3 * sizeof(double)) * block : 0;
}

template <int MODE>
static const void *dataflow_il_kernel()
{
    if constexpr (MODE == 0) return reinterpret_cast<const void *>(&k_route_dataflow_il_ref);
    else return reinterpret_cast<const void *>(&k_route_dataflow_il<MODE>);
}

template <int MODE>
static hipError_t launch_dataflow_il_t(const RouteArgs &a, int grid, int block, hipStream_t stream)
{
    RouteArgs copy = a;
    void *args[] = {&copy};
    return hipLaunchCooperativeKernel(dataflow_il_kernel<MODE>(), dim3(grid), dim3(block), args,
                                      dataflow_il_lds_bytes(MODE, block), stream);
}

hipError_t launch_route_dataflow_il(const RouteArgs &a, int mode, int grid, int block, hipStream_t stream)
{
    return mode == 0   ? launch_dataflow_il_t<0>(a, grid, block, stream)
           : mode == 1 ? launch_dataflow_il_t<1>(a, grid, block, stream)
           : mode == 2 ? launch_dataflow_il_t<2>(a, grid, block, stream)
           : mode == 3 ? launch_dataflow_il_t<3>(a, grid, block, stream)
                       : launch_dataflow_il_t<4>(a, grid, block, stream);
}

hipError_t route_dataflow_il_occupancy(int mode, int block, int *blocks_per_cu)
{
    const size_t lds = dataflow_il_lds_bytes(mode, block);
    return mode == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_dataflow_il_ref, block, lds)
           : mode == 1
               ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_dataflow_il<1>, block, lds)
           : mode == 2
               ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_dataflow_il<2>, block, lds)
           : mode == 3
               ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_dataflow_il<3>, block, 0)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_dataflow_il<4>, block, 0);
}

// Linearized MC coefficients, linear_mc.cu:57-100 (compute_mc_coeffs): float
// arithmetic at the reference depth h = max(0.5 bankfull, 0.1); each power is
// the float rounding of the double pow (correctly rounded division and square
// root are the defaults for float).  out [4][n]: C1, C2, C3, C4/ql.  fixed_x:
// the weighting X = 0.25 of the matrix form (linear_mc_matrix.cu:58-84) in
// place of the Cunge formula.
__global__ void k_linear_coeffs(const float *__restrict__ params, int n, float dt, int fixed_x,
                                float *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *p = params + (size_t)i * 8;
    const float dx = p[0], bw = p[1], tw = p[2], twcc = p[3], n_ch = p[4], cs = p[6], s0 = p[7];
    const float z = (cs == 0.0f) ? 1.0f : 1.0f / cs;
    float bfd;
    if (bw > tw) bfd = bw / 0.00001f;
    else if (bw == tw) bfd = bw / (2.0f * z);
    else bfd = (tw - bw) / (2.0f * z);
    const float h = fmaxf(0.5f * bfd, 0.1f);
    const float twl = bw + 2.0f * z * h;
    const float A = (bw + h * z) * h;
    const float WP = bw + 2.0f * h * sqrtf(1.0f + z * z);
    const float R = (WP > 0.0f) ? A / WP : 0.0f;
    const float r23 = (float)pow((double)R, (double)(2.0f / 3.0f));
    const float r53 = (float)pow((double)R, (double)(5.0f / 3.0f));
    const float Ck = fmaxf(1e-6f, (sqrtf(s0) / n_ch) * ((5.0f / 3.0f) * r23 - (2.0f / 3.0f) * r53 *
                                                        (2.0f * sqrtf(1.0f + z * z) / (bw + 2.0f * h * z))));
    const float Km = fmaxf(dt, dx / Ck);
    const float tw_use = (h > bfd && twcc > 0.0f) ? twcc : twl;
    float X = 0.25f;
    if (!fixed_x && Ck > 0.0f && tw_use * s0 * Ck * dx > 0.0f)
        X = fminf(0.5f, fmaxf(0.0f, 0.5f * (1.0f - 0.0f / (2.0f * tw_use * s0 * Ck * dx))));
    float D = Km * (1.0f - X) + dt / 2.0f;
    if (D == 0.0f) D = 1.0f;
    out[i] = (Km * X + dt / 2.0f) / D;
    out[(size_t)n + i] = (dt / 2.0f - Km * X) / D;
    out[2 * (size_t)n + i] = (Km * (1.0f - X) - dt / 2.0f) / D;
    out[3 * (size_t)n + i] = dt / D;
}

// Diffusive-wave step, owp_extended_kernels.cu:35-94 (cpu_diffusive_tridiag):
// one lane per reach, the reference's float operations in its order.  Node i of
// reach j is at j * kDiffMaxNodes + i in the node arrays; the elimination
// factors eei go to a node-major workspace (ws[i * nreach + j]: lanes of a
// wave touch adjacent words), the factors ffi to qp_out itself, which the back
// substitution then overwrites in place.
constexpr int kDiffMaxNodes = 500;
constexpr int kDiffChunk = 8;   // nodes whose inputs are loaded together

__global__ void __launch_bounds__(256) k_diffusive_tridiag(const DiffReach *__restrict__ rp,
                                                           const float *__restrict__ dx,
                                                           const float *__restrict__ diffusivity,
                                                           const float *__restrict__ qp_prev,
                                                           float *__restrict__ qp_out, float *__restrict__ ws,
                                                           int nreach, unsigned *bad)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nreach) return;
    const int nc = rp[j].ncomp;
    if (nc < 1 || nc > kDiffMaxNodes) {       // only device-resident reach records reach here unchecked
        if (bad) atomicAdd(bad, 1u);
        return;
    }
    const float theta = rp[j].theta, dtini = rp[j].dtini;
    const size_t base = (size_t)j * kDiffMaxNodes;
    const float *q = qp_prev + base;
    float *out = qp_out + base;
    float e = 1.0f, f = 0.0f;                 // eei[0], ffi[0]
    ws[j] = e;
    out[0] = f;
    float q_im1 = q[0], q_i = (nc > 1) ? q[1] : q[0];
    // the node inputs do not depend on the elimination: each chunk of
    // kDiffChunk nodes is loaded at once, then eliminated in order
    for (int i0 = 1; i0 < nc; i0 += kDiffChunk) {
        float dxv[kDiffChunk], dfv[kDiffChunk], qv[kDiffChunk];
#pragma unroll
        for (int u = 0; u < kDiffChunk; ++u) {
            const int i = i0 + u;
            dxv[u] = (i < nc) ? dx[base + i - 1] : 1.0f;
            dfv[u] = (i < nc) ? diffusivity[base + i] : 1.0f;
            qv[u] = (i < nc - 1) ? q[i + 1] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kDiffChunk; ++u) {
            const int i = i0 + u;
            if (i >= nc) break;
            float dxi = dxv[u];
            if (dxi < 1.0f) dxi = 1.0f;
            const float alpha = 1.0f;
            float diff_i = dfv[u];
            if (diff_i < 0.01f) diff_i = 0.01f;
            const float ppi = -theta * diff_i * dtini / (dxi * dxi) * 2.0f / (alpha * (alpha + 1.0f)) * alpha;
            const float qqi = 1.0f - ppi * (alpha + 1.0f) / alpha;
            const float rri = ppi / alpha;
            const float q_ip1 = (i < nc - 1) ? qv[u] : q_i;
            const float ssi = q_i + dtini * diff_i * (1.0f - theta) * (q_im1 - 2.0f * q_i + q_ip1) / (dxi * dxi);
            float denom = ppi * e + qqi;
            if (fabsf(denom) < 1e-20f) denom = 1e-20f;
            e = -rri / denom;
            f = (ssi - ppi * f) / denom;
            ws[(size_t)i * nreach + j] = e;
            out[i] = f;
            q_im1 = q_i;
            q_i = q_ip1;
        }
    }
    float qn = e * q[nc - 1] + f;             // downstream ghost node: qp_prev[nc - 1]
    out[nc - 1] = qn;
    for (int i0 = nc - 2; i0 >= 0; i0 -= kDiffChunk) {
        float ev[kDiffChunk], fv[kDiffChunk];
#pragma unroll
        for (int u = 0; u < kDiffChunk; ++u) {
            const int i = i0 - u;
            ev[u] = (i >= 0) ? ws[(size_t)i * nreach + j] : 0.0f;
            fv[u] = (i >= 0) ? out[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kDiffChunk; ++u) {
            const int i = i0 - u;
            if (i < 0) break;
            qn = ev[u] * qn + fv[u];
            out[i] = qn;
        }
    }
}

// One step of the matrix-form linearized MC, linear_mc_matrix_graph.cu:153-205
// (kernel_spmv_fused): y = P_Q x + P_q ql for every row, both products and
// their sum in one pass.  The two matrices share nearly all of their pattern
// (a row's columns are the reach and its ancestors), so a row is stored once
// as the union of its two CSR rows: a column and the pair {P_Q value, P_q
// value} (0 where a matrix has no entry), and the vector is stored as pairs
// z[c] = {Q[c], ql[c]}, so one 8-byte gather serves both products.  Per
// entry, in column order: s = s + vQ * Q[c], then s = s + vq * ql[c] (a
// product with a missing entry's 0 adds a zero, which leaves a sum started at
// +0 unchanged).  G lanes share a row: lane l takes union entries l, l + G,
// ...; the G partial sums are added by xor distance G/2, ..., 1, which every
// lane of the group computes alike, so the row's value is fixed by its G.
// Each lane issues kMatrixUnroll entry loads, then their gathers, then adds in
// order, so the loads of a row overlap.  The step writes z_out[r] = {y[r],
// ql of the next step at r}.
//
// Work layout: the host orders the rows into slots (rows of equal G together,
// in row order) with slot[i] = {row, first entry, end}; each 256-lane block
// takes rows of one G (block table {G, first slot, slots}), 256 / G rows each.
constexpr int kMatrixUnroll = 4;

template <int G>
__device__ __forceinline__ void matrix_rows(int slot0, int nslots, const int4 *__restrict__ slots,
                                            const int *__restrict__ col, const float2 *__restrict__ val,
                                            const float2 *__restrict__ z, const float *__restrict__ ql_next,
                                            float2 *__restrict__ z_out)
{
    const int i = threadIdx.x / G, l = threadIdx.x % G;
    const bool live = i < nslots;
    float s = 0.0f;
    int row = 0;
    if (live) {
        const int4 sl = slots[slot0 + i];
        row = sl.x;
        for (int k = sl.y + l; k < sl.z; k += G * kMatrixUnroll) {
            int c[kMatrixUnroll];
            float2 v[kMatrixUnroll], zc[kMatrixUnroll];
#pragma unroll
            for (int u = 0; u < kMatrixUnroll; ++u) {
                const bool in = k + u * G < sl.z;
                c[u] = in ? col[k + u * G] : 0;
                v[u] = in ? val[k + u * G] : make_float2(0.0f, 0.0f);
            }
#pragma unroll
            for (int u = 0; u < kMatrixUnroll; ++u) zc[u] = (k + u * G < sl.z) ? z[c[u]] : make_float2(0.0f, 0.0f);
#pragma unroll
            for (int u = 0; u < kMatrixUnroll; ++u)
                if (k + u * G < sl.z) {
                    s = s + v[u].x * zc[u].x;
                    s = s + v[u].y * zc[u].y;
                }
        }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) s = s + __shfl_xor(s, o, G);
    if (live && l == 0) z_out[row] = make_float2(s, ql_next[row]);
}

// The grid loops over the block table (at most 2^20 blocks launched), so the
// work-item count of one dispatch stays within 32 bits for any table size.
__global__ void __launch_bounds__(256) k_matrix_step(int n_blocks, const int4 *__restrict__ blocks,
                                                     const int4 *__restrict__ slots, const int *__restrict__ col,
                                                     const float2 *__restrict__ val, const float2 *__restrict__ z,
                                                     const float *__restrict__ ql_next, float2 *__restrict__ z_out)
{
    for (int b = blockIdx.x; b < n_blocks; b += gridDim.x) {
        const int4 bt = blocks[b];   // {G, first slot, slots, 0}
        switch (bt.x) {
        case 1: matrix_rows<1>(bt.y, bt.z, slots, col, val, z, ql_next, z_out); break;
        case 2: matrix_rows<2>(bt.y, bt.z, slots, col, val, z, ql_next, z_out); break;
        case 4: matrix_rows<4>(bt.y, bt.z, slots, col, val, z, ql_next, z_out); break;
        case 8: matrix_rows<8>(bt.y, bt.z, slots, col, val, z, ql_next, z_out); break;
        case 16: matrix_rows<16>(bt.y, bt.z, slots, col, val, z, ql_next, z_out); break;
        case 32: matrix_rows<32>(bt.y, bt.z, slots, col, val, z, ql_next, z_out); break;
        default: matrix_rows<64>(bt.y, bt.z, slots, col, val, z, ql_next, z_out); break;
        }
    }
}

hipError_t launch_matrix_step(int n_blocks, const int4 *blocks, const int4 *slots, const int *col, const float2 *val,
                              const float2 *z, const float *ql_next, float2 *z_out, hipStream_t stream, int max_grid)
{
    if (n_blocks <= 0) return hipSuccess;
    const int grid = std::min(n_blocks, max_grid > 0 ? std::min(max_grid, 1 << 20) : 1 << 20);
    hipLaunchKernelGGL(k_matrix_step, dim3(grid), dim3(256), 0, stream, n_blocks, blocks, slots, col, val, z,
                       ql_next, z_out);
    return hipGetLastError();
}

hipError_t launch_diffusive_tridiag(const DiffReach *rp, const float *dx, const float *diffusivity,
                                    const float *qp_prev, float *qp_out, float *ws, int nreach, hipStream_t stream,
                                    unsigned *bad)
{
    if (nreach <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_diffusive_tridiag, dim3((nreach + 63) / 64), dim3(64), 0, stream, rp, dx, diffusivity,
                       qp_prev, qp_out, ws, nreach, bad);
    return hipGetLastError();
}

hipError_t launch_linear_coeffs(const float *params, int n, float dt, int fixed_x, float *out, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_linear_coeffs, dim3((n + 255) / 256), dim3(256), 0, stream, params, n, dt, fixed_x, out);
    return hipGetLastError();
}

hipError_t launch_route_dataflow(const RouteArgs &a, int mode, int grid, int block, hipStream_t stream)
{
    return mode == 0   ? launch_dataflow_t<0>(a, grid, block, stream)
           : mode == 1 ? launch_dataflow_t<1>(a, grid, block, stream)
                       : launch_dataflow_t<2>(a, grid, block, stream);
}

hipError_t route_dataflow_occupancy(int mode, int block, int *blocks_per_cu)
{
    return mode == 0   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_dataflow<0>, block, 0)
           : mode == 1 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_dataflow<1>, block, 0)
                       : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_route_dataflow<2>, block, 0);
}

// blocks of 256 for a grid-stride element kernel over n elements: enough to
// fill the GPU, at most 2^20
static unsigned element_blocks(long long n)
{
    return (unsigned)std::min<long long>((n + 255) / 256, 1LL << 20);
}

hipError_t launch_hist_init(unsigned long long *hist, const double *ring, int src_slot, long long um, int T,
                            hipStream_t stream)
{
    const long long total = um * (long long)(T + 1);
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_hist_init, dim3(element_blocks(total)), dim3(256), 0, stream, hist, ring, src_slot, um, T);
    return hipGetLastError();
}

hipError_t launch_hist_finish(const unsigned long long *hist, double *ring, int dst_slot, long long um, int T,
                              hipStream_t stream)
{
    if (um <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_hist_finish, dim3(element_blocks(um)), dim3(256), 0, stream, hist, ring, dst_slot, um, T);
    return hipGetLastError();
}

hipError_t launch_hist_extract(const unsigned long long *hist, const int *reach_last_unit, int R, int U, int M,
                               int T, double *out, hipStream_t stream)
{
    const long long total = (long long)M * T * R;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_hist_extract, dim3(element_blocks(total)), dim3(256), 0, stream, hist, reach_last_unit, R,
                       U, M, T, out);
    return hipGetLastError();
}

hipError_t launch_step_advance(int *t_dev, hipStream_t stream)
{
    hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, stream, t_dev);
    return hipGetLastError();
}

hipError_t launch_reach_parallel(const double *segc, int n_seg, const double *ql, const int *rs,
                                 const int *rl, const double *qu, const double *qd, const double *dp,
                                 double *qo, double *vo, double *dpo, double uq, double dt, int nr,
                                 hipStream_t stream)
{
    if (nr <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_reach_parallel, dim3((nr + 255) / 256), dim3(256), 0, stream, segc, n_seg,
                       ql, rs, rl, qu, qd, dp, qo, vo, dpo, uq, dt, nr);
    return hipGetLastError();
}

hipError_t launch_solve_batch(const double *segc, int n, const double *in, double *out, int *iters,
                              double dt, int mode, double tol, int max_eval, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_solve_batch, dim3((n + 255) / 256), dim3(256), 0, stream, segc, n, in, out,
                       iters, dt, mode, tol, max_eval);
    return hipGetLastError();
}

hipError_t launch_scatter_forcing_f64(const double *src, int n_src, int T, const int *seg_of_dev,
                                      int nd, int M, int m, double *dst, hipStream_t stream)
{
    const long long total = (long long)T * nd;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_forcing<double>, dim3(element_blocks(total)), dim3(256), 0,
                       stream, src, n_src, total, seg_of_dev, nd, M, m, dst);
    return hipGetLastError();
}

hipError_t launch_scatter_forcing_f32(const float *src, int n_src, int T, const int *seg_of_dev,
                                      int nd, int M, int m, double *dst, hipStream_t stream)
{
    const long long total = (long long)T * nd;
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_scatter_forcing<float>, dim3(element_blocks(total)), dim3(256), 0,
                       stream, src, n_src, total, seg_of_dev, nd, M, m, dst);
    return hipGetLastError();
}

}  // namespace trmc
