// Device kernel interface shared by the HIP translation unit and the host glue.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace trmc {

// Structure-of-arrays slots of the per-segment constant table (stride n_seg).
enum SegSlot {
    SC_DX = 0, SC_BW, SC_TWCC, SC_N, SC_NCC, SC_S0,
    SC_Z,     // 1/cs (1 when cs == 0)
    SC_BFD,   // bankfull depth
    SC_KS,    // sqrt(s0) / n_ch
    SC_SQ,    // sqrt(1 + z*z)
    SC_SQS0,  // sqrt(s0)
    SC_COUNT
};


// Counter slots written by the routing kernels (unsigned long long each).
enum CounterSlot {
    CNT_ITERS = 0,      // secant iterations (REFERENCE) / residual evaluations
    CNT_CAPPED,         // CONVERGED/LEARNED solves that hit the cap
    CNT_FALLBACK,       // LEARNED solves restarted from the reference bracket
    CNT_WAIT_TICKS,     // barrier wait, wall-clock ticks summed over blocks
    CNT_BLOCK_TICKS,    // total wall-clock ticks summed over blocks
    CNT_MLP_CALLS,      // LEARNED: wave-wide network forwards
    CNT_MLP_TICKS,      // LEARNED: wall-clock ticks inside them, summed over waves
    CNT_SCAN_WAIT_TICKS,   // certified solves: ticks from request to answer, summed over requests
    CNT_SCAN_SOLVE_TICKS,  // certified solves: ticks a server lane spends solving, summed
    CNT_TRIPS,          // interleaved kernels: loop trips, summed over waves
    CNT_STEP_TRIPS,     // interleaved kernels: trips that ran a solve step, summed over waves
    CNT_STEP_LANES,     // interleaved kernels: lane solve steps, summed over lanes
    CNT_SCHED_TICKS,    // interleaved kernel: wall-clock ticks in scheduling phases, summed over waves
    CNT_NONCANON,       // CONVERGED/LEARNED: solves that kept their root estimate (no canonical cell)
    CNT_COUNT
};

// status codes the kernels can raise
enum { KSTATUS_OK = 0, KSTATUS_TIMEOUT = 1 };

// Bit pattern of a not-yet-written outflow in the dataflow history: a
// signaling NaN payload.  Arithmetic only produces quiet NaNs and outflows
// are always computed values, so no routed result can carry it.
static constexpr unsigned long long kPending = 0x7FF4A5A5A5A5A5A5ull;

// Upstream links per unit the dataflow kernel holds in registers; networks
// with wider confluences use the barrier schedule.
static constexpr int kDataflowMaxUp = 4;

struct RouteArgs {
    int n_units, n_levels, n_steps, n_members, ring, ring_shift, n_reaches;
    int n_dev_segs;           // stride of the per-segment constant table
    int t_base;               // forcing row of local step 0
    int k_begin, k_end;       // super-level range of one persistent launch
    int groups;               // barrier arrival groups
    int max_eval;             // CONVERGED/LEARNED evaluation cap
    int learned_budget;
    double dt, tol;
    long long timeout_ticks;  // barrier spin limit, wall-clock ticks
    const int *level_ptr, *unit_level, *unit_seg, *up_ptr, *up_idx, *unit_flags, *unit_reach;
    const double *segc;       // [SC_COUNT][n_dev_segs]
    const float *segc32;      // the same constants in float arithmetic (FP32 REFERENCE rule)
    const double *qlat;       // [forcing rows][n_dev_segs][M]
    double *q, *d, *v;        // [n_dev_segs][M], updated in place
    double *ring_buf;         // [ring][n_units][M]
    double *hist_out;         // [M][n_steps][n_reaches] or null
    const float *mlp;         // LEARNED weights or null
    unsigned *bar;            // barrier words (reset before each persistent launch)
    int *status;
    unsigned long long *counters;
    const int *t_dev;         // LEVEL schedule: current local step
    unsigned long long *trace_t;  // [super-levels] block 0 barrier release times, or null
    unsigned *trace_it;           // [super-levels] max iterations of an item, or null
    unsigned long long *flow_hist;  // DATAFLOW: [n_steps+1][n_units][M] outflow bits
    const int4 *unit_rec;           // DATAFLOW: [n_units][2] packed unit records
    int batch_iters;                // DATAFLOW interleaved: secant iterations per batch
    int guess_model;                // LEARNED: 0 network, 1 power law alone
    const double *lin;              // LINEAR: [4][n_dev_segs] C1, C2, C3, C4/ql
    // DATAFLOW interleaved, CONVERGED/LEARNED: blocks >= n_route_blocks are
    // server waves that run the certified solves routing lanes hand them
    // (solves whose F(mindepth) <= 0); see k_route_dataflow_il.
    int n_route_blocks;
    int scan_cap;                   // ring entries (a power of two)
    unsigned long long *scan_ctl;   // [0] tickets issued, [32] tickets claimed, [64] routing waves done
    unsigned long long *scan_ring;  // [scan_cap] (ticket + 1) << 20 | lane, 1 once read, 0 never used
    unsigned long long *scan_resp_seq;  // [routing lanes] ticket + 1 once answered
    double *scan_req;               // [routing lanes][5] qup, quc, qdp, ql, device segment
    double *scan_resp;              // [routing lanes][3] root depth, outflow, evaluations
};

// scan queue layout in one allocation of 64-bit words
static constexpr int kScanCtlWords = 96;
static constexpr int kScanLaneBits = 20;      // routing lanes < 2^20

// launchers (trmc_kernels.hip); all asynchronous on `stream`
hipError_t launch_stream_copy(const void *src, void *dst, size_t bytes, hipStream_t stream);
// diffusive wave: one reach's computational nodes and step parameters
struct DiffReach { int ncomp; float theta, dtini; };
// node arrays [nreach][500] floats; ws >= 500 * nreach floats; a reach whose
// ncomp lies outside 1..500 is skipped and counted in *bad (when non-null)
hipError_t launch_diffusive_tridiag(const DiffReach *rp, const float *dx, const float *diffusivity,
                                    const float *qp_prev, float *qp_out, float *ws, int nreach, hipStream_t stream,
                                    unsigned *bad = nullptr);
// params: n ChParams of 8 floats (dx, bw, tw, twcc, n_ch, ncc, cs, s0); out [4][n]
hipError_t launch_linear_coeffs(const float *params, int n, float dt, int fixed_x, float *out, hipStream_t stream);
// matrix-form linearized MC step over the union rows of P_Q and P_q:
// col[k], val[k] = {P_Q value, P_q value}; z[c] = {Q, ql}; blocks[b] = {G in
// 1, 2, 4, ..., 64, first slot, slots <= 256 / G, 0}; slots[i] = {row, first
// entry, end, 0}; writes z_out[row] = {new Q, ql_next[row]}
// max_grid > 0 caps the launched grid below its default of 2^20 blocks (the
// kernel loops over the block table either way)
hipError_t launch_matrix_step(int n_blocks, const int4 *blocks, const int4 *slots, const int *col, const float2 *val,
                              const float2 *z, const float *ql_next, float2 *z_out, hipStream_t stream,
                              int max_grid = 0);
hipError_t launch_gen_forcing(double *dst, const int *seg_of_dev, int nd, int N, int M, int T,
                              unsigned long long seed, hipStream_t stream);
hipError_t launch_reset(unsigned *bar, int nwords, hipStream_t stream);
hipError_t launch_route_skewed(const RouteArgs &a, int mode, int grid, int block,
                               bool cooperative, hipStream_t stream);
hipError_t launch_route_level(const RouteArgs &a, int mode, int level, long long items, int block,
                              hipStream_t stream);
hipError_t launch_step_advance(int *t_dev, hipStream_t stream);
hipError_t launch_route_dataflow(const RouteArgs &a, int mode, int grid, int block, hipStream_t stream);
hipError_t route_dataflow_occupancy(int mode, int block, int *blocks_per_cu);
// secant iterations interleaved across lanes (every mode)
hipError_t launch_route_dataflow_il(const RouteArgs &a, int mode, int grid, int block, hipStream_t stream);
hipError_t route_dataflow_il_occupancy(int mode, int block, int *blocks_per_cu);
// dynamic LDS of the interleaved kernel for `block` lanes: the converged
// rules keep the first kIlCacheSlots segments of a lane's current unit there
static constexpr int kIlCacheSlots = 2;
size_t dataflow_il_lds_bytes(int mode, int block);
// history slot 0 <- ring slot `src_slot`; slots 1..T <- kPending
hipError_t launch_hist_init(unsigned long long *hist, const double *ring, int src_slot, long long um,
                            int T, hipStream_t stream);
// ring slot `dst_slot` <- history slot T; reach outflows -> hist_out when non-null
hipError_t launch_hist_finish(const unsigned long long *hist, double *ring, int dst_slot, long long um,
                              int T, hipStream_t stream);
hipError_t launch_hist_extract(const unsigned long long *hist, const int *reach_last_unit, int R, int U,
                               int M, int T, double *out, hipStream_t stream);
hipError_t route_skewed_occupancy(int mode, int block, int *blocks_per_cu);
hipError_t launch_reach_parallel(const double *segc, int n_seg, const double *ql,
                                 const int *rs, const int *rl, const double *qu,
                                 const double *qd, const double *dp, double *qo, double *vo,
                                 double *dpo, double uq, double dt, int nr, hipStream_t stream);
hipError_t launch_solve_batch(const double *segc, int n, const double *in, double *out,
                              int *iters, double dt, int mode, double tol, int max_eval,
                              hipStream_t stream);
// forcing scatter: dst[(t*nd + ds)*M + m] = src[t*n_src + seg_of_dev[ds]] for one member
hipError_t launch_scatter_forcing_f64(const double *src, int n_src, int T, const int *seg_of_dev,
                                      int nd, int M, int m, double *dst, hipStream_t stream);
hipError_t launch_scatter_forcing_f32(const float *src, int n_src, int T, const int *seg_of_dev,
                                      int nd, int M, int m, double *dst, hipStream_t stream);

}  // namespace trmc
