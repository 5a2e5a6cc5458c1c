// C-ABI of the routing library: handles, residency, drop-in wrappers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "trmc.h"
#include "trmc_device.h"
#include "trmc_plan.h"
#include "trmc_util.h"

using trmc::fail;

namespace trmc {

static thread_local std::string g_err;

void set_error(const std::string &msg) { g_err = msg; }

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                        \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return fail(TRMC_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

// gfx950 check; the library ships code for gfx950 only.
static int check_device(int dev, int *cu_count, double *wall_khz)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(TRMC_ERR_NO_DEVICE, "no HIP device visible");
    if (dev < 0 || dev >= n) return fail(TRMC_ERR_NO_DEVICE, "device %d out of range (%d)", dev, n);
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(TRMC_ERR_NO_DEVICE, "device %d is %s, library built for gfx950", dev, prop.gcnArchName);
    HIPCHK(hipSetDevice(dev));
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (wall_khz) {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            khz = 100000;
        *wall_khz = (double)khz;
    }
    return TRMC_OK;
}

// Per-segment constant table in device segment order, SoA [SC_COUNT][nd].
static void build_segc(const trmc_chparams_f64 *p, const std::vector<int> &seg_of_dev,
                       std::vector<double> &out)
{
    const size_t nd = seg_of_dev.size();
    out.assign((size_t)SC_COUNT * nd, 0.0);
    for (size_t ds = 0; ds < nd; ++ds) {
        const trmc_chparams_f64 &s = p[seg_of_dev[ds]];
        double z = (s.cs == 0.0) ? 1.0 : 1.0 / s.cs;
        double bfd;
        if (s.bw > s.tw) bfd = s.bw / 0.00001;
        else if (s.bw == s.tw) bfd = s.bw / (2.0 * z);
        else bfd = (s.tw - s.bw) / (2.0 * z);
        out[SC_DX * nd + ds] = s.dx;
        out[SC_BW * nd + ds] = s.bw;
        out[SC_TWCC * nd + ds] = s.twcc;
        out[SC_N * nd + ds] = s.n_ch;
        out[SC_NCC * nd + ds] = s.ncc;
        out[SC_S0 * nd + ds] = s.s0;
        out[SC_Z * nd + ds] = z;
        out[SC_BFD * nd + ds] = bfd;
        out[SC_KS * nd + ds] = std::sqrt(s.s0) / s.n_ch;
        out[SC_SQ * nd + ds] = std::sqrt(1.0 + z * z);
        out[SC_SQS0 * nd + ds] = std::sqrt(s.s0);
    }
}

// The same table for the FP32 REFERENCE rule: the parameters rounded to float
// (the reference's ChParams are float, wavefront_mc.cu:41-44) and the derived
// constants in float arithmetic, as its float solve computes them inline
// (z and bankfull depth :73-76, sqrtf(s0)/n_ch :107, sqrtf(1+z*z) :58,
// sqrtf(s0) :124).
static void build_segc32(const trmc_chparams_f64 *p, const std::vector<int> &seg_of_dev,
                         std::vector<float> &out)
{
    const size_t nd = seg_of_dev.size();
    out.assign((size_t)SC_COUNT * nd, 0.0f);
    for (size_t ds = 0; ds < nd; ++ds) {
        const trmc_chparams_f64 &w = p[seg_of_dev[ds]];
        const float dx = (float)w.dx, bw = (float)w.bw, tw = (float)w.tw, twcc = (float)w.twcc;
        const float n_ch = (float)w.n_ch, ncc = (float)w.ncc, cs = (float)w.cs, s0 = (float)w.s0;
        const float z = (cs == 0.0f) ? 1.0f : 1.0f / cs;
        float bfd;
        if (bw > tw) bfd = bw / 0.00001f;
        else if (bw == tw) bfd = bw / (2.0f * z);
        else bfd = (tw - bw) / (2.0f * z);
        out[SC_DX * nd + ds] = dx;
        out[SC_BW * nd + ds] = bw;
        out[SC_TWCC * nd + ds] = twcc;
        out[SC_N * nd + ds] = n_ch;
        out[SC_NCC * nd + ds] = ncc;
        out[SC_S0 * nd + ds] = s0;
        out[SC_Z * nd + ds] = z;
        out[SC_BFD * nd + ds] = bfd;
        out[SC_KS * nd + ds] = sqrtf(s0) / n_ch;
        out[SC_SQ * nd + ds] = sqrtf(1.0f + z * z);
        out[SC_SQS0 * nd + ds] = sqrtf(s0);
    }
}

template <typename T>
static int dev_upload(T **dptr, const T *host, size_t n, hipStream_t s)
{
    if (*dptr) { (void)hipFree(*dptr); *dptr = nullptr; }
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    if (hipMalloc((void **)dptr, bytes) != hipSuccess) {
        *dptr = nullptr;
        return fail(TRMC_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    if (n) HIPCHK(hipMemcpyAsync(*dptr, host, n * sizeof(T), hipMemcpyHostToDevice, s));
    return TRMC_OK;
}

template <typename T>
static int dev_alloc(T **dptr, size_t n)
{
    if (*dptr) { (void)hipFree(*dptr); *dptr = nullptr; }
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    if (hipMalloc((void **)dptr, bytes) != hipSuccess) {
        *dptr = nullptr;
        return fail(TRMC_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    return TRMC_OK;
}

template <typename T>
static void dev_free(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace trmc

using namespace trmc;

// ------------------------------------------------------------------ handle
struct trmc_network {
    trmc_options opt;
    Plan plan;
    std::vector<int> reach_last_seg;   // original last segment of each reach or -1
    int M = 1;
    int cu_count = 0;
    double wall_khz = 100000.0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int *d_level_ptr = nullptr, *d_unit_level = nullptr, *d_unit_seg = nullptr, *d_up_ptr = nullptr,
        *d_up_idx = nullptr, *d_unit_flags = nullptr, *d_unit_reach = nullptr, *d_seg_of_dev = nullptr,
        *d_unit_rec = nullptr;
    double *d_segc = nullptr, *d_qlat = nullptr, *d_q = nullptr, *d_d = nullptr, *d_v = nullptr,
           *d_ring = nullptr, *d_hist = nullptr;
    double *d_q0 = nullptr, *d_d0 = nullptr, *d_ring0 = nullptr;   // initial state for reset
    unsigned long long *d_trace_t = nullptr, *d_flow_hist = nullptr;
    size_t flow_hist_n = 0;
    unsigned *d_trace_it = nullptr;
    int trace_n = 0;
    float *d_mlp = nullptr;
    float *d_segc32 = nullptr;           // FP32 REFERENCE constants [SC_COUNT][n_dev_segs]
    double *d_lin = nullptr;             // LINEAR coefficients [4][n_dev_segs]
    unsigned long long *d_scanq = nullptr;   // certified-solve queue: control words, ring, answer tickets
    double *d_scan_data = nullptr;           // requests [lanes][5], answers [lanes][3]
    size_t scan_lanes = 0;
    int scan_cap = 0;
    unsigned *d_bar = nullptr;
    // forcing read in place from page-locked host memory (not owned), or null
    const double *host_qlat = nullptr;
    int *d_status = nullptr, *d_t = nullptr;
    unsigned long long *d_counters = nullptr;
    int forcing_T = 0, t_cursor = 0, ring_shift = 0, hist_T = 0;
    size_t qlat_cap = 0;                 // doubles allocated at d_qlat
    long long steps_done = 0;
    bool have_params = false, have_state = false, have_lin = false;
    std::vector<double> init_q, init_d;  // [M][N] host copies (segments outside every reach)

    ~trmc_network()
    {
        dev_free(d_level_ptr); dev_free(d_unit_level); dev_free(d_unit_seg); dev_free(d_up_ptr);
        dev_free(d_up_idx); dev_free(d_unit_flags); dev_free(d_unit_reach); dev_free(d_seg_of_dev);
        dev_free(d_unit_rec);
        dev_free(d_segc); dev_free(d_qlat); dev_free(d_q); dev_free(d_d); dev_free(d_v);
        dev_free(d_ring); dev_free(d_hist); dev_free(d_mlp); dev_free(d_bar); dev_free(d_status);
        dev_free(d_t); dev_free(d_counters); dev_free(d_q0); dev_free(d_d0); dev_free(d_ring0);
        dev_free(d_trace_t); dev_free(d_trace_it); dev_free(d_flow_hist); dev_free(d_lin);
        dev_free(d_scanq); dev_free(d_scan_data); dev_free(d_segc32);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

static const int kBarWords = 64 + 32 * 64;   // up to 64 arrival groups

extern "C" {

const char *trmc_last_error(void) { return g_err.c_str(); }

int trmc_version(void) { return 10000; }

int trmc_device_count(int *count)
{
    if (!count) return fail(TRMC_ERR_ARG, "null count");
    *count = 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return TRMC_OK;
    for (int d = 0; d < n; ++d) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, d) == hipSuccess &&
            std::strncmp(prop.gcnArchName, "gfx950", 6) == 0)
            ++*count;
    }
    return TRMC_OK;
}

void trmc_default_options(trmc_options *opt)
{
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->mode = TRMC_MODE_REFERENCE;
    opt->schedule = TRMC_SCHED_AUTO;
    opt->unit_segments = 0;
    opt->block_size = 0;
    opt->n_members = 1;
    opt->record_history = 0;
    opt->tol = 1e-10;
    opt->max_iter = 60;
    opt->learned_budget = 12;
    opt->grid_blocks = 0;
    opt->dataflow_variant = 0;
    opt->device = 0;
    opt->trace = 0;
    opt->batch_iters = 0;
    opt->guess_model = 0;
}

int trmc_device_synchronize(int device)
{
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipDeviceSynchronize());
    return TRMC_OK;
}

// ------------------------------------------------------------ host plans
struct trmc_plan { Plan p; };

int trmc_plan_build(const int *reach_seg_start, const int *reach_seg_len, const int *reach_up_ptr,
                    const int *reach_up_idx, const int *level_ptr, const int *level_reach,
                    int n_reaches, int n_levels, int n_segments, int unit_segments, trmc_plan **out)
{
    if (!out) return fail(TRMC_ERR_ARG, "null out");
    *out = nullptr;
    std::unique_ptr<trmc_plan> h(new (std::nothrow) trmc_plan);
    if (!h) return fail(TRMC_ERR_NOMEM, "out of host memory");
    std::string err;
    int rc = build_plan(reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx, level_ptr,
                        level_reach, n_reaches, n_levels, n_segments, unit_segments, h->p, err);
    if (rc) return fail(rc, "%s", err.c_str());
    *out = h.release();
    return TRMC_OK;
}

void trmc_plan_destroy(trmc_plan *p) { delete p; }

static void fill_info(const Plan &p, int T, int M, int64_t *info, int n_info)
{
    int64_t v[8] = {p.n_units, p.n_levels, p.ring, T > 0 ? (int64_t)p.n_levels + T - 1 : 0,
                    max_window_items(p, T, M), p.n_dev_segments, p.max_span,
                    (int64_t)p.up_idx.size()};
    for (int i = 0; i < n_info && i < 8; ++i) info[i] = v[i];
}

int trmc_plan_info(const trmc_plan *p, int n_timesteps, int n_members, int64_t *info, int n_info)
{
    if (!p || !info || n_info < 0) return fail(TRMC_ERR_ARG, "null argument");
    fill_info(p->p, n_timesteps, n_members, info, n_info);
    return TRMC_OK;
}

int trmc_plan_export(const trmc_plan *h, int which, int *out, int n)
{
    if (!h || !out) return fail(TRMC_ERR_ARG, "null argument");
    const Plan &p = h->p;
    const std::vector<int> *src[] = {&p.level_ptr, &p.unit_level, &p.unit_seg, &p.up_ptr, &p.up_idx,
                                     &p.unit_flags, &p.unit_reach, &p.unit_init, &p.seg_of_dev,
                                     &p.reach_last_unit, &p.unit_rec};
    if (which < 0 || which >= (int)(sizeof(src) / sizeof(src[0])))
        return fail(TRMC_ERR_ARG, "unknown plan array %d", which);
    const std::vector<int> &v = *src[which];
    if (n < (int)v.size()) return fail(TRMC_ERR_ARG, "plan array %d needs %zu entries", which, v.size());
    std::copy(v.begin(), v.end(), out);
    return (int)v.size();
}

// ------------------------------------------------------------- networks
int trmc_network_create(const int *reach_seg_start, const int *reach_seg_len,
                        const int *reach_up_ptr, const int *reach_up_idx, const int *level_ptr,
                        const int *level_reach, int n_reaches, int n_levels, int n_segments,
                        const trmc_options *opt, trmc_network **out)
{
    if (!out) return fail(TRMC_ERR_ARG, "null out");
    *out = nullptr;
    std::unique_ptr<trmc_network> net(new (std::nothrow) trmc_network);
    if (!net) return fail(TRMC_ERR_NOMEM, "out of host memory");
    if (opt) net->opt = *opt; else trmc_default_options(&net->opt);
    trmc_options &o = net->opt;
    if (o.n_members < 1) return fail(TRMC_ERR_ARG, "n_members must be >= 1");
    if (o.block_size <= 0) o.block_size = 64;
    if (o.block_size % 64 != 0 || o.block_size > 256)
        return fail(TRMC_ERR_ARG, "block_size must be a multiple of 64, at most 256");
    if (o.mode < TRMC_MODE_REFERENCE || o.mode > TRMC_MODE_REFERENCE_F32)
        return fail(TRMC_ERR_ARG, "unknown mode %d", o.mode);
    // measured best on the 1M-segment Shreve network: the REFERENCE rule's
    // short solves want longer units, the converged rules shorter ones (more
    // items side by side), LINEAR (no solve) longer still
    if (o.unit_segments <= 0)
        o.unit_segments = (o.mode == TRMC_MODE_REFERENCE || o.mode == TRMC_MODE_REFERENCE_F32) ? 4
                          : o.mode == TRMC_MODE_LINEAR                                          ? 8
                                                                                                : 2;
    if (o.tol <= 0.0) o.tol = 1e-10;
    if (o.max_iter <= 0) o.max_iter = 60;
    if (o.learned_budget <= 0) o.learned_budget = 12;
    if (o.guess_model != 0 && o.guess_model != 1) return fail(TRMC_ERR_ARG, "guess_model must be 0 or 1");
    net->M = o.n_members;

    std::string err;
    int rc = build_plan(reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx, level_ptr,
                        level_reach, n_reaches, n_levels, n_segments, o.unit_segments, net->plan, err);
    if (rc) return fail(rc, "%s", err.c_str());
    net->reach_last_seg.resize(n_reaches);
    for (int r = 0; r < n_reaches; ++r)
        net->reach_last_seg[r] = reach_seg_len[r] > 0 ? reach_seg_start[r] + reach_seg_len[r] - 1 : -1;

    rc = check_device(o.device, &net->cu_count, &net->wall_khz);
    if (rc) return rc;
    HIPCHK(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&net->ev0));
    HIPCHK(hipEventCreate(&net->ev1));
    const Plan &P = net->plan;
    hipStream_t s = net->stream;
    if ((rc = dev_upload(&net->d_level_ptr, P.level_ptr.data(), P.level_ptr.size(), s))) return rc;
    if ((rc = dev_upload(&net->d_unit_level, P.unit_level.data(), P.unit_level.size(), s))) return rc;
    if ((rc = dev_upload(&net->d_unit_seg, P.unit_seg.data(), P.unit_seg.size(), s))) return rc;
    if ((rc = dev_upload(&net->d_up_ptr, P.up_ptr.data(), P.up_ptr.size(), s))) return rc;
    if ((rc = dev_upload(&net->d_up_idx, P.up_idx.data(), P.up_idx.size(), s))) return rc;
    if ((rc = dev_upload(&net->d_unit_flags, P.unit_flags.data(), P.unit_flags.size(), s))) return rc;
    if ((rc = dev_upload(&net->d_unit_reach, P.unit_reach.data(), P.unit_reach.size(), s))) return rc;
    if ((rc = dev_upload(&net->d_seg_of_dev, P.seg_of_dev.data(), P.seg_of_dev.size(), s))) return rc;
    if (!P.unit_rec.empty() && (rc = dev_upload(&net->d_unit_rec, P.unit_rec.data(), P.unit_rec.size(), s)))
        return rc;
    const size_t nd = (size_t)P.n_dev_segments * net->M;
    if ((rc = dev_alloc(&net->d_q, nd))) return rc;
    if ((rc = dev_alloc(&net->d_d, nd))) return rc;
    if ((rc = dev_alloc(&net->d_v, nd))) return rc;
    if ((rc = dev_alloc(&net->d_ring, (size_t)P.ring * P.n_units * net->M))) return rc;
    if ((rc = dev_alloc(&net->d_bar, (size_t)kBarWords))) return rc;
    if ((rc = dev_alloc(&net->d_status, 4))) return rc;
    if ((rc = dev_alloc(&net->d_t, 4))) return rc;
    if ((rc = dev_alloc(&net->d_counters, (size_t)CNT_COUNT))) return rc;
    HIPCHK(hipStreamSynchronize(s));
    *out = net.release();
    return TRMC_OK;
}

void trmc_network_destroy(trmc_network *net)
{
    if (!net) return;
    (void)hipSetDevice(net->opt.device);
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    delete net;
}

int trmc_network_set_params_f64(trmc_network *net, const trmc_chparams_f64 *params)
{
    if (!net || (!params && net->plan.n_segments > 0)) return fail(TRMC_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(net->opt.device));
    std::vector<double> segc;
    build_segc(params, net->plan.seg_of_dev, segc);
    int rc = dev_upload(&net->d_segc, segc.data(), segc.size(), net->stream);
    if (rc) return rc;
    std::vector<float> segc32;
    if (net->opt.mode == TRMC_MODE_REFERENCE_F32) {
        build_segc32(params, net->plan.seg_of_dev, segc32);
        if ((rc = dev_upload(&net->d_segc32, segc32.data(), segc32.size(), net->stream))) return rc;
    }
    HIPCHK(hipStreamSynchronize(net->stream));
    net->have_params = true;
    return TRMC_OK;
}

int trmc_network_set_params_f32(trmc_network *net, const trmc_chparams_f32 *params)
{
    if (!net || (!params && net->plan.n_segments > 0)) return fail(TRMC_ERR_ARG, "null argument");
    const int n = net->plan.n_segments;
    std::vector<trmc_chparams_f64> wide((size_t)n);
    for (int i = 0; i < n; ++i) {
        const trmc_chparams_f32 &p = params[i];
        wide[i] = {p.dx, p.bw, p.tw, p.twcc, p.n_ch, p.ncc, p.cs, p.s0};
    }
    return trmc_network_set_params_f64(net, wide.data());
}

int trmc_network_set_state_f64(trmc_network *net, const double *qdp0, const double *dp0,
                               const double *reach_last0)
{
    if (!net) return fail(TRMC_ERR_ARG, "null network");
    const Plan &P = net->plan;
    const int N = P.n_segments, R = P.n_reaches, M = net->M, U = P.n_units;
    const int nd = P.n_dev_segments;
    if (N > 0 && (!qdp0 || !dp0)) return fail(TRMC_ERR_ARG, "null initial state");
    HIPCHK(hipSetDevice(net->opt.device));
    std::vector<double> q((size_t)nd * M), d((size_t)nd * M), v((size_t)nd * M, 0.0);
    std::vector<double> ring0((size_t)U * M);
    for (int m = 0; m < M; ++m) {
        const double *qm = qdp0 + (size_t)m * N, *dm = dp0 + (size_t)m * N;
        for (int ds = 0; ds < nd; ++ds) {
            q[(size_t)ds * M + m] = qm[P.seg_of_dev[ds]];
            d[(size_t)ds * M + m] = dm[P.seg_of_dev[ds]];
        }
        for (int u = 0; u < U; ++u) {
            int src = P.unit_init[u];
            double val;
            if (src >= 0) {
                val = qm[src];
            } else {
                int r = -(src + 1);
                if (reach_last0) val = reach_last0[(size_t)m * R + r];
                else val = net->reach_last_seg[r] >= 0 ? qm[net->reach_last_seg[r]] : 0.0;
            }
            ring0[(size_t)u * M + m] = val;
        }
    }
    net->init_q.assign(qdp0, qdp0 + (size_t)N * M);
    net->init_d.assign(dp0, dp0 + (size_t)N * M);
    hipStream_t s = net->stream;
    if (nd > 0) {
        HIPCHK(hipMemcpyAsync(net->d_q, q.data(), q.size() * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(net->d_d, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(net->d_v, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, s));
    }
    if (U > 0)
        HIPCHK(hipMemcpyAsync(net->d_ring, ring0.data(), ring0.size() * sizeof(double),
                              hipMemcpyHostToDevice, s));
    // device-resident copy of the initial state for trmc_network_reset
    int rc;
    if ((rc = dev_upload(&net->d_q0, q.data(), q.size(), s))) return rc;
    if ((rc = dev_upload(&net->d_d0, d.data(), d.size(), s))) return rc;
    if ((rc = dev_upload(&net->d_ring0, ring0.data(), ring0.size(), s))) return rc;
    HIPCHK(hipStreamSynchronize(s));
    net->ring_shift = 0;
    net->steps_done = 0;
    net->have_state = true;
    return TRMC_OK;
}

}  // extern "C"

template <typename TIn>
static int set_forcing_impl(trmc_network *net, const TIn *qlat, int T)
{
    if (!net || T < 0 || (!qlat && T > 0 && net->plan.n_segments > 0))
        return fail(TRMC_ERR_ARG, "bad forcing arguments");
    HIPCHK(hipSetDevice(net->opt.device));
    const Plan &P = net->plan;
    const int N = P.n_segments, nd = P.n_dev_segments, M = net->M;
    int rc = dev_alloc(&net->d_qlat, (size_t)T * nd * M);
    if (rc) return rc;
    net->qlat_cap = (size_t)T * nd * M;
    net->host_qlat = nullptr;
    if ((size_t)T * nd > 0) {
        TIn *raw = nullptr;
        if (hipMalloc((void **)&raw, (size_t)T * N * sizeof(TIn)) != hipSuccess)
            return fail(TRMC_ERR_NOMEM, "hipMalloc of forcing staging failed");
        for (int m = 0; m < M; ++m) {
            hipError_t e = hipMemcpyAsync(raw, qlat + (size_t)m * T * N, (size_t)T * N * sizeof(TIn),
                                          hipMemcpyHostToDevice, net->stream);
            if (e == hipSuccess) {
                if (sizeof(TIn) == sizeof(double))
                    e = launch_scatter_forcing_f64((const double *)raw, N, T, net->d_seg_of_dev, nd, M, m,
                                                   net->d_qlat, net->stream);
                else
                    e = launch_scatter_forcing_f32((const float *)raw, N, T, net->d_seg_of_dev, nd, M, m,
                                                   net->d_qlat, net->stream);
            }
            if (e == hipSuccess) e = hipStreamSynchronize(net->stream);
            if (e != hipSuccess) {
                (void)hipFree(raw);
                return fail(TRMC_ERR_HIP, "forcing upload: %s", hipGetErrorString(e));
            }
        }
        (void)hipFree(raw);
    }
    net->forcing_T = T;
    net->t_cursor = 0;
    return TRMC_OK;
}

extern "C" {

int trmc_network_set_forcing_f64(trmc_network *net, const double *qlat, int n_timesteps)
{
    return set_forcing_impl(net, qlat, n_timesteps);
}

int trmc_network_set_forcing_f32(trmc_network *net, const float *qlat, int n_timesteps)
{
    return set_forcing_impl(net, qlat, n_timesteps);
}

int trmc_network_order_forcing_f64(const trmc_network *net, const double *qlat, int n_timesteps, double *out)
{
    if (!net || n_timesteps < 0 || (n_timesteps > 0 && (!qlat || !out)))
        return fail(TRMC_ERR_ARG, "bad forcing arguments");
    const Plan &P = net->plan;
    const size_t N = (size_t)P.n_segments, nd = (size_t)P.n_dev_segments, M = (size_t)net->M;
    const size_t T = (size_t)n_timesteps;
    for (size_t t = 0; t < T; ++t)
        for (size_t ds = 0; ds < nd; ++ds) {
            const size_t sg = (size_t)P.seg_of_dev[ds];
            double *o = out + (t * nd + ds) * M;
            for (size_t m = 0; m < M; ++m) o[m] = qlat[(m * T + t) * N + sg];
        }
    return TRMC_OK;
}

int trmc_host_alloc(size_t bytes, void **ptr)
{
    if (!ptr) return fail(TRMC_ERR_ARG, "null pointer");
    *ptr = nullptr;
    if (bytes == 0) return TRMC_OK;
    hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(TRMC_ERR_NOMEM, "hipHostMalloc of %zu bytes: %s", bytes, hipGetErrorString(e));
    }
    return TRMC_OK;
}

void trmc_host_free(void *ptr)
{
    if (ptr) (void)hipHostFree(ptr);
}

int trmc_network_set_forcing_mapped(trmc_network *net, const double *qlat_dev_order, int n_timesteps)
{
    if (!net || n_timesteps < 0 || (n_timesteps > 0 && !qlat_dev_order))
        return fail(TRMC_ERR_ARG, "bad forcing arguments");
    HIPCHK(hipSetDevice(net->opt.device));
    // the routing kernels read the forcing where it lies: the device's view
    // of the page-locked host allocation
    void *dp = nullptr;
    if (n_timesteps > 0) {
        hipError_t e = hipHostGetDevicePointer(&dp, const_cast<double *>(qlat_dev_order), 0);
        if (e != hipSuccess || !dp)
            return fail(TRMC_ERR_ARG, "forcing is not page-locked host memory (trmc_host_alloc): %s",
                        hipGetErrorString(e));
    }
    dev_free(net->d_qlat);
    net->qlat_cap = 0;
    net->host_qlat = static_cast<const double *>(dp);
    net->forcing_T = n_timesteps;
    net->t_cursor = 0;
    return TRMC_OK;
}

int trmc_network_generate_forcing(trmc_network *net, int T, uint64_t seed)
{
    if (!net || T < 0) return fail(TRMC_ERR_ARG, "bad forcing arguments");
    HIPCHK(hipSetDevice(net->opt.device));
    const Plan &P = net->plan;
    int rc = dev_alloc(&net->d_qlat, (size_t)T * P.n_dev_segments * net->M);
    if (rc) return rc;
    net->qlat_cap = (size_t)T * P.n_dev_segments * net->M;
    net->host_qlat = nullptr;
    HIPCHK(launch_gen_forcing(net->d_qlat, net->d_seg_of_dev, P.n_dev_segments, P.n_segments, net->M, T,
                              (unsigned long long)seed, net->stream));
    HIPCHK(hipStreamSynchronize(net->stream));
    net->forcing_T = T;
    net->t_cursor = 0;
    return TRMC_OK;
}

int trmc_network_set_linear_coeffs(trmc_network *net, const float *C1, const float *C2, const float *C3,
                                   const float *C4oq)
{
    if (!net) return fail(TRMC_ERR_ARG, "null network");
    const Plan &P = net->plan;
    const int nd = P.n_dev_segments;
    if (nd > 0 && (!C1 || !C2 || !C3 || !C4oq)) return fail(TRMC_ERR_ARG, "null coefficient array");
    HIPCHK(hipSetDevice(net->opt.device));
    // float coefficients widen exactly (linear_mc.cu:247-249), device order
    std::vector<double> lin((size_t)4 * std::max(nd, 1), 0.0);
    const float *src[4] = {C1, C2, C3, C4oq};
    for (int k = 0; k < 4; ++k)
        for (int ds = 0; ds < nd; ++ds) lin[(size_t)k * nd + ds] = (double)src[k][P.seg_of_dev[ds]];
    int rc = dev_upload(&net->d_lin, lin.data(), lin.size(), net->stream);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(net->stream));
    net->have_lin = true;
    return TRMC_OK;
}

int trmc_network_set_mlp(trmc_network *net, const float *weights)
{
    if (!net || !weights) return fail(TRMC_ERR_ARG, "null network or weights");
    for (int i = 0; i < TRMC_MLP_NPARAM; ++i)
        if (!std::isfinite(weights[i])) return fail(TRMC_ERR_ARG, "weight %d is not finite", i);
    HIPCHK(hipSetDevice(net->opt.device));
    int rc = dev_upload(&net->d_mlp, weights, (size_t)TRMC_MLP_NPARAM, net->stream);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(net->stream));
    return TRMC_OK;
}

}  // extern "C"

static RouteArgs make_args(trmc_network *net, int T, double dt)
{
    const Plan &P = net->plan;
    RouteArgs a;
    std::memset(&a, 0, sizeof(a));
    a.n_units = P.n_units;
    a.n_levels = P.n_levels;
    a.n_steps = T;
    a.n_members = net->M;
    a.ring = P.ring;
    a.ring_shift = net->ring_shift;
    a.n_reaches = P.n_reaches;
    a.n_dev_segs = P.n_dev_segments;
    a.t_base = net->t_cursor;
    a.k_begin = 0;
    a.k_end = T > 0 ? P.n_levels + T - 1 : 0;
    a.groups = 16;
    a.max_eval = net->opt.max_iter;
    a.learned_budget = net->opt.learned_budget;
    a.dt = dt;
    a.tol = net->opt.tol;
    a.timeout_ticks = (long long)(net->wall_khz * 1000.0 * 30.0);   // 30 s per barrier wait
    a.level_ptr = net->d_level_ptr;
    a.unit_level = net->d_unit_level;
    a.unit_seg = net->d_unit_seg;
    a.up_ptr = net->d_up_ptr;
    a.up_idx = net->d_up_idx;
    a.unit_flags = net->d_unit_flags;
    a.unit_reach = net->d_unit_reach;
    a.segc = net->d_segc;
    a.segc32 = net->d_segc32;
    a.qlat = net->host_qlat ? net->host_qlat : net->d_qlat;
    a.q = net->d_q;
    a.d = net->d_d;
    a.v = net->d_v;
    a.ring_buf = net->d_ring;
    a.hist_out = net->opt.record_history ? net->d_hist : nullptr;
    a.mlp = net->d_mlp;
    a.lin = net->d_lin;
    a.bar = net->d_bar;
    a.status = net->d_status;
    a.counters = net->d_counters;
    a.t_dev = net->d_t;
    a.unit_rec = reinterpret_cast<const int4 *>(net->d_unit_rec);
    a.batch_iters = net->opt.batch_iters > 0 ? net->opt.batch_iters : 8;
    a.guess_model = net->opt.guess_model;
    return a;
}

extern "C" {

int trmc_network_run(trmc_network *net, int T, double dt, trmc_run_stats *stats)
{
    if (!net) return fail(TRMC_ERR_ARG, "null network");
    if (T < 0) return fail(TRMC_ERR_ARG, "negative n_timesteps");
    if (net->opt.mode == TRMC_MODE_LINEAR) {
        if (!net->have_lin || !net->have_state)
            return fail(TRMC_ERR_STATE, "LINEAR mode needs coefficients (trmc_network_set_linear_coeffs) and state");
    } else if (!net->have_params || !net->have_state) {
        return fail(TRMC_ERR_STATE, "params and state must be set");
    }
    if (net->t_cursor + T > net->forcing_T)
        return fail(TRMC_ERR_STATE, "forcing has %d steps left, %d requested",
                    net->forcing_T - net->t_cursor, T);
    HIPCHK(hipSetDevice(net->opt.device));
    const auto w0 = std::chrono::steady_clock::now();
    const Plan &P = net->plan;
    const trmc_options &o = net->opt;
    trmc_run_stats st;
    std::memset(&st, 0, sizeof(st));
    st.block_size = o.block_size;
    hipStream_t s = net->stream;
    int rc;
    if (o.record_history) {
        if ((rc = dev_alloc(&net->d_hist, (size_t)net->M * T * P.n_reaches))) return rc;
        net->hist_T = T;
    }
    if (T == 0 || P.n_units == 0) {
        net->t_cursor += T;
        net->steps_done += T;
        if (stats) *stats = st;
        return TRMC_OK;
    }
    HIPCHK(launch_reset((unsigned *)net->d_counters, 2 * CNT_COUNT, s));
    HIPCHK(launch_reset((unsigned *)net->d_status, 1, s));
    RouteArgs a = make_args(net, T, dt);
    const int mode = o.mode;
    if (mode == TRMC_MODE_LEARNED && o.guess_model == 0 && !net->d_mlp)
        return fail(TRMC_ERR_STATE, "LEARNED mode needs weights (trmc_network_set_mlp)");
    // Forcing read from host memory: the REFERENCE, LINEAR and REFERENCE_F32
    // rules read a segment's inflow as they open it.  The converged rules'
    // interleaved kernel loads it ahead, with its unit cache, and gave wrong
    // outflows at some confluences with host-memory inflow (an unexplained
    // failure): refused for them.
    if (net->host_qlat && (mode == TRMC_MODE_CONVERGED || mode == TRMC_MODE_LEARNED))
        return fail(TRMC_ERR_ARG, "forcing in host memory is not supported for the CONVERGED and LEARNED rules");
    int sched = o.schedule;
    if (sched == TRMC_SCHED_AUTO)
        sched = P.max_upstream <= kDataflowMaxUp ? TRMC_SCHED_DATAFLOW : TRMC_SCHED_SKEWED;
    if (mode == TRMC_MODE_LINEAR && (sched != TRMC_SCHED_DATAFLOW || o.dataflow_variant != 0))
        return fail(TRMC_ERR_ARG, "LINEAR mode runs on the interleaved DATAFLOW schedule");
    if (mode == TRMC_MODE_REFERENCE_F32 && (sched != TRMC_SCHED_DATAFLOW || o.dataflow_variant != 0))
        return fail(TRMC_ERR_ARG, "REFERENCE_F32 mode runs on the interleaved DATAFLOW schedule");
    if (sched == TRMC_SCHED_DATAFLOW && P.max_upstream > kDataflowMaxUp)
        return fail(TRMC_ERR_ARG, "DATAFLOW schedule supports confluences of at most %d reaches (network has %d)",
                    kDataflowMaxUp, P.max_upstream);
    if (sched == TRMC_SCHED_DATAFLOW) {
        const size_t um = (size_t)P.n_units * net->M;
        const size_t need = um * ((size_t)T + 1);
        if (need > net->flow_hist_n) {
            if ((rc = dev_alloc(&net->d_flow_hist, need))) return rc;
            net->flow_hist_n = need;
        }
        const bool il = (o.dataflow_variant == 0);
        int bpcu = 0;
        if (il) HIPCHK(route_dataflow_il_occupancy(mode, o.block_size, &bpcu));
        else HIPCHK(route_dataflow_occupancy(mode, o.block_size, &bpcu));
        if (bpcu <= 0) return fail(TRMC_ERR_HIP, "dataflow kernel cannot be resident");
        long long resident = (long long)bpcu * net->cu_count;
        // CONVERGED/LEARNED on the interleaved kernel: the last blocks serve
        // the certified solves (one per 4 CUs, at least 2), resident with the rest
        const bool servers = il && (mode == TRMC_MODE_CONVERGED || mode == TRMC_MODE_LEARNED);
        const long long n_servers = servers ? std::max(2, net->cu_count / 4) : 0;
        if (resident <= n_servers) return fail(TRMC_ERR_HIP, "dataflow kernel cannot be resident");
        long long need_blocks = ((long long)um + o.block_size - 1) / o.block_size;
        long long grid = std::max(1LL, std::min(need_blocks, resident - n_servers));
        if (o.grid_blocks > 0) grid = std::min<long long>(o.grid_blocks, resident - n_servers);
        // every lane owns one member of a unit sequence: nthreads % M == 0
        while (grid > 1 && ((long long)grid * o.block_size) % net->M != 0) --grid;
        if (((long long)grid * o.block_size) % net->M != 0)
            return fail(TRMC_ERR_ARG, "n_members %d does not divide the dataflow grid", net->M);
        a.n_route_blocks = (int)grid;
        if (servers) {
            const size_t lanes = (size_t)grid * o.block_size;
            if (lanes >= ((size_t)1 << kScanLaneBits))
                return fail(TRMC_ERR_ARG, "dataflow grid of %zu lanes exceeds the certified-solve queue", lanes);
            int cap = 1024;
            while ((size_t)cap < 2 * lanes) cap *= 2;
            if (lanes > net->scan_lanes || cap > net->scan_cap) {
                if ((rc = dev_alloc(&net->d_scanq, (size_t)kScanCtlWords + cap + lanes))) return rc;
                if ((rc = dev_alloc(&net->d_scan_data, lanes * 8))) return rc;
                net->scan_lanes = lanes;
                net->scan_cap = cap;
            }
            // control words, ring and answer tickets start at 0 in every launch
            HIPCHK(launch_reset((unsigned *)net->d_scanq, (int)(2 * ((size_t)kScanCtlWords + cap + lanes)), s));
            a.scan_cap = cap;
            a.scan_ctl = net->d_scanq;
            a.scan_ring = net->d_scanq + kScanCtlWords;
            a.scan_resp_seq = net->d_scanq + kScanCtlWords + cap;
            a.scan_req = net->d_scan_data;
            a.scan_resp = net->d_scan_data + lanes * 5;
            grid += n_servers;
        }
        a.flow_hist = net->d_flow_hist;
        HIPCHK(launch_hist_init(net->d_flow_hist, net->d_ring, net->ring_shift % P.ring, (long long)um, T, s));
        HIPCHK(hipEventRecord(net->ev0, s));
        if (il) HIPCHK(launch_route_dataflow_il(a, mode, (int)grid, o.block_size, s));
        else HIPCHK(launch_route_dataflow(a, mode, (int)grid, o.block_size, s));
        HIPCHK(hipEventRecord(net->ev1, s));
        HIPCHK(launch_hist_finish(net->d_flow_hist, net->d_ring, (T + net->ring_shift) % P.ring, (long long)um, T, s));
        if (o.record_history) {
            int *d_rlu = nullptr;
            if ((rc = dev_upload(&d_rlu, P.reach_last_unit.data(), P.reach_last_unit.size(), s))) return rc;
            hipError_t e = launch_hist_extract(net->d_flow_hist, d_rlu, P.n_reaches, P.n_units, net->M, T,
                                               net->d_hist, s);
            if (e == hipSuccess) e = hipStreamSynchronize(s);
            (void)hipFree(d_rlu);
            if (e != hipSuccess) return fail(TRMC_ERR_HIP, "history extract: %s", hipGetErrorString(e));
        }
        st.launches = 1;
        st.super_levels = P.n_levels + T - 1;
        st.grid_blocks = (int)grid;
    } else if (sched == TRMC_SCHED_SKEWED) {
        int bpcu = 0;
        HIPCHK(route_skewed_occupancy(mode, o.block_size, &bpcu));
        if (bpcu <= 0) return fail(TRMC_ERR_HIP, "routing kernel cannot be resident");
        const long long resident = (long long)bpcu * net->cu_count;
        const long long need = (max_window_items(P, T, net->M) + o.block_size - 1) / o.block_size;
        long long grid = std::max(1LL, std::min(need, resident));
        if (o.grid_blocks > 0) grid = std::min<long long>(o.grid_blocks, resident);
        HIPCHK(launch_reset(net->d_bar, kBarWords, s));
        if (o.trace) {
            const int ns = a.k_end - a.k_begin;
            if ((rc = dev_alloc(&net->d_trace_t, (size_t)ns + 1))) return rc;
            if ((rc = dev_alloc(&net->d_trace_it, (size_t)ns))) return rc;
            HIPCHK(launch_reset(net->d_trace_it, ns, s));
            a.trace_t = net->d_trace_t;
            a.trace_it = net->d_trace_it;
            net->trace_n = ns;
        }
        HIPCHK(hipEventRecord(net->ev0, s));
        HIPCHK(launch_route_skewed(a, mode, (int)grid, o.block_size, true, s));
        HIPCHK(hipEventRecord(net->ev1, s));
        st.launches = 1;
        st.super_levels = a.k_end - a.k_begin;
        st.barriers = std::max(0, st.super_levels - 1);
        st.grid_blocks = (int)grid;
    } else if (sched == TRMC_SCHED_LEVEL) {
        // one step = one launch per level plus the step counter bump, captured
        // once and replayed T times
        HIPCHK(launch_reset((unsigned *)net->d_t, 1, s));
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t ce = hipSuccess;
        for (int L = 0; L < P.n_levels && ce == hipSuccess; ++L) {
            long long items = (long long)(P.level_ptr[L + 1] - P.level_ptr[L]) * net->M;
            ce = launch_route_level(a, mode, L, items, o.block_size, s);
        }
        if (ce == hipSuccess) ce = launch_step_advance(net->d_t, s);
        hipError_t ee = hipStreamEndCapture(s, &graph);
        if (ce != hipSuccess || ee != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return fail(TRMC_ERR_HIP, "graph capture failed: %s",
                        hipGetErrorString(ce != hipSuccess ? ce : ee));
        }
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return fail(TRMC_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ie));
        hipError_t le = hipEventRecord(net->ev0, s);
        for (int t = 0; t < T && le == hipSuccess; ++t) le = hipGraphLaunch(exec, s);
        if (le == hipSuccess) le = hipEventRecord(net->ev1, s);
        if (le == hipSuccess) le = hipStreamSynchronize(s);
        (void)hipGraphExecDestroy(exec);
        if (le != hipSuccess) return fail(TRMC_ERR_HIP, "graph launch: %s", hipGetErrorString(le));
        st.launches = T;
    } else {
        return fail(TRMC_ERR_ARG, "unknown schedule %d", o.schedule);
    }
    HIPCHK(hipStreamSynchronize(s));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, net->ev0, net->ev1));
    int status = 0;
    unsigned long long cnt[CNT_COUNT];
    HIPCHK(hipMemcpy(&status, net->d_status, sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cnt, net->d_counters, sizeof(cnt), hipMemcpyDeviceToHost));
    if (a.scan_ctl) {
        unsigned long long issued = 0;
        HIPCHK(hipMemcpy(&issued, a.scan_ctl, sizeof(issued), hipMemcpyDeviceToHost));
        st.certified_solves = (long long)issued;
        if (issued && std::getenv("TRMC_DEBUG_SCAN"))
            std::fprintf(stderr, "[trmc] certified solves %llu: request-to-answer %.1f us, server solve %.1f us (mean)\n",
                         issued, cnt[CNT_SCAN_WAIT_TICKS] / (net->wall_khz * 1e-3) / issued,
                         cnt[CNT_SCAN_SOLVE_TICKS] / (net->wall_khz * 1e-3) / issued);
    }
    if (cnt[CNT_TRIPS] && std::getenv("TRMC_DEBUG_TRIPS")) {
        const double us = 1e3 / net->wall_khz;      // microseconds per wall-clock tick
        std::fprintf(stderr,
                     "[trmc] trips %llu (%.2f us each, %.2f us of it scheduling), solve trips %llu (%.2f us each), "
                     "lanes per solve trip %.1f, wave time %.1f ms summed\n",
                     cnt[CNT_TRIPS], cnt[CNT_BLOCK_TICKS] * us * (o.block_size / 64.0) / (double)cnt[CNT_TRIPS],
                     cnt[CNT_SCHED_TICKS] * us / (double)cnt[CNT_TRIPS], cnt[CNT_STEP_TRIPS],
                     cnt[CNT_STEP_TRIPS] ? cnt[CNT_BLOCK_TICKS] * us * (o.block_size / 64.0) / cnt[CNT_STEP_TRIPS]
                                         : 0.0,
                     cnt[CNT_STEP_TRIPS] ? (double)cnt[CNT_STEP_LANES] / cnt[CNT_STEP_TRIPS] : 0.0,
                     cnt[CNT_BLOCK_TICKS] * us * 1e-3);
    }
    if (status == KSTATUS_TIMEOUT)
        return fail(TRMC_ERR_TIMEOUT, "persistent routing kernel: barrier wait exceeded 30 s");
    net->t_cursor += T;
    net->ring_shift = (net->ring_shift + T) % P.ring;
    net->steps_done += T;
    st.kernel_ms = ms;
    st.seg_steps = (long long)P.n_dev_segments * net->M * T;
    st.secant_iters = (long long)cnt[CNT_ITERS];
    st.not_converged = (long long)cnt[CNT_CAPPED];
    st.fallbacks = (long long)cnt[CNT_FALLBACK];
    st.noncanonical = (long long)cnt[CNT_NONCANON];
    st.barrier_wait_frac = cnt[CNT_BLOCK_TICKS] ? (double)cnt[CNT_WAIT_TICKS] / (double)cnt[CNT_BLOCK_TICKS] : 0.0;
    st.mlp_calls = (long long)cnt[CNT_MLP_CALLS];
    st.mlp_frac = cnt[CNT_BLOCK_TICKS]
                      ? (double)cnt[CNT_MLP_TICKS] / ((double)cnt[CNT_BLOCK_TICKS] * (o.block_size / 64))
                      : 0.0;
    st.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    if (stats) *stats = st;
    return TRMC_OK;
}

int trmc_network_reset(trmc_network *net)
{
    if (!net) return fail(TRMC_ERR_ARG, "null network");
    if (!net->have_state) return fail(TRMC_ERR_STATE, "state not set");
    HIPCHK(hipSetDevice(net->opt.device));
    const Plan &P = net->plan;
    const size_t nd = (size_t)P.n_dev_segments * net->M, nu = (size_t)P.n_units * net->M;
    hipStream_t s = net->stream;
    if (nd) {
        HIPCHK(hipMemcpyAsync(net->d_q, net->d_q0, nd * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(net->d_d, net->d_d0, nd * sizeof(double), hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemsetAsync(net->d_v, 0, nd * sizeof(double), s));
    }
    if (nu) HIPCHK(hipMemcpyAsync(net->d_ring, net->d_ring0, nu * sizeof(double), hipMemcpyDeviceToDevice, s));
    net->ring_shift = 0;
    net->steps_done = 0;
    net->t_cursor = 0;
    return TRMC_OK;
}

int trmc_network_get_state_f64(const trmc_network *net, double *qdc, double *depth, double *vel,
                               double *reach_last)
{
    if (!net) return fail(TRMC_ERR_ARG, "null network");
    if (!net->have_state) return fail(TRMC_ERR_STATE, "state not set");
    HIPCHK(hipSetDevice(net->opt.device));
    const Plan &P = net->plan;
    const int N = P.n_segments, R = P.n_reaches, M = net->M, nd = P.n_dev_segments, U = P.n_units;
    std::vector<double> buf;
    double *outs[3] = {qdc, depth, vel};
    double *devs[3] = {net->d_q, net->d_d, net->d_v};
    for (int k = 0; k < 3; ++k) {
        if (!outs[k]) continue;
        buf.resize((size_t)nd * M);
        if (nd) HIPCHK(hipMemcpy(buf.data(), devs[k], buf.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int m = 0; m < M; ++m)
            for (int sgi = 0; sgi < N; ++sgi) {
                int ds = P.dev_of_seg[sgi];
                double val;
                if (ds >= 0) val = buf[(size_t)ds * M + m];
                else if (net->steps_done > 0) val = 0.0;   // zero-filled by the reference roll
                else val = k == 0 ? net->init_q[(size_t)m * N + sgi] : k == 1 ? net->init_d[(size_t)m * N + sgi] : 0.0;
                outs[k][(size_t)m * N + sgi] = val;
            }
    }
    if (reach_last && R > 0) {
        buf.resize((size_t)U * M);
        const double *slot = net->d_ring + (size_t)(net->ring_shift % P.ring) * U * M;
        HIPCHK(hipMemcpy(buf.data(), slot, buf.size() * sizeof(double), hipMemcpyDeviceToHost));
        for (int m = 0; m < M; ++m)
            for (int r = 0; r < R; ++r)
                reach_last[(size_t)m * R + r] = buf[(size_t)P.reach_last_unit[r] * M + m];
    }
    return TRMC_OK;
}

int trmc_network_get_history_f64(const trmc_network *net, double *reach_hist)
{
    if (!net || !reach_hist) return fail(TRMC_ERR_ARG, "null argument");
    if (!net->opt.record_history || !net->d_hist) return fail(TRMC_ERR_STATE, "history not recorded");
    HIPCHK(hipSetDevice(net->opt.device));
    size_t n = (size_t)net->M * net->hist_T * net->plan.n_reaches;
    if (n) HIPCHK(hipMemcpy(reach_hist, net->d_hist, n * sizeof(double), hipMemcpyDeviceToHost));
    return TRMC_OK;
}

int trmc_network_get_trace(const trmc_network *net, double *us, int *max_iters, int n)
{
    if (!net || n < 0) return fail(TRMC_ERR_ARG, "bad argument");
    if (!net->opt.trace || !net->d_trace_t || net->trace_n == 0) return fail(TRMC_ERR_STATE, "no trace recorded");
    HIPCHK(hipSetDevice(net->opt.device));
    const int ns = net->trace_n;
    std::vector<unsigned long long> t((size_t)ns + 1);
    std::vector<unsigned> it((size_t)ns);
    HIPCHK(hipMemcpy(t.data(), net->d_trace_t, t.size() * sizeof(t[0]), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(it.data(), net->d_trace_it, it.size() * sizeof(it[0]), hipMemcpyDeviceToHost));
    const int m = std::min(n, ns);
    unsigned long long prev = t[(size_t)ns];
    for (int k = 0; k < m; ++k) {
        if (us) us[k] = (double)(t[k] - prev) * 1000.0 / net->wall_khz;
        if (max_iters) max_iters[k] = (int)it[k];
        prev = t[k];
    }
    return m;
}

int trmc_network_info(const trmc_network *net, int n_timesteps, int64_t *info, int n_info)
{
    if (!net || !info || n_info < 0) return fail(TRMC_ERR_ARG, "null argument");
    fill_info(net->plan, n_timesteps, net->M, info, n_info);
    return TRMC_OK;
}

// --------------------------------------------------------- drop-in calls
int trmc_wavefront_f64(const trmc_chparams_f64 *params, const int *reach_seg_start,
                       const int *reach_seg_len, const int *reach_up_ptr, const int *reach_up_idx,
                       const int *level_ptr, const int *level_reach, int n_reaches, int n_levels,
                       int n_segments, const double *qlat_ts, const double *qdp0_seg,
                       const double *dp0_seg, const double *reach_last0, double *qdc_seg_out,
                       double *reach_last_out, int n_timesteps, double dt, const trmc_options *opt)
{
    if (n_timesteps < 0) return fail(TRMC_ERR_ARG, "negative n_timesteps");
    if ((n_segments > 0 && !qdc_seg_out) || (n_reaches > 0 && !reach_last_out))
        return fail(TRMC_ERR_ARG, "null output");
    trmc_options o;
    if (opt) o = *opt; else trmc_default_options(&o);
    o.n_members = 1;
    trmc_network *net = nullptr;
    int rc = trmc_network_create(reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx,
                                 level_ptr, level_reach, n_reaches, n_levels, n_segments, &o, &net);
    if (rc) return rc;
    std::unique_ptr<trmc_network, void (*)(trmc_network *)> guard(net, trmc_network_destroy);
    if ((rc = trmc_network_set_params_f64(net, params))) return rc;
    if ((rc = trmc_network_set_state_f64(net, qdp0_seg, dp0_seg, reach_last0))) return rc;
    if ((rc = trmc_network_set_forcing_f64(net, qlat_ts, n_timesteps))) return rc;
    if ((rc = trmc_network_run(net, n_timesteps, dt, nullptr))) return rc;
    return trmc_network_get_state_f64(net, qdc_seg_out, nullptr, nullptr, reach_last_out);
}

int trmc_wavefront_f32in_f64(const trmc_chparams_f32 *params, const int *reach_seg_start,
                             const int *reach_seg_len, const int *reach_up_ptr,
                             const int *reach_up_idx, const int *level_ptr, const int *level_reach,
                             int n_reaches, int n_levels, int n_segments, const float *qlat_ts,
                             const float *qdp0_seg, const float *dp0_seg, const float *reach_last0,
                             double *qdc_seg_out_d, double *reach_last_out_d, int n_timesteps,
                             double dt)
{
    if (n_timesteps < 0 || n_segments < 0 || n_reaches < 0) return fail(TRMC_ERR_ARG, "negative size");
    if ((n_segments > 0 && (!qdc_seg_out_d || !qdp0_seg || !dp0_seg)) ||
        (n_reaches > 0 && (!reach_last_out_d || !reach_last0)))
        return fail(TRMC_ERR_ARG, "null array");
    trmc_options o;
    trmc_default_options(&o);
    trmc_network *net = nullptr;
    int rc = trmc_network_create(reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx,
                                 level_ptr, level_reach, n_reaches, n_levels, n_segments, &o, &net);
    if (rc) return rc;
    std::unique_ptr<trmc_network, void (*)(trmc_network *)> guard(net, trmc_network_destroy);
    // float inputs widen exactly to double (wavefront_mc.cu:507-513)
    std::vector<double> q0(qdp0_seg, qdp0_seg + n_segments), d0(dp0_seg, dp0_seg + n_segments);
    std::vector<double> rl0(reach_last0, reach_last0 + n_reaches);
    if ((rc = trmc_network_set_params_f32(net, params))) return rc;
    if ((rc = trmc_network_set_state_f64(net, q0.data(), d0.data(), rl0.data()))) return rc;
    if ((rc = trmc_network_set_forcing_f32(net, qlat_ts, n_timesteps))) return rc;
    if ((rc = trmc_network_run(net, n_timesteps, dt, nullptr))) return rc;
    return trmc_network_get_state_f64(net, qdc_seg_out_d, nullptr, nullptr, reach_last_out_d);
}

int trmc_wavefront_f32(const trmc_chparams_f32 *params, const int *reach_seg_start,
                       const int *reach_seg_len, const int *reach_up_ptr, const int *reach_up_idx,
                       const int *level_ptr, const int *level_reach, int n_reaches, int n_levels,
                       int n_segments, const float *qlat_ts, const float *qdp0_seg, const float *dp0_seg,
                       const float *reach_last0, float *qdc_seg_out, float *reach_last_out, int n_timesteps,
                       float dt)
{
    if (n_timesteps < 0 || n_segments < 0 || n_reaches < 0) return fail(TRMC_ERR_ARG, "negative size");
    if ((n_segments > 0 && (!qdc_seg_out || !qdp0_seg || !dp0_seg)) ||
        (n_reaches > 0 && (!reach_last_out || !reach_last0)))
        return fail(TRMC_ERR_ARG, "null array");
    trmc_options o;
    trmc_default_options(&o);
    o.mode = TRMC_MODE_REFERENCE_F32;
    o.schedule = TRMC_SCHED_DATAFLOW;
    trmc_network *net = nullptr;
    int rc = trmc_network_create(reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx,
                                 level_ptr, level_reach, n_reaches, n_levels, n_segments, &o, &net);
    if (rc) return rc;
    std::unique_ptr<trmc_network, void (*)(trmc_network *)> guard(net, trmc_network_destroy);
    // float values are held exactly in the network's double state
    std::vector<double> q0(qdp0_seg, qdp0_seg + n_segments), d0(dp0_seg, dp0_seg + n_segments);
    std::vector<double> rl0(reach_last0, reach_last0 + n_reaches);
    if ((rc = trmc_network_set_params_f32(net, params))) return rc;
    if ((rc = trmc_network_set_state_f64(net, q0.data(), d0.data(), rl0.data()))) return rc;
    if ((rc = trmc_network_set_forcing_f32(net, qlat_ts, n_timesteps))) return rc;
    if ((rc = trmc_network_run(net, n_timesteps, (double)dt, nullptr))) return rc;
    std::vector<double> qo((size_t)n_segments), ro((size_t)n_reaches);
    if ((rc = trmc_network_get_state_f64(net, qo.data(), nullptr, nullptr, ro.data()))) return rc;
    for (int i = 0; i < n_segments; ++i) qdc_seg_out[i] = (float)qo[i];
    for (int r = 0; r < n_reaches; ++r) reach_last_out[r] = (float)ro[r];
    return TRMC_OK;
}

static int diffusive_run(const trmc_diff_reach *rp, const float *dx, const float *diffusivity,
                         const float *qp_prev, float *qp_out, int nreach, int reps, double *avg_ms)
{
    if (nreach < 0) return fail(TRMC_ERR_ARG, "negative nreach");
    if (nreach == 0) return TRMC_OK;
    if (!rp || !dx || !diffusivity || !qp_prev || !qp_out) return fail(TRMC_ERR_ARG, "null array");
    for (int j = 0; j < nreach; ++j)
        if (rp[j].ncomp < 1 || rp[j].ncomp > TRMC_DIFF_MAX_NODES)
            return fail(TRMC_ERR_ARG, "reach %d has ncomp %d (1..%d)", j, rp[j].ncomp, TRMC_DIFF_MAX_NODES);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TRMC_ERR_NO_DEVICE, "no gfx950 device");
    static_assert(sizeof(trmc_diff_reach) == sizeof(DiffReach), "reach record layout");
    const size_t nn = (size_t)nreach * TRMC_DIFF_MAX_NODES;
    DiffReach *d_rp = nullptr;
    float *d_dx = nullptr, *d_df = nullptr, *d_qp = nullptr, *d_out = nullptr, *d_ws = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = dev_upload(&d_rp, reinterpret_cast<const DiffReach *>(rp), (size_t)nreach, nullptr);
    if (!rc) rc = dev_upload(&d_dx, dx, nn, nullptr);
    if (!rc) rc = dev_upload(&d_df, diffusivity, nn, nullptr);
    if (!rc) rc = dev_upload(&d_qp, qp_prev, nn, nullptr);
    if (!rc) rc = dev_alloc(&d_out, nn);
    if (!rc) rc = dev_alloc(&d_ws, nn);
    hipError_t e = hipSuccess;
    float ms = 0.0f;
    if (!rc) {
        // entries beyond a reach's ncomp keep the caller's values, as in place
        e = hipMemcpy(d_out, qp_out, nn * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipEventCreate(&e0);
        if (e == hipSuccess) e = hipEventCreate(&e1);
        if (e == hipSuccess) e = launch_diffusive_tridiag(d_rp, d_dx, d_df, d_qp, d_out, d_ws, nreach, nullptr);
        if (e == hipSuccess && reps > 0) {
            e = hipEventRecord(e0, nullptr);
            for (int r = 0; r < reps && e == hipSuccess; ++r)
                e = launch_diffusive_tridiag(d_rp, d_dx, d_df, d_qp, d_out, d_ws, nreach, nullptr);
            if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        }
        if (e == hipSuccess) e = hipMemcpy(qp_out, d_out, nn * sizeof(float), hipMemcpyDeviceToHost);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    dev_free(d_rp); dev_free(d_dx); dev_free(d_df); dev_free(d_qp); dev_free(d_out); dev_free(d_ws);
    if (rc) return rc;
    if (e != hipSuccess) return fail(TRMC_ERR_HIP, "diffusive step: %s", hipGetErrorString(e));
    if (avg_ms) *avg_ms = reps > 0 ? (double)ms / reps : 0.0;
    return TRMC_OK;
}

int trmc_diffusive_tridiag_f32(const trmc_diff_reach *rp, const float *dx, const float *diffusivity,
                               const float *celerity, const float *qp_prev, float *qp_out, int nreach)
{
    (void)celerity;   // computed into an unused Courant number by the reference
    return diffusive_run(rp, dx, diffusivity, qp_prev, qp_out, nreach, 0, nullptr);
}

int trmc_diffusive_tridiag_timed_f32(const trmc_diff_reach *rp, const float *dx, const float *diffusivity,
                                     const float *celerity, const float *qp_prev, float *qp_out, int nreach,
                                     int reps, double *avg_kernel_ms)
{
    (void)celerity;
    if (reps < 1 || !avg_kernel_ms) return fail(TRMC_ERR_ARG, "reps >= 1 and avg_kernel_ms required");
    return diffusive_run(rp, dx, diffusivity, qp_prev, qp_out, nreach, reps, avg_kernel_ms);
}

static int coeffs_run(const trmc_chparams_f32 *params, int n, float dt, int fixed_x, float *C1, float *C2,
                      float *C3, float *C4oq)
{
    if (n < 0) return fail(TRMC_ERR_ARG, "negative n");
    if (n == 0) return TRMC_OK;
    if (!params || !C1 || !C2 || !C3 || !C4oq) return fail(TRMC_ERR_ARG, "null array");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(TRMC_ERR_NO_DEVICE, "no gfx950 device");
    static_assert(sizeof(trmc_chparams_f32) == 8 * sizeof(float), "ChParams layout");
    float *d_p = nullptr, *d_out = nullptr;
    int rc = dev_upload(&d_p, reinterpret_cast<const float *>(params), (size_t)n * 8, nullptr);
    if (!rc) rc = dev_alloc(&d_out, (size_t)4 * n);
    hipError_t e = hipSuccess;
    if (!rc) e = launch_linear_coeffs(d_p, n, dt, fixed_x, d_out, nullptr);
    std::vector<float> out;
    if (!rc && e == hipSuccess) {
        out.resize((size_t)4 * n);
        e = hipMemcpy(out.data(), d_out, out.size() * sizeof(float), hipMemcpyDeviceToHost);
    }
    dev_free(d_p);
    dev_free(d_out);
    if (rc) return rc;
    if (e != hipSuccess) return fail(TRMC_ERR_HIP, "linear coefficients: %s", hipGetErrorString(e));
    float *dst[4] = {C1, C2, C3, C4oq};
    for (int k = 0; k < 4; ++k) std::copy(out.begin() + (size_t)k * n, out.begin() + (size_t)(k + 1) * n, dst[k]);
    return TRMC_OK;
}

int trmc_linear_coeffs_f32(const trmc_chparams_f32 *params, int n, float dt, float *C1, float *C2, float *C3,
                           float *C4oq)
{
    return coeffs_run(params, n, dt, 0, C1, C2, C3, C4oq);
}

int trmc_matrix_coeffs_f32(const trmc_chparams_f32 *params, int n, float dt, float *C1, float *C2, float *C3,
                           float *C4oq)
{
    return coeffs_run(params, n, dt, 1, C1, C2, C3, C4oq);
}

// Host checks of one propagation matrix: the kernel gathers x[col] for every
// entry, so every column must index the vector.
static int check_csr(const trmc_csr *m, int n, const char *name)
{
    if (!m) return fail(TRMC_ERR_ARG, "%s is null", name);
    if (m->n != n) return fail(TRMC_ERR_ARG, "%s has %d rows, expected %d", name, m->n, n);
    if (m->nnz < 0 || !m->rowptr || (m->nnz > 0 && (!m->col || !m->val)))
        return fail(TRMC_ERR_ARG, "%s: null array or negative nnz", name);
    if (m->rowptr[0] != 0 || m->rowptr[n] != m->nnz)
        return fail(TRMC_ERR_ARG, "%s: rowptr must run from 0 to nnz", name);
    for (int r = 0; r < n; ++r)
        if (m->rowptr[r + 1] < m->rowptr[r]) return fail(TRMC_ERR_ARG, "%s: rowptr decreases at row %d", name, r);
    for (int r = 0; r < n; ++r)
        for (int k = m->rowptr[r]; k < m->rowptr[r + 1]; ++k) {
            if (m->col[k] < 0 || m->col[k] >= n)
                return fail(TRMC_ERR_ARG, "%s: column %d out of range", name, m->col[k]);
            if (k > m->rowptr[r] && m->col[k] <= m->col[k - 1])
                return fail(TRMC_ERR_ARG, "%s: columns of row %d not ascending", name, r);
        }
    return TRMC_OK;
}

static int matrix_run(const trmc_csr *PQ, const trmc_csr *Pq, const float *qlat, const float *Q0, int T,
                      int lanes, float *Q_final, int reps, double *avg_ms)
{
    if (!PQ) return fail(TRMC_ERR_ARG, "P_Q is null");
    const int n = PQ->n;
    if (n < 0 || T < 0) return fail(TRMC_ERR_ARG, "negative size");
    int rc = check_csr(PQ, n, "P_Q");
    if (!rc) rc = check_csr(Pq, n, "P_q");
    if (rc) return rc;
    if (lanes < 0 || lanes > 64 || (lanes & (lanes - 1)))
        return fail(TRMC_ERR_ARG, "lanes_per_row %d is neither 0 nor a power of two in 1..64", lanes);
    if (n == 0) return TRMC_OK;
    if (!Q0 || !Q_final || (T > 0 && !qlat)) return fail(TRMC_ERR_ARG, "null array");
    if (T == 0) {
        std::copy(Q0, Q0 + n, Q_final);
        if (avg_ms) *avg_ms = 0.0;
        return TRMC_OK;
    }
    if ((rc = check_device(0, nullptr, nullptr))) return rc;
    // the union of each row's two CSR rows: column, {P_Q value, P_q value}
    std::vector<int> urow((size_t)n + 1), ucol;
    std::vector<float2> uval;
    ucol.reserve((size_t)std::max(PQ->nnz, Pq->nnz));
    uval.reserve((size_t)std::max(PQ->nnz, Pq->nnz));
    urow[0] = 0;
    for (int r = 0; r < n; ++r) {
        int a = PQ->rowptr[r], b = Pq->rowptr[r];
        const int ae = PQ->rowptr[r + 1], be = Pq->rowptr[r + 1];
        while (a < ae || b < be) {
            const int ca = a < ae ? PQ->col[a] : INT_MAX, cb = b < be ? Pq->col[b] : INT_MAX;
            const int c = std::min(ca, cb);
            ucol.push_back(c);
            uval.push_back(make_float2(ca == c ? PQ->val[a++] : 0.0f, cb == c ? Pq->val[b++] : 0.0f));
        }
        if (ucol.size() > (size_t)INT_MAX) return fail(TRMC_ERR_ARG, "union of the matrices exceeds 2^31 entries");
        urow[r + 1] = (int)ucol.size();
    }
    if (ucol.empty()) { ucol.push_back(0); uval.push_back(make_float2(0.0f, 0.0f)); }
    // lanes per row: fixed, or by the row's union length U (smallest power of
    // two G with 2 G >= U, at most 64); slots hold the rows of larger G first,
    // in row order within one G; blocks hold one G each
    std::vector<int> gof((size_t)n), order((size_t)n);
    for (int r = 0; r < n; ++r) {
        int G = lanes;
        if (G == 0) {
            const int U = urow[r + 1] - urow[r];
            for (G = 1; G < 64 && G * 2 < U; G *= 2) {}
        }
        gof[r] = G;
        order[r] = r;
    }
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return gof[a] > gof[b]; });
    std::vector<int4> slots((size_t)n), blocks;
    for (int i = 0; i < n; ++i) slots[i] = make_int4(order[i], urow[order[i]], urow[order[i] + 1], 0);
    for (int i = 0; i < n;) {
        const int G = gof[order[i]], per = 256 / G;
        int k = 0;
        while (k < per && i + k < n && gof[order[i + k]] == G) ++k;
        blocks.push_back(make_int4(G, i, k, 0));
        i += k;
    }
    // state pairs {Q, ql of the step about to run}
    std::vector<float2> z0((size_t)n);
    for (int r = 0; r < n; ++r) z0[r] = make_float2(Q0[r], qlat[r]);
    int4 *d_slots = nullptr, *d_blk = nullptr;
    int *d_col = nullptr;
    float2 *d_val = nullptr, *d_za = nullptr, *d_zb = nullptr;
    float *d_ql = nullptr;
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // TRMC_MATRIX_MAX_GRID (tests): a smaller grid, so that the step
    // kernel's loop over its block table takes more than one trip
    const char *cap = std::getenv("TRMC_MATRIX_MAX_GRID");
    const int max_grid = cap ? std::atoi(cap) : 0;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) return fail(TRMC_ERR_HIP, "stream: %s", hipGetErrorString(e));
    rc = dev_upload(&d_slots, slots.data(), slots.size(), st);
    if (!rc) rc = dev_upload(&d_blk, blocks.data(), blocks.size(), st);
    if (!rc) rc = dev_upload(&d_col, ucol.data(), ucol.size(), st);
    if (!rc) rc = dev_upload(&d_val, uval.data(), uval.size(), st);
    if (!rc) rc = dev_upload(&d_ql, qlat, (size_t)T * n, st);
    if (!rc) rc = dev_upload(&d_za, z0.data(), (size_t)n, st);
    if (!rc) rc = dev_alloc(&d_zb, (size_t)n);
    std::vector<float2> zf;
    float ms = 0.0f;
    if (!rc) {
        e = hipStreamSynchronize(st);
        // the T steps as one graph: step t reads one state buffer and writes
        // the other, with the lateral inflow of step t + 1 (of step 0 after
        // the last, so a replay starts from well-formed pairs)
        if (e == hipSuccess) e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            hipError_t le = hipSuccess;
            for (int t = 0; t < T && le == hipSuccess; ++t) {
                const float2 *zi = (t & 1) ? d_zb : d_za;
                float2 *zo = (t & 1) ? d_za : d_zb;
                le = launch_matrix_step((int)blocks.size(), d_blk, d_slots, d_col, d_val, zi,
                                        d_ql + (size_t)((t + 1) % T) * n, zo, st, max_grid);
            }
            e = hipStreamEndCapture(st, &graph);
            if (e == hipSuccess) e = le;
        }
        if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (e == hipSuccess) e = hipGraphLaunch(exec, st);
        if (e == hipSuccess) {
            zf.resize((size_t)n);
            e = hipMemcpyAsync(zf.data(), (T & 1) ? d_zb : d_za, (size_t)n * sizeof(float2), hipMemcpyDeviceToHost,
                               st);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        // timed replays continue from the state the previous one left: the
        // work of a step does not depend on the values
        if (e == hipSuccess && reps > 0) {
            e = hipEventCreate(&e0);
            if (e == hipSuccess) e = hipEventCreate(&e1);
            if (e == hipSuccess) e = hipEventRecord(e0, st);
            for (int r = 0; r < reps && e == hipSuccess; ++r) e = hipGraphLaunch(exec, st);
            if (e == hipSuccess) e = hipEventRecord(e1, st);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        }
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(st);
    dev_free(d_slots); dev_free(d_blk); dev_free(d_col); dev_free(d_val); dev_free(d_ql); dev_free(d_za);
    dev_free(d_zb);
    (void)hipStreamDestroy(st);
    if (rc) return rc;
    if (e != hipSuccess) return fail(TRMC_ERR_HIP, "matrix route: %s", hipGetErrorString(e));
    for (int r = 0; r < n; ++r) Q_final[r] = zf[r].x;
    if (avg_ms) *avg_ms = reps > 0 ? (double)ms / reps : 0.0;
    return TRMC_OK;
}

int trmc_matrix_route_f32(const trmc_csr *P_Q, const trmc_csr *P_q, const float *qlat, const float *Q0,
                          int n_timesteps, int lanes_per_row, float *Q_final)
{
    return matrix_run(P_Q, P_q, qlat, Q0, n_timesteps, lanes_per_row, Q_final, 0, nullptr);
}

int trmc_matrix_route_timed_f32(const trmc_csr *P_Q, const trmc_csr *P_q, const float *qlat, const float *Q0,
                                int n_timesteps, int lanes_per_row, float *Q_final, int reps, double *avg_pass_ms)
{
    if (reps < 1 || !avg_pass_ms) return fail(TRMC_ERR_ARG, "reps >= 1 and avg_pass_ms required");
    return matrix_run(P_Q, P_q, qlat, Q0, n_timesteps, lanes_per_row, Q_final, reps, avg_pass_ms);
}

int trmc_linear_mc_f64(const float *C1, const float *C2, const float *C3, const float *C4oq,
                       const int *reach_seg_start, const int *reach_seg_len, const int *reach_up_ptr,
                       const int *reach_up_idx, const int *level_ptr, const int *level_reach, int n_reaches,
                       int n_levels, int n_segments, const float *qlat_ts, const float *Q0, double *Q_final,
                       int n_timesteps)
{
    if (n_timesteps < 0 || n_segments < 0 || n_reaches < 0) return fail(TRMC_ERR_ARG, "negative size");
    if (n_segments > 0 && (!Q0 || !Q_final || !C1 || !C2 || !C3 || !C4oq)) return fail(TRMC_ERR_ARG, "null array");
    for (int r = 0; r < n_reaches; ++r)
        if (reach_seg_len && reach_seg_len[r] < 1)
            return fail(TRMC_ERR_TOPOLOGY, "reach %d has no segment (the linear routine writes its first one)", r);
    trmc_options o;
    trmc_default_options(&o);
    o.mode = TRMC_MODE_LINEAR;
    trmc_network *net = nullptr;
    int rc = trmc_network_create(reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx, level_ptr, level_reach,
                                 n_reaches, n_levels, n_segments, &o, &net);
    if (rc) return rc;
    std::unique_ptr<trmc_network, void (*)(trmc_network *)> guard(net, trmc_network_destroy);
    // every segment starts at Q0; an upstream reach's previous-step outflow is
    // the flow of its last segment (linear_mc.cu:255-256, :266)
    std::vector<double> q0(Q0, Q0 + n_segments), d0((size_t)n_segments, 0.0), rl0((size_t)n_reaches, 0.0);
    for (int r = 0; r < n_reaches; ++r) rl0[r] = (double)Q0[reach_seg_start[r] + reach_seg_len[r] - 1];
    if ((rc = trmc_network_set_linear_coeffs(net, C1, C2, C3, C4oq))) return rc;
    if ((rc = trmc_network_set_state_f64(net, q0.data(), d0.data(), rl0.data()))) return rc;
    if ((rc = trmc_network_set_forcing_f32(net, qlat_ts, n_timesteps))) return rc;
    if ((rc = trmc_network_run(net, n_timesteps, 300.0, nullptr))) return rc;
    return trmc_network_get_state_f64(net, Q_final, nullptr, nullptr, nullptr);
}

int trmc_reach_parallel_f64(const trmc_chparams_f64 *p, const double *ql, const int *rs,
                            const int *rl, const double *qu, const double *qd, const double *dp,
                            double *qo, double *vo, double *dpo, double uq, double dt, int nr)
{
    if (nr < 0) return fail(TRMC_ERR_ARG, "negative nr");
    if (nr == 0) return TRMC_OK;
    if (!p || !ql || !rs || !rl || !qu || !qd || !dp || !qo || !vo || !dpo)
        return fail(TRMC_ERR_ARG, "null array");
    int n_seg = 0;
    for (int r = 0; r < nr; ++r) {
        if (rs[r] < 0 || rl[r] < 0) return fail(TRMC_ERR_ARG, "reach %d has a negative range", r);
        n_seg = std::max(n_seg, rs[r] + rl[r]);
    }
    std::vector<unsigned char> seen((size_t)n_seg, 0);
    for (int r = 0; r < nr; ++r)
        for (int i = rs[r]; i < rs[r] + rl[r]; ++i) {
            if (seen[i]) return fail(TRMC_ERR_TOPOLOGY, "segment %d belongs to two reaches", i);
            seen[i] = 1;
        }
    int cu = 0;
    double khz = 0;
    int rc = check_device(0, &cu, &khz);
    if (rc) return rc;
    std::vector<int> ident((size_t)n_seg);
    for (int i = 0; i < n_seg; ++i) ident[i] = i;
    std::vector<double> segc;
    build_segc(p, ident, segc);
    hipStream_t s = nullptr;
    double *d_segc = nullptr, *d_ql = nullptr, *d_qu = nullptr, *d_qd = nullptr, *d_dp = nullptr,
           *d_qo = nullptr, *d_vo = nullptr, *d_dpo = nullptr;
    int *d_rs = nullptr, *d_rl = nullptr;
    auto cleanup = [&]() {
        dev_free(d_segc); dev_free(d_ql); dev_free(d_qu); dev_free(d_qd); dev_free(d_dp);
        dev_free(d_qo); dev_free(d_vo); dev_free(d_dpo); dev_free(d_rs); dev_free(d_rl);
    };
    const size_t n = (size_t)n_seg;
    rc = dev_upload(&d_segc, segc.data(), segc.size(), s);
    if (!rc) rc = dev_upload(&d_ql, ql, n, s);
    if (!rc) rc = dev_upload(&d_qu, qu, n, s);
    if (!rc) rc = dev_upload(&d_qd, qd, n, s);
    if (!rc) rc = dev_upload(&d_dp, dp, n, s);
    if (!rc) rc = dev_upload(&d_qo, qo, n, s);     // untouched entries keep the caller's values
    if (!rc) rc = dev_upload(&d_vo, vo, n, s);
    if (!rc) rc = dev_upload(&d_dpo, dpo, n, s);
    if (!rc) rc = dev_upload(&d_rs, rs, (size_t)nr, s);
    if (!rc) rc = dev_upload(&d_rl, rl, (size_t)nr, s);
    hipError_t e = hipSuccess;
    if (!rc) e = launch_reach_parallel(d_segc, n_seg, d_ql, d_rs, d_rl, d_qu, d_qd, d_dp, d_qo, d_vo,
                                       d_dpo, uq, dt, nr, s);
    if (!rc && e == hipSuccess) e = hipMemcpy(qo, d_qo, n * sizeof(double), hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess) e = hipMemcpy(vo, d_vo, n * sizeof(double), hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess) e = hipMemcpy(dpo, d_dpo, n * sizeof(double), hipMemcpyDeviceToHost);
    cleanup();
    if (rc) return rc;
    if (e != hipSuccess) return fail(TRMC_ERR_HIP, "reach-parallel: %s", hipGetErrorString(e));
    return TRMC_OK;
}

int trmc_mc_solve_f64(const trmc_chparams_f64 *params, const double *in, double *out, int *iters,
                      int n, double dt, int mode)
{
    if (n < 0) return fail(TRMC_ERR_ARG, "negative n");
    if (n == 0) return TRMC_OK;
    if (!params || !in || !out) return fail(TRMC_ERR_ARG, "null array");
    if (mode != TRMC_MODE_REFERENCE && mode != TRMC_MODE_CONVERGED)
        return fail(TRMC_ERR_ARG, "mode must be REFERENCE or CONVERGED");
    int cu = 0;
    double khz = 0;
    int rc = check_device(0, &cu, &khz);
    if (rc) return rc;
    std::vector<int> ident((size_t)n);
    for (int i = 0; i < n; ++i) ident[i] = i;
    std::vector<double> segc;
    build_segc(params, ident, segc);
    double *d_segc = nullptr, *d_in = nullptr, *d_out = nullptr;
    int *d_it = nullptr;
    rc = dev_upload(&d_segc, segc.data(), segc.size(), nullptr);
    if (!rc) rc = dev_upload(&d_in, in, (size_t)n * 5, nullptr);
    if (!rc) rc = dev_alloc(&d_out, (size_t)n * 3);
    if (!rc) rc = dev_alloc(&d_it, (size_t)n);
    hipError_t e = hipSuccess;
    if (!rc) e = launch_solve_batch(d_segc, n, d_in, d_out, d_it, dt, mode == TRMC_MODE_REFERENCE ? 0 : 1,
                                    1e-13, 60, nullptr);
    if (!rc && e == hipSuccess) e = hipMemcpy(out, d_out, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost);
    if (!rc && e == hipSuccess && iters) e = hipMemcpy(iters, d_it, (size_t)n * sizeof(int), hipMemcpyDeviceToHost);
    dev_free(d_segc); dev_free(d_in); dev_free(d_out); dev_free(d_it);
    if (rc) return rc;
    if (e != hipSuccess) return fail(TRMC_ERR_HIP, "solve batch: %s", hipGetErrorString(e));
    return TRMC_OK;
}

// -------------------------------------------------------- copy reference
int trmc_stream_copy(int device, size_t bytes, int reps, double *gbps)
{
    if (bytes < 16 || reps <= 0) return fail(TRMC_ERR_ARG, "need bytes >= 16 and reps > 0");
    int cu = 0;
    double khz = 0.0;
    int rc = check_device(device, &cu, &khz);
    if (rc) return rc;
    HIPCHK(hipSetDevice(device));
    bytes &= ~(size_t)15;
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return fail(TRMC_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(TRMC_ERR_NOMEM, "hipMalloc of %zu bytes failed", bytes);
    }
    hipEvent_t e0, e1;
    hipError_t e = hipMemset(a, 0, bytes);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = launch_stream_copy(a, b, bytes, nullptr);      // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_stream_copy(a, b, bytes, nullptr);
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(TRMC_ERR_HIP, "stream copy: %s", hipGetErrorString(e));
    if (gbps) *gbps = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
    return TRMC_OK;
}

// ------------------------------------------------------------- autotuner
// Candidate schedules are timed on the first n_probe steps (after one warm-up
// run of the same steps): unit size (one network each), then batch size and
// workgroup size on that network.  The fastest kernel time wins.
int trmc_autotune(const int *reach_seg_start, const int *reach_seg_len, const int *reach_up_ptr,
                  const int *reach_up_idx, const int *level_ptr, const int *level_reach, int n_reaches,
                  int n_levels, int n_segments, const trmc_chparams_f64 *params, const double *qdp0_seg,
                  const double *dp0_seg, const double *reach_last0, const double *qlat, uint64_t forcing_seed,
                  int n_probe, double dt, const float *mlp_weights, const trmc_options *base, trmc_options *best,
                  double *best_ms_per_step)
{
    if (!params || !qdp0_seg || !dp0_seg || !best || n_probe <= 0) return fail(TRMC_ERR_ARG, "bad argument");
    trmc_options o0;
    if (base) o0 = *base; else trmc_default_options(&o0);
    // The schedules searched (when base leaves the schedule to the library):
    //   tile      DATAFLOW, interleaved: each lane owns a tile of units (one
    //             member each) and the wave interleaves their solves; solve
    //             batch and workgroup size searched;
    //   whole     DATAFLOW, one whole item per lane pass (variant 1);
    //   wavefront SKEWED: the level wavefront of the reference, one lane per
    //             (unit, member) per super-level between device-scope barriers.
    // LINEAR and REFERENCE_F32 run on the tile schedule only.
    struct Cand { int sched, variant, batch, block; };
    static const int kUnits[] = {2, 4, 8};
    std::vector<Cand> cands;
    const bool any_sched = o0.schedule == TRMC_SCHED_AUTO;
    const bool tile_only = o0.mode == TRMC_MODE_LINEAR || o0.mode == TRMC_MODE_REFERENCE_F32;
    const bool want_df = any_sched || o0.schedule == TRMC_SCHED_DATAFLOW;
    const bool want_sk = !tile_only && (any_sched || o0.schedule == TRMC_SCHED_SKEWED);
    if (want_df)
        for (int bi : {6, 8, 12})
            for (int blk : {64, 128}) cands.push_back({TRMC_SCHED_DATAFLOW, 0, bi, blk});
    if (want_df && !tile_only)
        for (int blk : {64, 128}) cands.push_back({TRMC_SCHED_DATAFLOW, 1, 0, blk});
    if (want_sk)
        for (int blk : {64, 256}) cands.push_back({TRMC_SCHED_SKEWED, 0, 0, blk});
    if (cands.empty()) cands.push_back({o0.schedule, o0.dataflow_variant, 0, o0.block_size});
    double best_ms = 1e300;
    trmc_options winner = o0;
    for (int us : kUnits) {
        trmc_options o = o0;
        o.unit_segments = us;
        o.record_history = 0;
        trmc_network *net = nullptr;
        int rc = trmc_network_create(reach_seg_start, reach_seg_len, reach_up_ptr, reach_up_idx, level_ptr,
                                     level_reach, n_reaches, n_levels, n_segments, &o, &net);
        if (!rc) rc = trmc_network_set_params_f64(net, params);
        if (!rc) rc = trmc_network_set_state_f64(net, qdp0_seg, dp0_seg, reach_last0);
        if (!rc) rc = qlat ? trmc_network_set_forcing_f64(net, qlat, n_probe)
                           : trmc_network_generate_forcing(net, n_probe, forcing_seed);
        if (!rc && o.mode == TRMC_MODE_LEARNED && mlp_weights) rc = trmc_network_set_mlp(net, mlp_weights);
        for (const Cand &cd : cands) {
            if (rc) break;
            if (cd.sched == TRMC_SCHED_DATAFLOW && net->plan.max_upstream > kDataflowMaxUp) continue;
            net->opt.schedule = cd.sched;
            net->opt.dataflow_variant = cd.variant;
            if (cd.batch > 0) net->opt.batch_iters = cd.batch;
            net->opt.block_size = cd.block;
            trmc_run_stats st;
            rc = trmc_network_run(net, n_probe, dt, &st);
            if (!rc) rc = trmc_network_reset(net);
            if (!rc) rc = trmc_network_run(net, n_probe, dt, &st);
            if (!rc) rc = trmc_network_reset(net);
            if (!rc && st.kernel_ms < best_ms) {
                best_ms = st.kernel_ms;
                winner = net->opt;
            }
        }
        if (net) trmc_network_destroy(net);
        if (rc) return rc;
    }
    *best = winner;
    if (best_ms_per_step) *best_ms_per_step = best_ms / n_probe;
    return TRMC_OK;
}

// ------------------------------------------------------ formats, generator
int trmc_load_topo(const char *path, trmc_topo *out)
{
    if (!path || !out) return fail(TRMC_ERR_ARG, "null argument");
    std::string err;
    int rc = load_topo(path, out, err);
    if (rc) return fail(rc, "%s", err.c_str());
    return TRMC_OK;
}

int trmc_save_topo(const char *path, const trmc_topo *t)
{
    if (!path || !t) return fail(TRMC_ERR_ARG, "null argument");
    int rc = save_topo(path, t);
    if (rc) return fail(rc, "cannot write %s", path);
    return TRMC_OK;
}

int trmc_topo_concat(const trmc_topo *nets, int n, trmc_topo *out, int *seg_offset, int *reach_offset)
{
    if (!out || n < 0 || (n > 0 && !nets)) return fail(TRMC_ERR_ARG, "bad argument");
    int rc = concat_topo(nets, n, out, seg_offset, reach_offset);
    if (rc) {
        trmc_free_topo(out);
        return fail(rc, rc == TRMC_ERR_NOMEM ? "out of host memory" : "inconsistent or oversized topology");
    }
    return TRMC_OK;
}

void trmc_free_topo(trmc_topo *t)
{
    if (!t) return;
    int **arrs[] = {&t->reach_seg_start, &t->reach_seg_len, &t->reach_level, &t->reach_n_up,
                    &t->reach_up_ptr, &t->reach_up_idx, &t->level_ptr, &t->level_reach};
    for (int **a : arrs) { std::free(*a); *a = nullptr; }
}

void trmc_free(void *p) { std::free(p); }

int trmc_load_params(const char *path, int n_segments, trmc_chparams_f32 *out)
{
    if (!path || !out || n_segments < 0) return fail(TRMC_ERR_ARG, "bad argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(TRMC_ERR_IO, "cannot open %s", path);
    size_t got = std::fread(out, sizeof(trmc_chparams_f32), (size_t)n_segments, f);
    std::fclose(f);
    if (got != (size_t)n_segments) return fail(TRMC_ERR_IO, "short read %s: %zu of %d", path, got, n_segments);
    return TRMC_OK;
}

int trmc_save_params(const char *path, int n_segments, const trmc_chparams_f32 *p)
{
    if (!path || !p || n_segments < 0) return fail(TRMC_ERR_ARG, "bad argument");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(TRMC_ERR_IO, "cannot open %s", path);
    size_t put = std::fwrite(p, sizeof(trmc_chparams_f32), (size_t)n_segments, f);
    int cl = std::fclose(f);
    if (put != (size_t)n_segments || cl) return fail(TRMC_ERR_IO, "short write %s", path);
    return TRMC_OK;
}

int trmc_load_forcings(const char *path, int n_reaches, int n_segments, int *n_timesteps, float *qlat,
                       float *qup0, float *qdp0, float *dp0)
{
    if (!path || !n_timesteps || n_reaches < 0 || n_segments < 0) return fail(TRMC_ERR_ARG, "bad argument");
    FILE *f = std::fopen(path, "rb");
    if (!f) return fail(TRMC_ERR_IO, "cannot open %s", path);
    int T = 0;
    if (std::fread(&T, 4, 1, f) != 1 || T < 0) { std::fclose(f); return fail(TRMC_ERR_IO, "bad header %s", path); }
    *n_timesteps = T;
    if (!qlat && !qup0 && !qdp0 && !dp0) { std::fclose(f); return TRMC_OK; }
    struct Part { float *dst; size_t n; } parts[] = {
        {qlat, (size_t)T * n_segments}, {qup0, (size_t)n_reaches}, {qdp0, (size_t)n_segments},
        {dp0, (size_t)n_segments}};
    for (auto &pt : parts) {
        if (pt.dst) {
            if (std::fread(pt.dst, 4, pt.n, f) != pt.n) { std::fclose(f); return fail(TRMC_ERR_IO, "short read %s", path); }
        } else if (std::fseek(f, (long)(pt.n * 4), SEEK_CUR) != 0) {
            std::fclose(f);
            return fail(TRMC_ERR_IO, "seek failed in %s", path);
        }
    }
    std::fclose(f);
    return TRMC_OK;
}

int trmc_save_forcings(const char *path, int n_reaches, int n_segments, int n_timesteps,
                       const float *qlat, const float *qup0, const float *qdp0, const float *dp0)
{
    if (!path || n_reaches < 0 || n_segments < 0 || n_timesteps < 0) return fail(TRMC_ERR_ARG, "bad argument");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(TRMC_ERR_IO, "cannot open %s", path);
    bool ok = std::fwrite(&n_timesteps, 4, 1, f) == 1;
    std::vector<float> zeros;
    auto put = [&](const float *p, size_t n) {
        if (!ok || n == 0) return;
        if (!p) { zeros.assign(n, 0.0f); p = zeros.data(); }
        ok = std::fwrite(p, 4, n, f) == n;
    };
    put(qlat, (size_t)n_timesteps * n_segments);
    put(qup0, (size_t)n_reaches);
    put(qdp0, (size_t)n_segments);
    put(dp0, (size_t)n_segments);
    ok = (std::fclose(f) == 0) && ok;
    return ok ? TRMC_OK : fail(TRMC_ERR_IO, "short write %s", path);
}

int trmc_generate_network(int kind, int n_segments, int n_reaches, uint64_t seed, trmc_topo *out)
{
    int rc = generate_network(kind, n_segments, n_reaches, seed, out);
    if (rc) return fail(rc, "generate_network(kind=%d, n=%d) failed", kind, n_segments);
    return TRMC_OK;
}

int trmc_generate_params(int n_segments, uint64_t seed, trmc_chparams_f64 *out)
{
    int rc = generate_params(n_segments, seed, out);
    return rc ? fail(rc, "generate_params failed") : TRMC_OK;
}

int trmc_generate_forcing(int n_segments, int n_timesteps, int n_members, uint64_t seed, double *qlat)
{
    int rc = generate_forcing(n_segments, n_timesteps, n_members, seed, qlat);
    return rc ? fail(rc, "generate_forcing failed") : TRMC_OK;
}

int trmc_generate_forcing_member(int n_segments, int n_timesteps, int member, uint64_t seed, double *qlat)
{
    int rc = generate_forcing_members(n_segments, n_timesteps, member, 1, seed, qlat);
    return rc ? fail(rc, "generate_forcing_member failed") : TRMC_OK;
}

int trmc_generate_state(int n_segments, const trmc_chparams_f64 *params, double q0, double *qdp0,
                        double *dp0)
{
    int rc = generate_state(n_segments, params, q0, qdp0, dp0);
    return rc ? fail(rc, "generate_state failed") : TRMC_OK;
}

int trmc_compute_stats(const double *pred, const double *truth, int n, double *out)
{
    int rc = compute_stats(pred, truth, n, out);
    return rc ? fail(rc, "compute_stats: bad argument") : TRMC_OK;
}

}  // extern "C"
