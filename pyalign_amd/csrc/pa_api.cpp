// pa_api.cpp -- the C ABI of libpyalign_hip.so (include/pyalign_hip.h).
//
// Host side of the solver: options and solver objects mirroring the
// reference's pybind11 surface (pyalign/algorithm/pyalign.h:12-102,
// factory.h:414-508, 1243-1278), the border vectors of the reference's
// solver constructors, the scheduler that turns a ragged bag of problems
// into wavefront groups (the autotuned work queue, or the reference's
// Morton batches for the plain form), the launches, and the result objects
// (scores, one-optimal paths, all-optimal enumeration over the GPU's
// traceback bits).
#include "pyalign_hip.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <numeric>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "pa_internal.h"

using namespace pa;

namespace {

thread_local std::string g_err;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] void fail(int code, const std::string &msg) { throw Error(code, msg); }

#define HIPCHK(x)                                                                                 \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) fail(PA_EHIP, std::string(#x) + ": " + hipGetErrorString(e_));      \
    } while (0)

template <class F>
int guard(F &&f) {
    try {
        f();
        return PA_OK;
    } catch (const Error &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_err = "out of host memory";
        return PA_ERUNTIME;
    } catch (const std::exception &e) {
        g_err = e.what();
        return PA_ERUNTIME;
    }
}

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    void ensure(size_t bytes) {
        if (bytes <= cap) return;
        if (p) HIPCHK(hipFree(p));
        p = nullptr;
        cap = 0;
        const size_t c = std::max<size_t>(bytes + bytes / 4, 256);
        HIPCHK(hipMalloc(&p, c));
        cap = c;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

inline float worst(int dir) {
    return dir == PA_MAXIMIZE ? -std::numeric_limits<float>::infinity() : std::numeric_limits<float>::infinity();
}

// AffineCost::vector (solver.h:2876-2880): linspace<float>(0, (n-1)*u, n) + v,
// w[0] = 0, with xtensor's double-precision generator (start + k*step,
// stop at k = n-1).  Equals fl32(fl32(k*u) + v) whenever (n-1)*u is exact.
std::vector<float> affine_vector(float u, float v, int64_t n) {
    std::vector<float> w(static_cast<size_t>(n));
    const float stop = float(n - 1) * u;
    const double step = double(stop - 0.0f) / std::max(1.0, double(n - 1));
    for (int64_t k = 0; k < n; k++) {
        const float lin = (n > 1 && k == n - 1) ? stop : float(0.0 + double(k) * step);
        w[size_t(k)] = lin + v;
    }
    w[0] = 0.0f;
    return w;
}

uint64_t part1by1(uint64_t x) {
    x &= 0xffffffffull;
    x = (x | (x << 16)) & 0x0000ffff0000ffffull;
    x = (x | (x << 8)) & 0x00ff00ff00ff00ffull;
    x = (x | (x << 4)) & 0x0f0f0f0f0f0f0f0full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

// Morton code of a problem shape (pymorton.interleave2: x in the even bits).
uint64_t morton(uint64_t x, uint64_t y) { return part1by1(x) | (part1by1(y) << 1); }

}  // namespace

// ---------------------------------------------------------------------------
// options / solver objects
// ---------------------------------------------------------------------------

struct pa_options {
    pa_options_desc d{};
    int type = 0, dir = 0, loc = 0, goal = 0, gap = 0;
    float lin_s = 0, lin_t = 0;
    float aff_u_s = 0, aff_v_s = 0, aff_u_t = 0, aff_v_t = 0;  // AffineCost(u=extend, v=open-extend)
    bool remove_dup = true;
    int index_bits = 16;
};

namespace {

struct Plan {
    std::vector<Prob> probs;    // in group order
    std::vector<Group> groups;  // per launch, concatenated
    std::vector<std::pair<int, std::pair<int, int>>> launches;  // R, [first group, end group)
    int64_t rowbuf = 0;         // float2 elements
    int64_t tb_words = 0;
    int64_t dvals = 0;
    double cells = 0.0;
    int64_t band_slots = 0;            // strips of band-mode problems
    std::vector<int32_t> band_probs;   // their indices in probs (long-mode problems first)
    int32_t n_long = 0;                // long-mode problems (pa_long_kernel): band_probs[0 .. n_long)
    int32_t long_bands = 0;            // their bands (workgroup tickets)
    int64_t xbuf_begin = 0, xbuf_end = 0;  // float2 range of the row buffer holding their tagged hand-offs
};

// The last call's schedule, reused (with its device copy) while the next
// call has the same lengths, offsets and job: repeated batch shapes skip
// planning and the descriptor upload.
struct PlanCache {
    bool valid = false;
    bool uploaded = false;  // b_probs / b_groups / b_band_probs hold this plan
    std::vector<int64_t> key;
    std::vector<int32_t> len_s, len_t;
    std::vector<int64_t> off_s, off_t;
    Plan plan;
};

}  // namespace

struct pa_solver {
    pa_options opt;
    PlanCache last_plan;
    std::vector<std::pair<std::vector<int>, int>> occ_cache;
    int64_t lim_s = 0, lim_t = 0;
    std::vector<float> border_s, border_t;
    float go_s = 0, ge_s = 0, go_t = 0, ge_t = 0;
    hipStream_t stream = nullptr;      // where the solver's work goes: own_stream or the caller's
    hipStream_t own_stream = nullptr;
    // checkpointed backward: walks run on a second stream while the next
    // chunk is recomputed; events order the two (created on first use)
    hipStream_t aux_stream = nullptr;
    std::vector<hipEvent_t> ck_events;
    bool external_stream = false;      // pa_solver_set_stream: results ordered on the caller's stream
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cus = 256;
    bool has_device = false;
    DevBuf b_border_s, b_border_t, b_probs, b_groups, b_queue, b_cs, b_ct, b_table, b_dense, b_rowbuf;
    DevBuf b_scores, b_seeds, b_tb, b_path, b_path_off, b_path_len, b_dvals;
    DevBuf b_progress, b_partials, b_band_err, b_band_probs;
    // general gap costs: w(k) * sgn for k = 0 .. lim (GeneralGapCostSolver's tensors)
    std::vector<float> ws_sgn, wt_sgn;
    bool gen_ready = false;
    DevBuf b_ws, b_wt;
    // raw symbols encoded on the device (pa_ragged.raw_width)
    DevBuf b_raw_s, b_raw_t, b_lut, b_bad;
    DevBuf b_safe, b_sink;  // KArgs::safe, KArgs::sink
    DevBuf b_finite;        // KArgs::finite (series form)
    DevBuf b_ckpt_tb;       // checkpointed backward: recomputed tracebacks (scratch)
    DevBuf b_tape_borders;  // a checkpointed tape's copy of the border vectors (recycled)
    // pinned staging of border uploads after a gap change (pa_solver_set_gap_params)
    float *border_stage = nullptr;
    size_t border_stage_cap = 0;
    hipEvent_t border_ev = nullptr;
    bool border_pending = false;
    // checkpointed backward: scratch words per chunk when set (> 0), else
    // sized from the free device memory
    int64_t ckpt_words = std::getenv("PYALIGN_AMD_CKPT_SCRATCH_WORDS")
                             ? std::atoll(std::getenv("PYALIGN_AMD_CKPT_SCRATCH_WORDS"))
                             : 0;
    int64_t last_bad[2] = {-1, -1};  // first symbol outside the alphabet, per side
    uint32_t *band_err_host = nullptr;  // pinned: the last band-mode launch's failure flag
    // pinned staging of the plan upload: from pageable memory hipMemcpyAsync
    // would wait for the stream; staged, the host returns at once, and the
    // next upload waits only for the previous one's copies (stage_ev)
    char *stage = nullptr;
    size_t stage_cap = 0;
    hipEvent_t stage_ev = nullptr;
    bool stage_pending = false;
    bool band_check = false;            // a band-mode flag is on its way to band_err_host
    // launch sequences replayed from a hipGraph when they repeat
    bool use_graphs = true;
    // fast kernel: single-symbol LDS table even where the pair table fits
    bool no_pair = std::getenv("PYALIGN_AMD_NO_PAIR_TABLE") != nullptr;
    // dense input: per-cell loads even where rows could be staged in LDS
    bool no_drows = std::getenv("PYALIGN_AMD_NO_DENSE_STAGING") != nullptr;
    // a tape's global pairs longer than this (max of the two lengths) keep
    // only strip boundary rows and are recomputed strip by strip in the
    // backward; negative: never
    int64_t ckpt_min = 2048;
    std::string prev_key, graph_key;
    hipGraphExec_t graph_exec = nullptr;
    double last_ms = 0.0;
    int last_launches = 0;
    double last_cells = 0.0;
    int last_groups = 0, last_band = 0, last_ckpt = 0;
    // tapes made by this solver and not yet freed: a freed tape hands its
    // device buffers back (the next forward reuses them instead of a
    // hipMalloc / hipFree pair per step); the solver going first detaches them
    std::vector<pa_tape *> tapes;

    ~pa_solver();
};

namespace {

// One problem's all-optimal enumeration state (TracebackIterators<true>,
// solver.h:1568-1743).
struct IterState {
    struct Entry {
        float val;
        int32_t u, v;
        int64_t path_len;
    };
    std::vector<Entry> stack;
    std::vector<std::pair<int32_t, int32_t>> path;
    std::set<std::vector<std::pair<int32_t, int32_t>>> emitted;
    std::vector<std::pair<int32_t, int32_t>> cur_path;
    bool started = false;
};

}  // namespace

struct pa_result {
    int64_t n = 0;
    int dir = 0, loc = 0, gap = 0, goal = 0;
    int index_bits = 16;
    bool remove_dup = true;
    std::vector<float> scores;
    std::vector<int32_t> seeds;
    std::vector<int32_t> len_s, len_t;
    // one-optimal paths
    std::vector<int64_t> path_off;
    std::vector<int32_t> path_len, paths;
    // all-optimal: traceback bits and local D values; Solution matrices:
    // the traceback bits and every layer's values, plus the borders
    std::vector<uint32_t> tb;
    std::vector<float> dvals;
    bool matrices = false;
    int layers = 1, tr = TR_SCORE;
    std::vector<float> border_s, border_t;
    std::vector<Prob> probs;  // indexed by problem
    std::vector<std::unique_ptr<IterState>> iters;
    // general gap costs: dvals holds WSB_PLANES planes per problem; w(k) * sgn
    std::vector<float> ws_sgn, wt_sgn;
};

// A forward pass kept for differentiation: the schedule's problem
// descriptors, the seeds and the traceback bits, moved out of the solver's
// scratch buffers so later solves cannot overwrite them.
struct pa_tape {
    pa_solver *solver = nullptr;
    int64_t n = 0, n_probs = 0;
    int loc = 0, gap = 0, form = 0, dir = 0, dim = 0;
    DevBuf probs, seeds, tb;
    const float *dense = nullptr;
    const float *xs = nullptr, *xt = nullptr;
    // checkpointed problems: their strip boundary rows, the forward's kernel
    // arguments (to recompute a strip) and their descriptors
    DevBuf rowbuf;
    DevBuf borders;  // the forward's border vectors (the solver's may change before the backward)
    pa::KArgs kargs{};
    std::vector<pa::Prob> ckpt;
    int64_t band_slots = 0;
    // backward scratch
    DevBuf s_probs, s_groups, s_queue, s_progress, s_partials, s_err, s_state, s_strip, s_scores;
};

pa_solver::~pa_solver() {
    for (pa_tape *t : tapes) t->solver = nullptr;
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (band_err_host) (void)hipHostFree(band_err_host);
    if (stage_pending) (void)hipEventSynchronize(stage_ev);
    if (stage) (void)hipHostFree(stage);
    if (stage_ev) (void)hipEventDestroy(stage_ev);
    if (border_pending) (void)hipEventSynchronize(border_ev);
    if (border_stage) (void)hipHostFree(border_stage);
    if (border_ev) (void)hipEventDestroy(border_ev);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (own_stream) (void)hipStreamDestroy(own_stream);
    if (aux_stream) (void)hipStreamDestroy(aux_stream);
    for (hipEvent_t e : ck_events) (void)hipEventDestroy(e);
}

namespace {

void make_borders(pa_solver &s) {
    const pa_options &o = s.opt;
    const int64_t ns = s.lim_s + 1, nt = s.lim_t + 1;
    const float sgn = o.dir == PA_MINIMIZE ? 1.0f : -1.0f;
    s.border_s.assign(size_t(ns), 0.0f);
    s.border_t.assign(size_t(nt), 0.0f);
    if (o.type == PA_TYPE_DTW) {
        // DynamicTimeSolver: every border cell is the worst value except
        // D(-1,-1) = 0 (solver.h:3205-3210)
        std::fill(s.border_s.begin(), s.border_s.end(), worst(o.dir));
        std::fill(s.border_t.begin(), s.border_t.end(), worst(o.dir));
        s.border_s[0] = s.border_t[0] = 0.0f;
        return;
    }
    if (o.loc != PA_GLOBAL) return;  // Local/Semiglobal fill 0 (solver.h:2021-2029, 2253-2261)
    if (o.gap == GAP_GEN) {
        // w(k) * sgn (solver.h:3114-3123); zeros until the costs are given
        if (s.gen_ready) {
            for (int64_t k = 0; k < ns; k++) s.border_s[size_t(k)] = s.ws_sgn[size_t(k)];
            for (int64_t k = 0; k < nt; k++) s.border_t[size_t(k)] = s.wt_sgn[size_t(k)];
            s.border_s[0] = s.border_t[0];
        }
        return;
    }
    if (o.gap == GAP_LIN) {
        // arange<index>(0, n) * gap * sgn (solver.h:2806-2814)
        for (int64_t k = 0; k < ns; k++) s.border_s[size_t(k)] = (float(k) * o.lin_s) * sgn;
        for (int64_t k = 0; k < nt; k++) s.border_t[size_t(k)] = (float(k) * o.lin_t) * sgn;
    } else {
        // w(k) * sgn for D's border column/row (solver.h:2943-2951)
        const std::vector<float> ws = affine_vector(o.aff_u_s, o.aff_v_s, ns);
        const std::vector<float> wt = affine_vector(o.aff_u_t, o.aff_v_t, nt);
        for (int64_t k = 0; k < ns; k++) s.border_s[size_t(k)] = ws[size_t(k)] * sgn;
        for (int64_t k = 0; k < nt; k++) s.border_t[size_t(k)] = wt[size_t(k)] * sgn;
    }
    // D(-1,-1) is written last by the row initialisation
    s.border_s[0] = s.border_t[0];
}

// After a synchronisation: a band-mode wait that gave up leaves wrong
// results behind, so the call fails.
void check_band(pa_solver &s) {
    if (!s.band_check) return;
    s.band_check = false;
    if (*s.band_err_host)
        fail(PA_ERUNTIME, "band-mode strip pipeline stalled (a strip waited too long for the one above it)");
}

void ensure_device(pa_solver &s) {
    if (s.has_device) return;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0)
        fail(PA_EHIP, "no HIP device is available: the pyalign_amd solver runs only on an AMD GPU");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    s.cus = prop.multiProcessorCount;
    HIPCHK(hipStreamCreate(&s.own_stream));
    s.stream = s.own_stream;
    HIPCHK(hipEventCreate(&s.ev0));
    HIPCHK(hipEventCreate(&s.ev1));
    HIPCHK(hipEventCreateWithFlags(&s.stage_ev, hipEventDisableTiming));
    HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&s.band_err_host), 4, hipHostMallocDefault));
    *s.band_err_host = 0;
    s.b_border_s.ensure(s.border_s.size() * 4);
    s.b_border_t.ensure(s.border_t.size() * 4);
    {
        float safe[SAFE_FLOATS] = {};
        safe[0] = -std::numeric_limits<float>::infinity();
        safe[1] = std::numeric_limits<float>::infinity();
        s.b_safe.ensure(sizeof(safe));
        HIPCHK(hipMemcpy(s.b_safe.p, safe, sizeof(safe), hipMemcpyHostToDevice));
        s.b_sink.ensure(SINK_WORDS * 4);
    }
    HIPCHK(hipMemcpy(s.b_border_s.p, s.border_s.data(), s.border_s.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s.b_border_t.p, s.border_t.data(), s.border_t.size() * 4, hipMemcpyHostToDevice));
    s.has_device = true;
}

// ---------------------------------------------------------------------------
// scheduling
// ---------------------------------------------------------------------------

int ilog2(int x) {
    int r = 0;
    while ((1 << r) < x) r++;
    return r;
}

// Choose (L, R) per problem.  A lane-step costs STEP_OVERHEAD + R cell
// units (the per-step shifts, symbol load and masking are worth about four
// cells), so the layout cost is L * strips * (len_t + L - 1) * (4 + R) and a
// segment's serial time is that over L.  Take the cheapest layout whose
// serial time stays under `cap` (which keeps several wavefronts of work per
// SIMD and the longest problem off the critical path); problems that fit no
// layout take the one with the least serial time.
#ifndef PA_STEP_OVERHEAD
#define PA_STEP_OVERHEAD 4.0
#endif
constexpr double STEP_OVERHEAD = PA_STEP_OVERHEAD;

bool choose_layout(int64_t ls, int64_t lt, const std::vector<int> &Rs, double cap, int &L_out, int &R_out,
                   double *cost_out = nullptr, double *serial_out = nullptr) {
    double best_cost = 1e300, best_serial = 1e300;
    int bL = 64, bR = Rs.back();
    double fb_serial = 1e300, fb_cost = 1e300;
    int fL = 64, fR = Rs.back();
    for (int L = 1; L <= 64; L <<= 1) {
        for (int R : Rs) {
            const int64_t rps = int64_t(L) * R;
            const int64_t nst = (ls + rps - 1) / rps;
            const double serial = double(nst) * double(lt + L - 1) * (STEP_OVERHEAD + R);
            const double cost = serial * L;
            if (serial < fb_serial || (serial == fb_serial && cost < fb_cost)) {
                fb_serial = serial;
                fb_cost = cost;
                fL = L;
                fR = R;
            }
            if (serial > cap) continue;
            if (cost < best_cost || (cost == best_cost && serial < best_serial)) {
                best_cost = cost;
                best_serial = serial;
                bL = L;
                bR = R;
            }
        }
    }
    const bool fit = best_cost != 1e300;
    if (!fit) {
        bL = fL;
        bR = fR;
        best_cost = fb_cost;
        best_serial = fb_serial;
    }
    L_out = bL;
    R_out = bR;
    if (cost_out) *cost_out = best_cost;
    if (serial_out) *serial_out = best_serial;
    return fit;
}

// Band mode for a problem no layout brings under the cap: strips of 64 * 8
// rows run concurrently, each starting about 64 + BAND_PUB steps after the
// strip above, so the serial time is (len_t + 63 + (strips - 1) * (64 +
// BAND_PUB)) steps instead of strips * (len_t + 63).
constexpr int BAND_L = 64, BAND_R = 8;

double band_serial(int64_t ls, int64_t lt, int R, int *nst_out) {
    const int64_t rps = int64_t(BAND_L) * R;
    const int64_t nst = (ls + rps - 1) / rps;
    *nst_out = int(nst);
    return double(lt + BAND_L - 1 + (nst - 1) * (BAND_L + BAND_PUB)) * (STEP_OVERHEAD + R);
}

struct Job {
    int loc, dir, gap, form, tr;
    bool fast;
};

// ckpt[k] != 0: problem k keeps no traceback (a checkpointed forward): one
// wavefront sweeps its strips of BAND_L x BAND_R rows in order and keeps
// every strip's bottom row, from which the backward recomputes it.
// Long mode (pa_long_kernel) for a problem that would otherwise be the
// critical path: its strips of 64 rows on the waves of workgroups,
// long_dw phases of LONG_T steps apart within a workgroup (rows handed down
// through LDS) and a little more between workgroups, so the serial time is
// (len_t + 63 + (strips - 1) * long_dw * LONG_T * ~1.25) steps of one row.
// A step of one row costs about 50 ns (measured), three cell units of the
// general kernel (R = 8: a 12-unit step in about 200 ns) and four of the
// fast kernel's cheaper cells.
double long_serial(int64_t ls, int64_t lt, int form, bool fast, int *nst_out) {
    const int64_t nst = (ls + WAVE * LONG_R - 1) / (WAVE * LONG_R);
    *nst_out = int(nst);
    return double(lt + WAVE - 1 + double(nst - 1) * long_dw(form) * LONG_T * 1.25) * (fast ? 4.0 : 3.0);
}

Plan plan_auto(const Job &job, const int32_t *ls, const int32_t *lt, int64_t n, int simds,
               const uint8_t *ckpt = nullptr, bool allow_band = true, bool allow_long = true) {
    std::vector<int> Rs;
    if (job.fast) {
        Rs.assign(std::begin(FAST_RS), std::end(FAST_RS));
    } else {
        for (int R : {1, 2, 3, 4, 5, 6, 7, 8, 10, 12, 14, 16, 20, 24, 28, 32})
            if (supported_R(job.form, job.tr, job.fast, R)) Rs.push_back(R);
    }
    // total wavefront time at the cheapest layouts (no cap) sets the cap:
    // an eighth of a SIMD's share of the work for the fast kernel, a
    // sixteenth for the general kernel (both measured best)
    double total = 0.0;
    for (int64_t k = 0; k < n; k++) {
        if (ls[k] <= 0 || lt[k] <= 0) continue;
        int L, R;
        double cost = 0.0;
        choose_layout(ls[k], lt[k], Rs, 1e300, L, R, &cost);
        total += cost / 64.0;
    }
#ifndef PA_CAP_DIV
#define PA_CAP_DIV 8.0
#endif
#ifndef PA_CAP_DIV_GENERAL
#define PA_CAP_DIV_GENERAL 16.0
#endif
    const double cap = std::max(256.0, total / (double(simds) * (job.fast ? PA_CAP_DIV : PA_CAP_DIV_GENERAL)));
    // band mode only for a problem that would otherwise be the critical path
    // (serial time above the balanced per-SIMD total); its strips run in the
    // general kernel, several times slower per cell than the fast one
#ifndef PA_BAND_CRIT
#define PA_BAND_CRIT 4.0
#endif
    const double crit = PA_BAND_CRIT * cap;
    const double band_penalty = job.fast ? 3.0 : 1.0;

    struct Item {
        int64_t k;
        int L, R, nst;
        int64_t lt;
        double band;  // > 0: band mode, its serial time
        bool ckpt;    // a checkpointed problem (strip rows kept, no traceback)
    };
    const bool band_ok = allow_band && job.tr != TR_ALL && supported_R(job.form, job.tr, false, BAND_R);
    const bool long_ok = allow_band && allow_long && !std::getenv("PYALIGN_AMD_NO_LONG_MODE") &&
                         long_supported(job.gap, job.form, job.tr);
    std::vector<Item> items(static_cast<size_t>(n));
    std::vector<Item> bands, ckpts, longs;
    for (int64_t k = 0; k < n; k++) {
        int L = 64, R = Rs.back();
        double serial = 0.0;
        bool fit = true;
        if (ls[k] > 0 && lt[k] > 0) fit = choose_layout(ls[k], lt[k], Rs, cap, L, R, nullptr, &serial);
        else { L = 1; R = Rs.front(); }
        const int64_t rps = int64_t(L) * R;
        const int nst = ls[k] > 0 && lt[k] > 0 ? int((ls[k] + rps - 1) / rps) : 0;
        if (ckpt && ckpt[k]) {
            // checkpointed: one wavefront sweeps the strips of BAND_L x BAND_R
            // rows in order (the layout the backward's strip recomputation
            // uses), keeping every strip's bottom row and no traceback
            const int64_t crs = int64_t(BAND_L) * BAND_R;
            ckpts.push_back(Item{k, BAND_L, BAND_R, int((ls[k] + crs - 1) / crs), lt[k], 0.0, true});
            items[size_t(k)] = Item{k, 0, 0, 0, 0, 0.0, false};
            continue;
        }
        // where the long mode serves the job it replaces band mode (whose
        // strips hand rows down through published column counts, several
        // times slower per step than the long mode's LDS rings)
        if (!fit && long_ok && serial > crit) {
            int nl = 0;
            const double sl = long_serial(ls[k], lt[k], job.form, job.fast, &nl);
            // (at least one strip per compute wave of a workgroup: shorter
            // problems leave its waves idle)
            if (nl >= LONG_WAVES && sl < serial) {
                longs.push_back(Item{k, WAVE, LONG_R, nl, lt[k], sl, false});
                items[size_t(k)] = Item{k, 0, 0, 0, 0, 0.0, false};
                continue;
            }
        }
        if (!fit && band_ok && !long_ok && serial > crit) {
            int nb = 0;
            const int bR = band_R(job.form, job.tr);
            const double sb = band_serial(ls[k], lt[k], bR, &nb) * band_penalty;
            if (nb >= 2 && sb < serial) {
                bands.push_back(Item{k, BAND_L, bR, nb, lt[k], sb, false});
                items[size_t(k)] = Item{k, 0, 0, 0, 0, 0.0, false};
                continue;
            }
        }
        items[size_t(k)] = Item{k, L, R, nst, lt[k], 0.0, false};
    }
    if (!bands.empty() || !ckpts.empty() || !longs.empty())
        items.erase(std::remove_if(items.begin(), items.end(), [](const Item &it) { return it.L == 0; }),
                    items.end());
    // group identical layouts with similar shape: sort by R, L, strips, len_t
    std::stable_sort(items.begin(), items.end(), [](const Item &a, const Item &b) {
        if (a.R != b.R) return a.R < b.R;
        if (a.L != b.L) return a.L < b.L;
        if (a.nst != b.nst) return a.nst > b.nst;
        return a.lt > b.lt;
    });
    Plan plan;
    plan.probs.reserve(static_cast<size_t>(n));
    struct GroupR {
        Group g;
        int R;
    };
    std::vector<GroupR> gs;
    size_t p = 0;
    while (p < items.size()) {
        const int L = items[p].L, R = items[p].R;
        const int per = 64 / L;
        Group g{};
        g.first = int32_t(plan.probs.size());
        g.L = L;
        g.lgL = ilog2(L);
        g.strip = -1;
        int cnt = 0;
        while (p < items.size() && cnt < per && items[p].L == L && items[p].R == R) {
            Prob pr{};
            pr.out = int32_t(items[p].k);
            pr.L = L;
            pr.R = R;
            pr.band = -1;
            plan.probs.push_back(pr);
            g.max_lt = std::max<int32_t>(g.max_lt, int32_t(items[p].lt));
            g.max_nst = std::max<int32_t>(g.max_nst, items[p].nst);
            cnt++;
            p++;
        }
        g.count = cnt;
        g.R = R;
        g.cost = int64_t(double(g.max_nst) * double(g.max_lt + L - 1) * (STEP_OVERHEAD + R));
        gs.push_back(GroupR{g, job.fast ? 0 : R});
    }
    // checkpointed problems: a group each; band slots only for the
    // backward's strip recomputation (progress counters and partials)
    for (const Item &it : ckpts) {
        Prob pr{};
        pr.out = int32_t(it.k);
        pr.L = BAND_L;
        pr.R = BAND_R;
        pr.band = int32_t(plan.band_slots);  // also: keep every strip's bottom row
        pr.tb_off = NO_TB_OFF;
        plan.band_slots += it.nst;
        Group g{};
        g.first = int32_t(plan.probs.size());
        g.count = 1;
        g.L = BAND_L;
        g.lgL = ilog2(BAND_L);
        g.max_lt = int32_t(it.lt);
        g.max_nst = it.nst;
        g.R = BAND_R;
        g.strip = -1;
        g.cost = int64_t(double(it.nst) * double(it.lt + BAND_L - 1) * (STEP_OVERHEAD + BAND_R));
        gs.push_back(GroupR{g, job.fast ? 1000 + BAND_R : BAND_R});
        plan.probs.push_back(pr);
    }
    // long-mode problems: no groups; pa_long_kernel takes their bands in
    // order; listed first among the band problems (their strips' partials
    // are folded like band-mode strips')
    std::stable_sort(longs.begin(), longs.end(), [](const Item &a, const Item &b) { return a.band > b.band; });
    for (const Item &it : longs) {
        Prob pr{};
        pr.out = int32_t(it.k);
        pr.L = WAVE;
        pr.R = LONG_R;
        pr.band = int32_t(plan.band_slots);
        plan.band_slots += it.nst;
        plan.band_probs.push_back(int32_t(plan.probs.size()));
        plan.long_bands += (it.nst + LONG_WAVES - 1) / LONG_WAVES;
        plan.n_long++;
        plan.probs.push_back(pr);
    }
    // band-mode problems: one group per strip, strips in order, run by the
    // general kernel (its launch goes first: it holds the longest problems)
    std::stable_sort(bands.begin(), bands.end(), [](const Item &a, const Item &b) { return a.band > b.band; });
    for (const Item &it : bands) {
        Prob pr{};
        pr.out = int32_t(it.k);
        pr.L = BAND_L;
        pr.R = it.R;
        pr.band = int32_t(plan.band_slots);
        if (it.ckpt) pr.tb_off = NO_TB_OFF;
        plan.band_slots += it.nst;
        plan.band_probs.push_back(int32_t(plan.probs.size()));
        for (int st = 0; st < it.nst; st++) {
            Group g{};
            g.first = int32_t(plan.probs.size());
            g.count = 1;
            g.L = BAND_L;
            g.lgL = ilog2(BAND_L);
            g.max_lt = int32_t(it.lt);
            g.max_nst = it.nst;
            g.R = it.R;
            g.strip = st;
            g.cost = int64_t(it.band);
            // with the other groups of its strip height, or (fast jobs) in a
            // general-kernel launch of its own, marked by 1000 + R
            gs.push_back(GroupR{g, job.fast ? 1000 + it.R : it.R});
        }
        plan.probs.push_back(pr);
    }
    // one launch per kernel (the fast kernel serves every R); inside a
    // launch the longest groups go first (LPT)
    std::stable_sort(gs.begin(), gs.end(), [](const GroupR &a, const GroupR &b) {
        if (a.R != b.R) return a.R > b.R;
        return a.g.cost > b.g.cost;
    });
    size_t q = 0;
    while (q < gs.size()) {
        const int R = gs[q].R;
        const int first = int(plan.groups.size());
        while (q < gs.size() && gs[q].R == R) plan.groups.push_back(gs[q++].g);
        plan.launches.push_back({R, {first, int(plan.groups.size())}});
    }
    return plan;
}

// The plain form: the reference's own batching (ProblemBag.batches,
// instance.py:230-275) with 64 lanes, one problem per lane (L = 1), every
// lane padded to its batch's master shape, groups in generator order.
Plan plan_plain(const Job &job, const int32_t *ls, const int32_t *lt, int64_t n) {
    const int B = 64;
    const int R = (job.form == F_TABLE && job.tr == TR_SCORE) ? 16 : 8;
    std::vector<uint64_t> code(static_cast<size_t>(n));
    for (int64_t k = 0; k < n; k++) code[size_t(k)] = morton(uint64_t(ls[k]), uint64_t(lt[k]));
    std::vector<int64_t> sorted(static_cast<size_t>(n));
    std::iota(sorted.begin(), sorted.end(), 0);
    std::stable_sort(sorted.begin(), sorted.end(),
                     [&](int64_t a, int64_t b) { return code[size_t(a)] < code[size_t(b)]; });
    std::vector<char> used(size_t(n), 0);
    Plan plan;
    while (true) {
        while (!sorted.empty() && used[size_t(sorted.back())]) sorted.pop_back();
        if (sorted.empty()) break;
        const int64_t m = sorted.back();
        sorted.pop_back();
        std::vector<int64_t> idx{m};
        const int64_t ms = ls[m], mt = lt[m];
        for (size_t kk = 1; kk <= sorted.size(); kk++) {
            const int64_t i = sorted[sorted.size() - kk];
            if (used[size_t(i)]) continue;
            if (ls[i] <= ms && lt[i] <= mt) {
                idx.push_back(i);
                used[size_t(i)] = 1;
                if (int(idx.size()) == B) break;
            }
        }
        Group g{};
        g.first = int32_t(plan.probs.size());
        g.count = int32_t(idx.size());
        g.strip = -1;
        g.L = 1;
        g.lgL = 0;
        g.max_lt = int32_t(mt);
        g.max_nst = int32_t((ms + R - 1) / R);
        g.R = R;
        g.cost = int64_t(g.max_nst) * mt * R;
        for (int64_t i : idx) {
            Prob pr{};
            pr.out = int32_t(i);
            pr.band = -1;
            pr.L = 1;
            pr.R = R;
            plan.probs.push_back(pr);
        }
        plan.groups.push_back(g);
    }
    // same kernel as the batched form: only the schedule differs
    plan.launches.push_back({job.fast ? 0 : R, {0, int(plan.groups.size())}});
    return plan;
}

// ---------------------------------------------------------------------------
// a solve
// ---------------------------------------------------------------------------

struct Input {
    int form;  // PA_FORM_*
    bool on_device;
    const void *codes_s, *codes_t;
    int64_t bytes_s, bytes_t;  // payload sizes (host uploads)
    const int64_t *off_s, *off_t;
    const int32_t *len_s, *len_t;
    int64_t n;
    std::vector<float> table;  // (K+1)^2 with pad, kernel layout
    int K1 = 1;
    float eq = 0, ne = 0;
    const float *dense = nullptr;
    int64_t dense_elems = 0;
    const int64_t *dense_off = nullptr;
    int32_t dense_si_fixed = 0, dense_sj_fixed = 0;  // reference layout strides (0 = row-major)
    int dim = 0;
    int schedule = PA_SCHED_AUTO;
    bool nan_free = true;
    bool matrices = false;  // keep every layer's cell values (Solution matrices)
    // raw symbols to encode on the device: 1 or 4 bytes each (0: codes given)
    int raw_width = 0;
    const uint8_t *lut = nullptr;  // indexed form: symbol -> code, 255 = none (s side)
    int32_t lut_len = 0;
    const uint8_t *lut_t = nullptr;  // t side's table (null: the s side's)
    int32_t lut_t_len = 0;
    int64_t syms_s = 0, syms_t = 0;  // symbols per side
    // identity tables that check host uint8 codes on the device (s codes
    // below the table's rows, t codes below its columns)
    std::vector<uint8_t> lut_store, lut_t_store;
    const char *bad_s = "item not in alphabet", *bad_t = "item not in alphabet";
    // indexed form: the table is read from global memory with uint16 codes
    // (F_TABLE_G) instead of staged in LDS with uint8 codes
    bool table_g = false;
};

int goal_to_tr(int goal) { return goal == PA_GOAL_SCORE ? TR_SCORE : (goal == PA_GOAL_ONE ? TR_ONE : TR_ALL); }

template <class T>
void upload(hipStream_t st, DevBuf &buf, const T *src, size_t count) {
    buf.ensure(std::max<size_t>(count * sizeof(T), 4));
    if (count) HIPCHK(hipMemcpyAsync(buf.p, src, count * sizeof(T), hipMemcpyHostToDevice, st));
}

// Device copies of the problem payload (uploaded, or on the device already)
// and, for raw symbols, their codes.  Fails with ValueError on a symbol
// outside the alphabet and records where (pa_solver_last_bad_symbol).
void stage_payload(pa_solver &s, const Input &in, hipStream_t st, const void *&codes_s, const void *&codes_t,
                   const float *&dense) {
    codes_s = in.codes_s;
    codes_t = in.codes_t;
    dense = in.dense;
    if (in.form == PA_FORM_DENSE) {
        if (!in.on_device) {
            upload(st, s.b_dense, in.dense, size_t(in.dense_elems));
            dense = s.b_dense.as<float>();
        }
        return;
    }
    if (!in.raw_width) {
        if (!in.on_device) {
            upload(st, s.b_cs, static_cast<const uint8_t *>(in.codes_s), size_t(in.bytes_s));
            upload(st, s.b_ct, static_cast<const uint8_t *>(in.codes_t), size_t(in.bytes_t));
            codes_s = s.b_cs.p;
            codes_t = s.b_ct.p;
        }
        return;
    }
    const void *raw_s = in.codes_s, *raw_t = in.codes_t;
    if (!in.on_device) {
        upload(st, s.b_raw_s, static_cast<const uint8_t *>(in.codes_s), size_t(in.bytes_s));
        upload(st, s.b_raw_t, static_cast<const uint8_t *>(in.codes_t), size_t(in.bytes_t));
        raw_s = s.b_raw_s.p;
        raw_t = s.b_raw_t.p;
    }
    const bool indexed = in.form == PA_FORM_INDEXED;
    const int cw = indexed ? (in.table_g ? 2 : 1) : 4;  // code width
    s.b_cs.ensure(size_t(std::max<int64_t>(in.syms_s, 1)) * cw);
    s.b_ct.ensure(size_t(std::max<int64_t>(in.syms_t, 1)) * cw);
    const uint8_t *lut = nullptr, *lut_t = nullptr;
    int32_t lut_t_len = in.lut_len;
    if (indexed) {
        const size_t at = (size_t(in.lut_len) + 255) / 256 * 256;
        lut_t_len = in.lut_t ? in.lut_t_len : in.lut_len;
        s.b_lut.ensure(at + size_t(std::max(lut_t_len, 1)));
        HIPCHK(hipMemcpyAsync(s.b_lut.p, in.lut, size_t(in.lut_len), hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(static_cast<char *>(s.b_lut.p) + at, in.lut_t ? in.lut_t : in.lut, size_t(lut_t_len),
                              hipMemcpyHostToDevice, st));
        lut = s.b_lut.as<uint8_t>();
        lut_t = lut + at;
    }
    s.b_bad.ensure(16);
    unsigned long long *bad = s.b_bad.as<unsigned long long>();
    HIPCHK(launch_fill_u32(s.b_bad.as<uint32_t>(), 4, 0xffffffffu, st));
    HIPCHK(launch_translate(raw_s, in.raw_width, in.syms_s, lut, in.lut_len, s.b_cs.p, cw, bad, s.cus, st));
    HIPCHK(launch_translate(raw_t, in.raw_width, in.syms_t, lut_t, lut_t_len, s.b_ct.p, cw, bad + 1, s.cus, st));
    codes_s = s.b_cs.p;
    codes_t = s.b_ct.p;
    if (indexed) {
        unsigned long long first[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(first, bad, sizeof(first), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int side = 0; side < 2; side++) s.last_bad[side] = first[side] == ~0ull ? -1 : int64_t(first[side]);
        if (s.last_bad[0] >= 0) fail(PA_EINVAL, in.bad_s);
        if (s.last_bad[1] >= 0) fail(PA_EINVAL, in.bad_t);
    }
}


// General gap costs (Waterman-Smith-Beyer): one workgroup per problem, no
// strip schedule.  The kernel leaves D, S and the traceback codes of every
// cell in device memory; all-optimal enumeration and Solution matrices read
// them on the host.
void run_wsb(pa_solver &s, const Input &in, const Job &job, std::unique_ptr<pa_result> *result,
             float *scores_device) {
    const pa_options &o = s.opt;
    const int64_t n = in.n;
    if (!s.gen_ready) fail(PA_EINVAL, "general gap costs were not given to the solver");
    hipStream_t st = s.stream;
    std::vector<Prob> probs(static_cast<size_t>(n));
    int64_t words = 0;
    double cells = 0;
    for (int64_t k = 0; k < n; k++) {
        Prob &pr = probs[size_t(k)];
        std::memset(&pr, 0, sizeof(pr));
        pr.out = int32_t(k);
        pr.len_s = in.len_s[k];
        pr.len_t = in.len_t[k];
        pr.band = -1;
        if (in.form == PA_FORM_DENSE) {
            pr.off_s = in.dense_off ? in.dense_off[k] : k;
            pr.si = in.dense_si_fixed ? in.dense_si_fixed : pr.len_t;
            pr.sj = in.dense_sj_fixed ? in.dense_sj_fixed : 1;
        } else {
            pr.off_s = in.off_s[k];
            pr.off_t = in.off_t[k];
        }
        pr.dv_off = words;
        words += int64_t(WSB_PLANES) * pr.len_s * pr.len_t;
        cells += double(pr.len_s) * double(pr.len_t);
    }
    upload(st, s.b_probs, probs.data(), probs.size());
    s.last_plan.uploaded = false;  // b_probs now holds these descriptors
    upload(st, s.b_ws, s.ws_sgn.data(), s.ws_sgn.size());
    upload(st, s.b_wt, s.wt_sgn.data(), s.wt_sgn.size());
    const void *codes_s = nullptr, *codes_t = nullptr;
    const float *dense = nullptr;
    stage_payload(s, in, st, codes_s, codes_t, dense);
    if (job.form == F_TABLE || job.form == F_TABLE_G) upload(st, s.b_table, in.table.data(), in.table.size());
    s.b_dvals.ensure(size_t(std::max<int64_t>(words, 1)) * 4);
    float *scores = scores_device;
    if (!scores) {
        s.b_scores.ensure(size_t(std::max<int64_t>(n, 1)) * 4);
        scores = s.b_scores.as<float>();
    }
    s.b_seeds.ensure(size_t(std::max<int64_t>(n, 1)) * 8);

    KArgs a;
    std::memset(&a, 0, sizeof(a));
    a.probs = s.b_probs.as<Prob>();
    a.codes_s = codes_s;
    a.codes_t = codes_t;
    a.table = s.b_table.as<float>();
    a.K1 = in.K1;
    a.eq = in.eq;
    a.ne = in.ne;
    a.dense = dense;
    a.border_s = s.b_border_s.as<float>();
    a.border_t = s.b_border_t.as<float>();
    a.safe = s.b_safe.as<float>();
    a.scores = scores;
    a.seeds = s.b_seeds.as<int32_t>();
    a.dvals = s.b_dvals.as<float>();
    a.val_layers = 1;

    std::unique_ptr<pa_result> r;
    int32_t *path = nullptr, *path_len = nullptr;
    const int64_t *path_off = nullptr;
    if (result) {
        r = std::make_unique<pa_result>();
        if (job.tr == TR_ONE) {
            r->path_off.resize(size_t(n) + 1);
            r->path_off[0] = 0;
            for (int64_t k = 0; k < n; k++)
                r->path_off[size_t(k) + 1] = r->path_off[size_t(k)] + in.len_s[k] + in.len_t[k] + 2;
            upload(st, s.b_path_off, r->path_off.data(), r->path_off.size());
            s.b_path.ensure(size_t(std::max<int64_t>(r->path_off.back(), 1)) * 8);
            s.b_path_len.ensure(size_t(std::max<int64_t>(n, 1)) * 4);
            path = s.b_path.as<int32_t>();
            path_off = s.b_path_off.as<int64_t>();
            path_len = s.b_path_len.as<int32_t>();
        }
    }
    // workgroup size: 256 threads when the problems fill the CUs, else up to
    // 1024 (one thread per cell of the longest anti-diagonal) for more loads
    // in flight per problem
    int threads = 256;
    if (n < s.cus) {
        int64_t diag = 0;
        for (int64_t k = 0; k < n; k++) diag = std::max<int64_t>(diag, std::min(in.len_s[k], in.len_t[k]));
        threads = int(std::min<int64_t>(1024, std::max<int64_t>(256, (diag + 63) / 64 * 64)));
    }
    HIPCHK(hipEventRecord(s.ev0, st));
    const hipError_t le = launch_wsb(job.loc, job.dir, job.form, a, n, s.b_ws.as<float>(), s.b_wt.as<float>(),
                                     path, path_off, path_len, threads, st);
    if (le != hipSuccess) {
        std::ostringstream m;
        m << "kernel launch failed (general gap costs, loc " << job.loc << " dir " << job.dir << " form "
          << job.form << "): " << hipGetErrorString(le);
        fail(PA_EHIP, m.str());
    }
    HIPCHK(hipEventRecord(s.ev1, st));
    s.last_launches = 1;
    s.last_cells = cells;
    s.last_groups = 0;
    s.last_band = 0;
    s.last_ckpt = 0;
    if (!result) return;  // device scores: caller synchronises
    r->n = n;
    r->dir = o.dir;
    r->loc = job.loc;
    r->gap = job.gap;
    r->goal = o.goal;
    r->index_bits = o.index_bits;
    r->remove_dup = o.remove_dup;
    r->len_s.assign(in.len_s, in.len_s + n);
    r->len_t.assign(in.len_t, in.len_t + n);
    r->scores.resize(size_t(n));
    r->seeds.resize(size_t(n) * 2);
    if (job.tr == TR_ONE) {
        r->path_len.resize(size_t(n));
        r->paths.resize(size_t(r->path_off.back()) * 2);
        HIPCHK(hipMemcpyAsync(r->path_len.data(), path_len, size_t(n) * 4, hipMemcpyDeviceToHost, st));
        if (!r->paths.empty())
            HIPCHK(hipMemcpyAsync(r->paths.data(), path, r->paths.size() * 4, hipMemcpyDeviceToHost, st));
    }
    if (n) {
        HIPCHK(hipMemcpyAsync(r->scores.data(), scores, size_t(n) * 4, hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(r->seeds.data(), a.seeds, size_t(n) * 8, hipMemcpyDeviceToHost, st));
    }
    r->matrices = in.matrices;
    r->layers = 1;
    r->tr = job.tr;
    if (job.tr == TR_ALL || in.matrices) {
        r->border_s.assign(s.border_s.begin(), s.border_s.end());
        r->border_t.assign(s.border_t.begin(), s.border_t.end());
        r->ws_sgn = s.ws_sgn;
        r->wt_sgn = s.wt_sgn;
        r->dvals.resize(size_t(words));
        if (words) HIPCHK(hipMemcpyAsync(r->dvals.data(), a.dvals, size_t(words) * 4, hipMemcpyDeviceToHost, st));
        r->probs = probs;
        if (job.tr == TR_ALL) r->iters.resize(size_t(n));
    }
    HIPCHK(hipStreamSynchronize(st));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, s.ev0, s.ev1));
    s.last_ms = ms;
    *result = std::move(r);
}

bool is_long(const Plan &pl, const Prob &pr) {
    for (int32_t q = 0; q < pl.n_long; q++)
        if (&pl.probs[size_t(pl.band_probs[size_t(q)])] == &pr) return true;
    return false;
}

// A plan's problems get their lengths and offsets: payload, traceback
// words, strip-boundary row buffer and cell values.
void place_plan(Plan &pl, const Input &in, int bits, bool keep_vals, int val_layers) {
    for (Prob &pr : pl.probs) {
        const int64_t k = pr.out;
        pr.len_s = in.len_s[k];
        pr.len_t = in.len_t[k];
        if (in.form == PA_FORM_DENSE) {
            pr.off_s = in.dense_off ? in.dense_off[k] : k;
            pr.off_t = 0;
            pr.si = in.dense_si_fixed ? in.dense_si_fixed : pr.len_t;
            pr.sj = in.dense_sj_fixed ? in.dense_sj_fixed : 1;
        } else {
            pr.off_s = in.off_s[k];
            pr.off_t = in.off_t[k];
        }
        const int64_t rps = int64_t(pr.L) * pr.R;
        const int64_t nst = (pr.len_s > 0 && pr.len_t > 0) ? (pr.len_s + rps - 1) / rps : 0;
        pr.pack = 0;
        if (bits && pr.tb_off != NO_TB_OFF) {
            // 16-byte aligned for problems whose steps are whole 64-word
            // rows (the walk loads them four words at a time)
            if (pr.L == WAVE && pr.R == 1) pl.tb_words = (pl.tb_words + 3) & ~int64_t(3);
            pr.tb_off = pl.tb_words;
            if (is_long(pl, pr)) {  // codes packed along the steps
                pr.pack = long_pack(bits);
                pl.tb_words += nst * long_strip_words(pr.len_t + WAVE - 1, bits) * WAVE;
            } else {
                pl.tb_words += nst * (pr.len_t + pr.L - 1) * pr.L * ((bits * pr.R + 31) / 32);
            }
        }
        if (nst > 1 && !is_long(pl, pr)) {
            pr.rb_off = pl.rowbuf;
            pl.rowbuf += pr.band >= 0 ? (nst - 1) * int64_t(pr.len_t) : int64_t(pr.len_t);
        }
        if (keep_vals) {
            pr.dv_off = pl.dvals;
            pl.dvals += int64_t(val_layers) * pr.len_s * pr.len_t;
        }
        pl.cells += double(pr.len_s) * double(pr.len_t);
    }
    // long-mode problems hand each band's bottom row to the band below as
    // two tagged 8-byte words per column (two float2 slots), in one range
    // that is cleared before every launch
    pl.xbuf_begin = pl.rowbuf;
    for (int32_t q = 0; q < pl.n_long; q++) {
        Prob &pr = pl.probs[size_t(pl.band_probs[size_t(q)])];
        const int64_t nst = (pr.len_s + int64_t(WAVE) * LONG_R - 1) / (int64_t(WAVE) * LONG_R);
        const int64_t nb = (nst + LONG_WAVES - 1) / LONG_WAVES;
        pr.rb_off = pl.rowbuf;
        pl.rowbuf += 2 * std::max<int64_t>(nb - 1, 0) * pr.len_t;
    }
    pl.xbuf_end = pl.rowbuf;
}

// The DP launches of a plan: one spec (kernel, strip height, grid, LDS) and
// argument block per launch.  `a` holds the buffers; queue_base the launches'
// ticket counters.
void build_launches(pa_solver &s, const Job &job, const Input &in, const Plan &plan, const KArgs &a,
                    uint32_t *queue_base, std::vector<LaunchSpec> &specs, std::vector<KArgs> &largs) {
    // dense rows with unit column stride and >= DROWS_MIN_LT columns in every
    // non-empty problem: the kernel stages each lane's rows through LDS (F_DROWS)
    const int64_t n = in.n;
    bool drows = job.form == F_DENSE && (in.dense_sj_fixed == 0 || in.dense_sj_fixed == 1) && !s.no_drows;
    for (int64_t k = 0; k < n && drows; k++)
        if (in.len_s[k] > 0 && in.len_t[k] > 0 && in.len_t[k] < DROWS_MIN_LT) drows = false;
    for (size_t li = 0; li < plan.launches.size(); li++) {
        int R = plan.launches[li].first;
        bool fastmax = job.fast;
        if (R >= 1000) {  // band-mode strips of a fast job: the general kernel
            R -= 1000;
            fastmax = false;
        }
        const int g0 = plan.launches[li].second.first, g1 = plan.launches[li].second.second;
        if (g1 <= g0) continue;
        LaunchSpec spec;
        std::memset(&spec, 0, sizeof(spec));
        spec.loc = job.loc;
        spec.dir = job.dir;
        spec.gap = job.gap;
        spec.form = job.form;
        spec.tr = job.tr;
        spec.R = R;
        spec.fastmax = fastmax;
        spec.lds_bytes = job.form == F_TABLE ? size_t(in.K1) * in.K1 * 4 : 0;
        if (drows && R == DROWS_WIN) {
            spec.form = F_DROWS;
            spec.lds_bytes = DROWS_LDS;
        }
        if (fastmax && R == 0 && in.K1 >= 2) {
            const size_t pair_bytes = size_t(in.K1) * in.K1 * (in.K1 - 1) * 8;
            if (pair_bytes <= PAIR_LDS_MAX && !s.no_pair) {
                spec.pair = 1;
                spec.lds_bytes = pair_bytes;
            }
        }
        const std::vector<int> okey = {spec.loc, spec.dir, spec.gap, spec.form, spec.tr, spec.R,
                                       int(spec.fastmax), int(spec.lds_bytes), spec.pair, in.dim};
        int waves_per_cu = 0;
        for (auto &e : s.occ_cache)
            if (e.first == okey) waves_per_cu = e.second;
        if (!waves_per_cu) {
            waves_per_cu = std::max(1, occupancy_waves(spec, in.dim));
            s.occ_cache.push_back({okey, waves_per_cu});
        }
        // the fast kernel's large workgroups only where the groups fill every
        // CU with them; fewer groups spread over more, smaller workgroups
        spec.block = fastmax && R == 0 && int64_t(g1 - g0) >= int64_t(s.cus) * (FAST_BLOCK / WAVE) ? FAST_BLOCK : 256;
        const int wpb = spec.block / WAVE;  // waves per workgroup
        const int64_t wg_needed = (int64_t(g1 - g0) + wpb - 1) / wpb;
        spec.grid = int(std::max<int64_t>(1, std::min<int64_t>(wg_needed, int64_t(s.cus) * std::max(1, waves_per_cu / wpb))));
        KArgs la = a;
        la.groups = a.groups + g0;
        la.n_groups = g1 - g0;
        la.queue = queue_base + li;
        specs.push_back(spec);
        largs.push_back(la);
    }
}

void run(pa_solver &s, const Input &in, std::unique_ptr<pa_result> *result, float *scores_device,
         pa_tape *tape = nullptr) {
    ensure_device(s);
    const pa_options &o = s.opt;
    const int64_t n = in.n;
    for (int64_t k = 0; k < n; k++) {
        if (in.len_s[k] < 0 || in.len_t[k] < 0) fail(PA_EINVAL, "negative sequence length");
        if (in.len_s[k] > s.lim_s) {
            std::ostringstream e;
            e << "sequence s of length " << in.len_s[k] << " exceeds configured maximum length " << s.lim_s;
            fail(PA_ELENGTH, e.str());
        }
        if (in.len_t[k] > s.lim_t) {
            std::ostringstream e;
            e << "sequence t of length " << in.len_t[k] << " exceeds configured maximum length " << s.lim_t;
            fail(PA_ELENGTH, e.str());
        }
    }
    Job job;
    job.loc = o.type == PA_TYPE_DTW ? LOC_GLOBAL : o.loc;
    job.dir = o.dir == PA_MAXIMIZE ? DIR_MAX : DIR_MIN;
    job.gap = o.gap;
    job.tr = goal_to_tr(o.goal);
    switch (in.form) {
        case PA_FORM_DENSE: job.form = F_DENSE; break;
        case PA_FORM_INDEXED: job.form = in.table_g ? F_TABLE_G : F_TABLE; break;
        case PA_FORM_BINARY: job.form = F_BINARY; break;
        case PA_FORM_SERIES: job.form = F_SERIES; break;
        default: fail(PA_EINVAL, "unknown problem form");
    }
    if (job.form == F_SERIES && o.type != PA_TYPE_DTW) fail(PA_EINVAL, "the series form is for DTW solvers");
    if (o.type == PA_TYPE_DTW && job.form != F_DENSE && job.form != F_SERIES)
        fail(PA_EINVAL, "DTW solvers take dense costs or series");
    if (job.form == F_SERIES && (in.dim < 1 || in.dim > 4)) fail(PA_EINVAL, "series dimension must be 1..4");
    if (tape) job.tr = TR_ONE;  // differentiation walks the one-optimal path
    if (in.matrices && job.tr == TR_SCORE) job.tr = TR_ONE;
    if (job.gap == GAP_GEN) {
        if (tape) fail(PA_EINVAL, "differentiation is not available for general gap costs");
        run_wsb(s, in, job, result, scores_device);
        return;
    }
    const int val_layers = in.matrices && job.gap == GAP_AFF ? 3 : 1;
    const bool keep_vals = in.matrices || (job.tr == TR_ALL && job.loc == LOC_LOCAL);
    job.fast = job.form == F_TABLE && job.tr == TR_SCORE && job.loc == LOC_LOCAL && job.dir == DIR_MAX &&
               o.gap != GAP_DTW && in.nan_free;

    hipStream_t st = s.stream;
    // ---- schedule (reused while lengths, offsets and job repeat) ------------
    const int bits = tb_bits(job.gap, job.tr);
    const int64_t *o_s = in.form == PA_FORM_DENSE ? in.dense_off : in.off_s;
    const int64_t *o_t = in.form == PA_FORM_DENSE ? nullptr : in.off_t;
    // checkpointed forward (a tape of problems on a device payload): pairs
    // longer than the solver's threshold keep only their strip rows (and,
    // local and semiglobal, the seed the forward found)
    std::vector<uint8_t> ckpt;
    int64_t n_ckpt = 0;
    if (tape && (job.form == F_DENSE || job.form == F_SERIES) && !in.matrices && s.ckpt_min >= 0 &&
        in.schedule == PA_SCHED_AUTO) {
        ckpt.assign(size_t(n), 0);
        const int64_t rps = int64_t(BAND_L) * BAND_R;
        for (int64_t k = 0; k < n; k++) {
            const int64_t m = std::max(in.len_s[k], in.len_t[k]);
            if (m > s.ckpt_min && in.len_t[k] > 0 && in.len_s[k] > rps) {
                ckpt[size_t(k)] = 1;
                n_ckpt++;
            }
        }
    }
    // long mode packs its traceback codes along the steps, a layout only the
    // walk of a solve reads: a tape's backward walks the per-lane layout
    const bool allow_long = !keep_vals && !tape;
    const std::vector<int64_t> key = {n, job.loc, job.dir, job.gap, job.form, job.tr, int64_t(job.fast),
                                      in.schedule, in.dense_si_fixed, in.dense_sj_fixed, s.cus,
                                      int64_t(in.matrices), n_ckpt ? s.ckpt_min : -1, int64_t(allow_long)};
    PlanCache &pc = s.last_plan;
    auto same = [&]() {
        if (!pc.valid || pc.key != key) return false;
        if (std::memcmp(pc.len_s.data(), in.len_s, size_t(n) * 4) != 0) return false;
        if (std::memcmp(pc.len_t.data(), in.len_t, size_t(n) * 4) != 0) return false;
        if ((o_s == nullptr) != pc.off_s.empty() || (o_t == nullptr) != pc.off_t.empty()) return false;
        if (o_s && std::memcmp(pc.off_s.data(), o_s, size_t(n) * 8) != 0) return false;
        if (o_t && std::memcmp(pc.off_t.data(), o_t, size_t(n) * 8) != 0) return false;
        return true;
    };
    const bool reuse = same();
    if (!reuse) {
        pc.valid = false;
        pc.plan = in.schedule == PA_SCHED_PLAIN ? plan_plain(job, in.len_s, in.len_t, n)
                                                : plan_auto(job, in.len_s, in.len_t, n, s.cus * 4,
                                                            n_ckpt ? ckpt.data() : nullptr, true, allow_long);
        place_plan(pc.plan, in, bits, keep_vals, val_layers);
        pc.uploaded = false;
        pc.key = key;
        pc.len_s.assign(in.len_s, in.len_s + n);
        pc.len_t.assign(in.len_t, in.len_t + n);
        pc.off_s.clear();
        if (o_s) pc.off_s.assign(o_s, o_s + n);
        pc.off_t.clear();
        if (o_t) pc.off_t.assign(o_t, o_t + n);
        pc.valid = true;
    }
    if (!pc.uploaded) {
        const Plan &pl = pc.plan;
        const size_t nb[3] = {pl.probs.size() * sizeof(Prob), pl.groups.size() * sizeof(Group),
                              pl.band_probs.size() * sizeof(int32_t)};
        const size_t al = 256, total = (nb[0] + al) + (nb[1] + al) + (nb[2] + al);
        if (s.stage_pending) HIPCHK(hipEventSynchronize(s.stage_ev));
        s.stage_pending = false;
        if (s.stage_cap < total) {
            if (s.stage) HIPCHK(hipHostFree(s.stage));
            s.stage = nullptr;
            s.stage_cap = 0;
            HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&s.stage), total + total / 4, hipHostMallocDefault));
            s.stage_cap = total + total / 4;
        }
        const void *src[3] = {pl.probs.data(), pl.groups.data(), pl.band_probs.data()};
        DevBuf *dst[3] = {&s.b_probs, &s.b_groups, &s.b_band_probs};
        size_t at = 0;
        for (int q = 0; q < 3; q++) {
            dst[q]->ensure(std::max<size_t>(nb[q], 4));
            if (!nb[q]) continue;
            std::memcpy(s.stage + at, src[q], nb[q]);
            HIPCHK(hipMemcpyAsync(dst[q]->p, s.stage + at, nb[q], hipMemcpyHostToDevice, st));
            at += (nb[q] + al - 1) / al * al;
        }
        HIPCHK(hipEventRecord(s.stage_ev, st));
        s.stage_pending = true;
        pc.uploaded = true;
    }
    const Plan &plan = pc.plan;

    // ---- device buffers ----------------------------------------------------
    // queue tickets: one per launch, then the long-mode band tickets
    s.b_queue.ensure((plan.launches.size() + 1) * 4);
    const void *codes_s = nullptr, *codes_t = nullptr;
    const float *dense = nullptr;
    stage_payload(s, in, st, codes_s, codes_t, dense);
    if (job.form == F_TABLE || job.form == F_TABLE_G) upload(st, s.b_table, in.table.data(), in.table.size());
    s.b_rowbuf.ensure(size_t(std::max<int64_t>(plan.rowbuf, 1)) * sizeof(float2));
    float *scores = scores_device;
    if (!scores) {
        s.b_scores.ensure(size_t(std::max<int64_t>(n, 1)) * 4);
        scores = s.b_scores.as<float>();
    }
    int32_t *seeds = nullptr;
    if (job.tr != TR_SCORE) {
        s.b_seeds.ensure(size_t(std::max<int64_t>(n, 1)) * 8);
        seeds = s.b_seeds.as<int32_t>();
        s.b_tb.ensure(size_t(std::max<int64_t>(plan.tb_words, 1)) * 4);
    }
    if (plan.dvals > 0) s.b_dvals.ensure(size_t(plan.dvals) * 4);

    KArgs a;
    std::memset(&a, 0, sizeof(a));  // padding included: the bytes key the graph cache
    a.probs = s.b_probs.as<Prob>();
    a.groups = s.b_groups.as<Group>();
    a.codes_s = codes_s;
    a.codes_t = codes_t;
    a.table = s.b_table.as<float>();
    a.K1 = in.K1;
    a.dim = in.dim;
    a.eq = in.eq;
    a.ne = in.ne;
    a.dense = dense;
    a.rowbuf = s.b_rowbuf.as<float2>();
    a.border_s = s.b_border_s.as<float>();
    a.border_t = s.b_border_t.as<float>();
    a.safe = s.b_safe.as<float>();
    a.sink = s.b_sink.as<uint32_t>();
    // DTW over series: a check of the series' values before the DP launches
    // lets the kernels fold with min / max where every value is finite
    int64_t finite_nx = 0, finite_ny = 0;
    if (job.form == F_SERIES && job.gap == GAP_DTW && in.off_s && in.off_t) {
        s.b_finite.ensure(4);
        a.finite = s.b_finite.as<uint32_t>();
        for (int64_t k = 0; k < n; k++) {
            if (in.len_s[k] > 0) finite_nx = std::max<int64_t>(finite_nx, (in.off_s[k] + in.len_s[k]) * in.dim);
            if (in.len_t[k] > 0) finite_ny = std::max<int64_t>(finite_ny, (in.off_t[k] + in.len_t[k]) * in.dim);
        }
    }
    a.go_s = s.go_s;
    a.ge_s = s.ge_s;
    a.go_t = s.go_t;
    a.ge_t = s.ge_t;
    a.scores = scores;
    a.seeds = seeds;
    a.tb = job.tr != TR_SCORE ? s.b_tb.as<uint32_t>() : nullptr;
    a.dvals = plan.dvals > 0 ? s.b_dvals.as<float>() : nullptr;
    a.val_layers = val_layers;
    if (plan.band_slots > 0) {
        s.b_progress.ensure(size_t(plan.band_slots) * 4);
        s.b_partials.ensure(size_t(plan.band_slots) * sizeof(BandPartial));
        s.b_band_err.ensure(4);
        a.progress = s.b_progress.as<uint32_t>();
        a.partials = s.b_partials.as<BandPartial>();
        a.band_err = s.b_band_err.as<uint32_t>();
        a.band_sync = 1;
    }

    // ---- launches: specs first (occupancy queries stay outside any capture)
    std::vector<LaunchSpec> specs;
    std::vector<KArgs> largs;
    build_launches(s, job, in, plan, a, s.b_queue.as<uint32_t>(), specs, largs);
    const int n_band = int(plan.band_probs.size());
    LongArgs la{};
    int long_grid = 0;
    // the long kernel folds cells with max / min (and takes traceback codes
    // from equality tests) where that keeps the strict fold's bits:
    // similarities finite (or the pad's worst value) and never -0, and, for
    // maximising global problems, gap costs above zero (borders below -0), so
    // that no candidate is -0 or NaN
    bool fold_max = false;
    if (plan.n_long > 0) {
        auto exact = [](float x) { return std::isfinite(x) && !(x == 0.0f && std::signbit(x)); };
        bool good = true;
        if (job.form == F_BINARY) {
            good = exact(in.eq) && exact(in.ne);
        } else {
            const float w = worst(o.dir);
            for (float x : in.table)
                if (!exact(x) && x != w) {
                    good = false;
                    break;
                }
        }
        if (job.loc == LOC_GLOBAL && job.dir == DIR_MAX && !(s.go_s < 0.0f && s.go_t < 0.0f)) good = false;
        fold_max = good;
    }
    if (plan.n_long > 0) {
        la.idx = s.b_band_probs.as<int32_t>();
        la.n = plan.n_long;
        la.total_bands = plan.long_bands;
        la.ticket = s.b_queue.as<uint32_t>() + plan.launches.size();
        long_grid = std::max(1, std::min(plan.long_bands, s.cus));
    }
    // the whole sequence: queue tickets, band counters and long-mode
    // hand-off tags cleared by kernels (no memset nodes), the long-mode
    // launch (the critical path), the DP launches, the band fold
    auto enqueue = [&]() -> hipError_t {
        hipError_t e = launch_zero_u32(s.b_queue.as<uint32_t>(), int64_t(plan.launches.size()) + 1, st);
        if (e == hipSuccess && a.finite)
            e = launch_all_finite(static_cast<const float *>(codes_s), finite_nx, static_cast<const float *>(codes_t),
                                  finite_ny, s.b_finite.as<uint32_t>(), st);
        if (e == hipSuccess && plan.band_slots > 0) e = launch_zero_u32(s.b_progress.as<uint32_t>(), plan.band_slots, st);
        if (e == hipSuccess && plan.band_slots > 0) e = launch_zero_u32(s.b_band_err.as<uint32_t>(), 1, st);
        if (e == hipSuccess && plan.xbuf_end > plan.xbuf_begin)
            e = launch_zero_u32(reinterpret_cast<uint32_t *>(s.b_rowbuf.as<float2>() + plan.xbuf_begin),
                                2 * (plan.xbuf_end - plan.xbuf_begin), st);
        if (e == hipSuccess && plan.n_long > 0)
            e = launch_long(job.loc, job.dir, job.gap, job.form, job.tr, fold_max, a, la, long_grid, st);
        for (size_t q = 0; q < specs.size() && e == hipSuccess; q++) e = launch_solve(specs[q], largs[q], st);
        if (e == hipSuccess && n_band > 0)
            e = launch_band_combine(job.loc, job.dir, job.tr, s.b_probs.as<Prob>(), s.b_band_probs.as<int32_t>(),
                                    n_band, s.b_partials.as<BandPartial>(), scores, seeds, st);
        return e;
    };
    // A launch sequence repeated with identical arguments (same shapes, same
    // buffers) is captured once as a hipGraph and replayed; the kernels read
    // the plan and inputs uploaded just before, so a replay sees new data.
    std::string gkey(reinterpret_cast<const char *>(specs.data()), specs.size() * sizeof(LaunchSpec));
    gkey.append(reinterpret_cast<const char *>(largs.data()), largs.size() * sizeof(KArgs));
    const int64_t extra[10] = {int64_t(plan.launches.size()), plan.band_slots, n_band,
                               int64_t(reinterpret_cast<intptr_t>(scores)), int64_t(reinterpret_cast<intptr_t>(seeds)),
                               int64_t(job.loc * 100 + job.dir * 10 + job.tr), plan.n_long, plan.long_bands,
                               (plan.xbuf_end - plan.xbuf_begin) * 2 + int64_t(fold_max),
                               int64_t(reinterpret_cast<intptr_t>(s.b_band_probs.p))};
    gkey.append(reinterpret_cast<const char *>(extra), sizeof(extra));

    HIPCHK(hipEventRecord(s.ev0, st));
    hipError_t le = hipSuccess;
    // graphs are captured on the solver's own stream only (a caller's stream
    // may be the null stream, or under a capture of its own)
    const bool graphs = s.use_graphs && !s.external_stream;
    if (graphs && s.graph_exec && gkey == s.graph_key) {
        le = hipGraphLaunch(s.graph_exec, st);
    } else if (graphs && gkey == s.prev_key) {
        if (s.graph_exec) {
            (void)hipGraphExecDestroy(s.graph_exec);
            s.graph_exec = nullptr;
            s.graph_key.clear();
        }
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        le = enqueue();
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(st, &graph);
        if (le == hipSuccess) le = ce;
        if (le == hipSuccess) le = hipGraphInstantiate(&s.graph_exec, graph, nullptr, nullptr, 0);
        if (graph) (void)hipGraphDestroy(graph);
        if (le == hipSuccess) {
            s.graph_key = gkey;
            le = hipGraphLaunch(s.graph_exec, st);
        } else {
            s.graph_exec = nullptr;
        }
    } else {
        le = enqueue();
    }
    s.prev_key = gkey;
    if (le != hipSuccess) {
        std::ostringstream m;
        m << "kernel launch failed (loc " << job.loc << " dir " << job.dir << " gap " << job.gap << " form "
          << job.form << " tr " << job.tr << "): " << hipGetErrorString(le);
        fail(PA_EHIP, m.str());
    }
    const int launches = int(specs.size());
    if (n_band > 0) {
        HIPCHK(hipMemcpyAsync(s.band_err_host, s.b_band_err.p, 4, hipMemcpyDeviceToHost, st));
        s.band_check = true;
    }
    HIPCHK(hipEventRecord(s.ev1, st));
    s.last_launches = launches;
    s.last_cells = plan.cells;
    s.last_groups = int(plan.groups.size());
    s.last_band = int(plan.band_probs.size());
    s.last_ckpt = int(std::count_if(plan.probs.begin(), plan.probs.end(),
                                    [](const Prob &pr) { return pr.tb_off == NO_TB_OFF; }));

    if (tape) {
        tape->solver = &s;
        s.tapes.push_back(tape);
        tape->n = n;
        tape->n_probs = int64_t(plan.probs.size());
        tape->loc = job.loc;
        tape->gap = job.gap;
        tape->form = job.form;
        tape->dir = job.dir;
        tape->dim = in.dim;
        tape->dense = dense;
        tape->xs = static_cast<const float *>(codes_s);
        tape->xt = static_cast<const float *>(codes_t);
        std::swap(tape->probs.p, s.b_probs.p);
        std::swap(tape->probs.cap, s.b_probs.cap);
        std::swap(tape->seeds.p, s.b_seeds.p);
        std::swap(tape->seeds.cap, s.b_seeds.cap);
        std::swap(tape->tb.p, s.b_tb.p);
        std::swap(tape->tb.cap, s.b_tb.cap);
        tape->ckpt.clear();
        for (const Prob &pr : plan.probs)
            if (pr.tb_off == NO_TB_OFF) tape->ckpt.push_back(pr);
        if (!tape->ckpt.empty()) {
            std::swap(tape->rowbuf.p, s.b_rowbuf.p);
            std::swap(tape->rowbuf.cap, s.b_rowbuf.cap);
            tape->kargs = a;  // a.rowbuf now points into the tape's row buffer
            tape->band_slots = plan.band_slots;
            // the strip recomputation reads the borders of this forward's gap
            // costs: the tape keeps a copy (a recycled buffer), since the
            // solver's can be changed by pa_solver_set_gap_params before the
            // backward runs
            const size_t ns = s.border_s.size(), nt = s.border_t.size();
            std::swap(tape->borders.p, s.b_tape_borders.p);
            std::swap(tape->borders.cap, s.b_tape_borders.cap);
            tape->borders.ensure((ns + nt + 1) * 4);
            HIPCHK(hipMemcpyAsync(tape->borders.p, s.b_border_s.p, ns * 4, hipMemcpyDeviceToDevice, st));
            HIPCHK(hipMemcpyAsync(tape->borders.as<float>() + ns, s.b_border_t.p, nt * 4, hipMemcpyDeviceToDevice,
                                  st));
            tape->kargs.border_s = tape->borders.as<float>();
            tape->kargs.border_t = tape->borders.as<float>() + ns;
            if (a.finite) {  // and of this forward's series check (after the borders)
                uint32_t *f = reinterpret_cast<uint32_t *>(tape->borders.as<float>() + ns + nt);
                HIPCHK(hipMemcpyAsync(f, a.finite, 4, hipMemcpyDeviceToDevice, st));
                tape->kargs.finite = f;
            }
        }
        s.last_plan.uploaded = false;  // its descriptors moved to the tape
        return;
    }
    if (!result) {
        return;  // device scores: caller synchronises
    }
    auto r = std::make_unique<pa_result>();
    r->n = n;
    r->dir = o.dir;
    r->loc = job.loc;
    r->gap = job.gap;
    r->goal = o.goal;
    r->index_bits = o.index_bits;
    r->remove_dup = o.remove_dup;
    r->len_s.assign(in.len_s, in.len_s + n);
    r->len_t.assign(in.len_t, in.len_t + n);
    r->scores.resize(static_cast<size_t>(n));
    if (job.tr == TR_ONE) {
        r->path_off.resize(size_t(n) + 1);
        r->path_off[0] = 0;
        for (int64_t k = 0; k < n; k++) r->path_off[size_t(k) + 1] = r->path_off[size_t(k)] + in.len_s[k] + in.len_t[k] + 2;
        upload(st, s.b_path_off, r->path_off.data(), r->path_off.size());
        s.b_path.ensure(size_t(std::max<int64_t>(r->path_off.back(), 1)) * 8);
        s.b_path_len.ensure(size_t(std::max<int64_t>(n, 1)) * 4);
        HIPCHK(launch_walk(job.gap, s.b_probs.as<Prob>(), int64_t(plan.probs.size()), s.b_tb.as<uint32_t>(), seeds,
                           s.b_path.as<int32_t>(), s.b_path_off.as<int64_t>(), s.b_path_len.as<int32_t>(), st));
        r->path_len.resize(static_cast<size_t>(n));
        r->paths.resize(size_t(r->path_off.back()) * 2);
        HIPCHK(hipMemcpyAsync(r->path_len.data(), s.b_path_len.p, size_t(n) * 4, hipMemcpyDeviceToHost, st));
        if (!r->paths.empty())
            HIPCHK(hipMemcpyAsync(r->paths.data(), s.b_path.p, r->paths.size() * 4, hipMemcpyDeviceToHost, st));
    }
    if (job.tr != TR_SCORE) {
        r->seeds.resize(size_t(n) * 2);
        HIPCHK(hipMemcpyAsync(r->seeds.data(), seeds, size_t(n) * 8, hipMemcpyDeviceToHost, st));
    }
    r->matrices = in.matrices;
    r->layers = val_layers;
    r->tr = job.tr;
    if (in.matrices) {
        r->border_s.assign(s.border_s.begin(), s.border_s.end());
        r->border_t.assign(s.border_t.begin(), s.border_t.end());
    }
    if (job.tr == TR_ALL || in.matrices) {
        r->tb.resize(size_t(plan.tb_words));
        if (!r->tb.empty())
            HIPCHK(hipMemcpyAsync(r->tb.data(), s.b_tb.p, r->tb.size() * 4, hipMemcpyDeviceToHost, st));
        r->dvals.resize(size_t(plan.dvals));
        if (!r->dvals.empty())
            HIPCHK(hipMemcpyAsync(r->dvals.data(), s.b_dvals.p, r->dvals.size() * 4, hipMemcpyDeviceToHost, st));
        r->probs.resize(static_cast<size_t>(n));
        for (const Prob &pr : plan.probs) r->probs[size_t(pr.out)] = pr;
        if (job.tr == TR_ALL) r->iters.resize(static_cast<size_t>(n));
    }
    if (n) HIPCHK(hipMemcpyAsync(r->scores.data(), scores, size_t(n) * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    check_band(s);
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, s.ev0, s.ev1));
    s.last_ms = ms;
    *result = std::move(r);
}

// kernel table layout: (K+1) x (K+1), row = s symbol, last row/column = pad
// symbol with the worst value so padded rows never win
std::vector<float> make_table(const float *sim, int64_t K0, int64_t K1, int dir, bool *nan_free) {
    const int64_t K = std::max(K0, K1);
    const int64_t S = K + 1;
    std::vector<float> t(size_t(S * S), worst(dir));
    bool ok = true;
    for (int64_t i = 0; i < K0; i++)
        for (int64_t j = 0; j < K1; j++) {
            const float x = sim[i * K1 + j];
            if (std::isnan(x)) ok = false;
            t[size_t(i * S + j)] = x;
        }
    *nan_free = ok;
    return t;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

extern "C" {

const char *pa_last_error(void) { return g_err.c_str(); }

int pa_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int pa_create_options(const pa_options_desc *d, pa_options **out) {
    return guard([&] {
        if (!d || !out) fail(PA_EINVAL, "null argument");
        auto o = std::make_unique<pa_options>();
        o->d = *d;
        o->type = d->solver_type;
        o->dir = d->direction;
        o->remove_dup = !d->return_dup;
        o->index_bits = d->index_bits;
        if (o->type != PA_TYPE_ALIGNMENT && o->type != PA_TYPE_DTW) fail(PA_EINVAL, "illegal solver type");
        if (o->dir != PA_MINIMIZE && o->dir != PA_MAXIMIZE) fail(PA_EINVAL, "illegal direction");
        if (o->index_bits != 16 && o->index_bits != 32) fail(PA_EINVAL, "index bits must be 16 or 32");
        if (o->type == PA_TYPE_DTW) {
            // DTW solvers are always one-optimal (factory.h:1072-1099)
            o->loc = PA_GLOBAL;
            o->goal = PA_GOAL_ONE;
            o->gap = GAP_DTW;
            *out = o.release();
            return;
        }
        o->loc = d->locality;
        if (o->loc != PA_LOCAL && o->loc != PA_GLOBAL && o->loc != PA_SEMIGLOBAL) fail(PA_EINVAL, "invalid locality");
        if (d->count == PA_COUNT_ONE) {
            if (d->detail == PA_DETAIL_SCORE) o->goal = PA_GOAL_SCORE;
            else if (d->detail == PA_DETAIL_ALIGNMENT || d->detail == PA_DETAIL_SOLUTION) o->goal = PA_GOAL_ONE;
            else fail(PA_EINVAL, "invalid detail");
        } else if (d->count == PA_COUNT_ALL) {
            if (d->detail < PA_DETAIL_SCORE || d->detail > PA_DETAIL_SOLUTION) fail(PA_EINVAL, "invalid detail");
            o->goal = PA_GOAL_ALL;
        } else {
            fail(PA_EINVAL, "invalid count");
        }
        auto kind = [](int k) { return k == PA_GAP_NONE ? PA_GAP_LINEAR : k; };
        const int ks = kind(d->gap_kind_s), kt = kind(d->gap_kind_t);
        if (ks == PA_GAP_LINEAR && kt == PA_GAP_LINEAR) {
            o->gap = GAP_LIN;
            o->lin_s = d->gap_kind_s == PA_GAP_NONE ? 0.0f : float(d->gap_s[0]);
            o->lin_t = d->gap_kind_t == PA_GAP_NONE ? 0.0f : float(d->gap_t[0]);
            if (o->lin_s < 0 || o->lin_t < 0) fail(PA_EINVAL, "gap costs must be non-negative");
        } else if (ks == PA_GAP_AFFINE && kt == PA_GAP_AFFINE) {
            o->gap = GAP_AFF;
            // flip (u, v) of pyalign.gaps into AffineCost(u = v, v = u) (factory.h:40-52)
            o->aff_u_s = float(d->gap_s[1]);
            o->aff_v_s = float(d->gap_s[0]);
            o->aff_u_t = float(d->gap_t[1]);
            o->aff_v_t = float(d->gap_t[0]);
            if (o->aff_u_s < 0 || o->aff_v_s < 0 || o->aff_u_t < 0 || o->aff_v_t < 0)
                fail(PA_EINVAL, "gap costs must be non-negative");
        } else if (ks == PA_GAP_GENERAL && kt == PA_GAP_GENERAL) {
            // GeneralGapCostSolver (factory.h:1126-1133); the cost vectors come
            // with pa_solver_set_gap_costs
            o->gap = GAP_GEN;
        } else {
            // the reference dereferences an empty general cost here (factory.h:1126-1133)
            fail(PA_EINVAL,
                 "gap costs of different kinds for s and t: use two linear, two affine or two general costs");
        }
        *out = o.release();
    });
}

void pa_options_free(pa_options *o) { delete o; }

int pa_create_solver(int64_t max_len_s, int64_t max_len_t, const pa_options *opts, pa_solver **out) {
    return guard([&] {
        if (!opts || !out) fail(PA_EINVAL, "null argument");
        const int64_t imax = (opts->index_bits == 16 ? int64_t(INT16_MAX) : int64_t(INT32_MAX)) >> 1;
        auto check = [&](const char *name, int64_t len) {  // solver.h:496-503
            if (len > imax) {
                std::ostringstream e;
                e << "requested maximum length " << len << " for " << name
                  << " exceeds maximum supported sequence length in this implementation " << imax;
                fail(PA_ELENGTH, e.str());
            }
        };
        check("s", max_len_s);
        check("t", max_len_t);
        if (max_len_s < 0 || max_len_t < 0) fail(PA_EINVAL, "negative maximum length");
        auto s = std::make_unique<pa_solver>();
        s->opt = *opts;
        s->use_graphs = std::getenv("PYALIGN_AMD_NO_GRAPHS") == nullptr;
        s->lim_s = max_len_s;
        s->lim_t = max_len_t;
        make_borders(*s);
        const float sgn = opts->dir == PA_MINIMIZE ? 1.0f : -1.0f;
        if (opts->gap == GAP_LIN) {
            s->go_s = opts->lin_s * sgn;
            s->go_t = opts->lin_t * sgn;
        } else if (opts->gap == GAP_AFF) {
            s->go_s = (opts->aff_u_s + opts->aff_v_s) * sgn;  // w1 = u + v (solver.h:2872-2874)
            s->ge_s = opts->aff_u_s * sgn;
            s->go_t = (opts->aff_u_t + opts->aff_v_t) * sgn;
            s->ge_t = opts->aff_u_t * sgn;
        }
        *out = s.release();
    });
}

void pa_solver_free(pa_solver *s) { delete s; }

int pa_solver_batch_size(const pa_solver *s) {
    // batch=True solvers take one wavefront of lanes per call; unbatched take 1
    return s && s->opt.d.batch ? 64 : 1;
}

static void check_batch(const pa_solver *s, int64_t B) {  // factory.h:387-394
    const int bs = pa_solver_batch_size(s);
    if (B != bs) {
        std::ostringstream e;
        e << "problem of batch size " << B << " given to solver that was configured to batch size " << bs;
        fail(PA_EINVAL, e.str());
    }
}

static void lanes_from_len(const int32_t *len, int64_t B, int64_t Ls, int64_t Lt, std::vector<int32_t> &ls,
                           std::vector<int32_t> &lt) {
    ls.assign(len, len + B);
    lt.assign(len + B, len + 2 * B);
    for (int64_t k = 0; k < B; k++)
        if (ls[size_t(k)] > Ls || lt[size_t(k)] > Lt || ls[size_t(k)] < 0 || lt[size_t(k)] < 0)
            fail(PA_EINVAL, "problem length exceeds the batch's matrix shape");
}

static void check_master(const pa_solver *s, int64_t Ls, int64_t Lt) {  // solver.h:725-726
    if (Ls > s->lim_s) {
        std::ostringstream e;
        e << "sequence s of length " << Ls << " exceeds configured maximum length " << s->lim_s;
        fail(PA_ELENGTH, e.str());
    }
    if (Lt > s->lim_t) {
        std::ostringstream e;
        e << "sequence t of length " << Lt << " exceeds configured maximum length " << s->lim_t;
        fail(PA_ELENGTH, e.str());
    }
}

int pa_solve_dense(pa_solver *s, const float *sim, int64_t Ls, int64_t Lt, int64_t B, const int32_t *len,
                   pa_result **out) {
    return guard([&] {
        if (!s || !len || !out || (!sim && Ls * Lt * B > 0)) fail(PA_EINVAL, "null argument");
        check_batch(s, B);
        check_master(s, Ls, Lt);
        std::vector<int32_t> ls, lt;
        lanes_from_len(len, B, Ls, Lt, ls, lt);
        Input in{};
        in.form = PA_FORM_DENSE;
        in.n = B;
        in.len_s = ls.data();
        in.len_t = lt.data();
        in.dense = sim;
        in.dense_elems = Ls * Lt * B;
        in.dense_si_fixed = int32_t(Lt * B);  // sim[i, j, k] at (i*Lt + j)*B + k
        in.dense_sj_fixed = int32_t(B);
        std::unique_ptr<pa_result> r;
        run(*s, in, &r, nullptr);
        *out = r.release();
    });
}

// Table sizes the indexed form serves: 1 .. 32,767 symbols per side (codes
// times the padded table stride stay within int32 in the kernels).
static void check_alphabet(int64_t K0, int64_t K1) {
    if (K0 < 1 || K1 < 1) fail(PA_EINVAL, "the similarity table needs at least one symbol per side");
    if (K0 > 32767 || K1 > 32767) fail(PA_EINVAL, "alphabets above 32,767 symbols are not supported");
}

static void indexed_input(Input &in, const uint32_t *a, int64_t La, const uint32_t *b, int64_t Lb, int64_t B,
                          const float *sim, int64_t K0, int64_t K1, int dir, std::vector<uint8_t> &cs,
                          std::vector<uint8_t> &ct, std::vector<int64_t> &os, std::vector<int64_t> &ot) {
    uint32_t amax = 0, bmax = 0;
    for (int64_t k = 0; k < B * La; k++) amax = std::max(amax, a[k]);
    for (int64_t k = 0; k < B * Lb; k++) bmax = std::max(bmax, b[k]);
    // indexed_matrix_form::check (factory.h:602-614)
    if (B * La > 0 && amax >= uint64_t(K0)) fail(PA_EINVAL, "out of bounds index in a");
    if (B * Lb > 0 && bmax >= uint64_t(K1)) fail(PA_EINVAL, "out of bounds index in b");
    check_alphabet(K0, K1);
    // up to 128 symbols: uint8 codes, table staged in LDS; beyond: uint16
    // codes, table read from global memory
    in.table_g = std::max(K0, K1) + 1 > MAX_LDS_TABLE;
    const size_t cw = in.table_g ? 2 : 1;
    cs.resize((size_t(B * La) + 1) * cw);
    ct.resize((size_t(B * Lb) + 1) * cw);
    for (int64_t k = 0; k < B * La; k++) {
        if (cw == 2) reinterpret_cast<uint16_t *>(cs.data())[k] = uint16_t(a[k]);
        else cs[size_t(k)] = uint8_t(a[k]);
    }
    for (int64_t k = 0; k < B * Lb; k++) {
        if (cw == 2) reinterpret_cast<uint16_t *>(ct.data())[k] = uint16_t(b[k]);
        else ct[size_t(k)] = uint8_t(b[k]);
    }
    os.resize(static_cast<size_t>(B));
    ot.resize(static_cast<size_t>(B));
    for (int64_t k = 0; k < B; k++) {
        os[size_t(k)] = k * La;
        ot[size_t(k)] = k * Lb;
    }
    in.form = PA_FORM_INDEXED;
    in.codes_s = cs.data();
    in.codes_t = ct.data();
    in.bytes_s = int64_t(cs.size());
    in.bytes_t = int64_t(ct.size());
    in.off_s = os.data();
    in.off_t = ot.data();
    bool nan_free = true;
    in.table = make_table(sim, K0, K1, dir, &nan_free);
    in.nan_free = nan_free;
    in.K1 = int(std::max(K0, K1) + 1);
}

int pa_solve_indexed(pa_solver *s, const uint32_t *a, int64_t La, const uint32_t *b, int64_t Lb, int64_t B,
                     const float *sim, int64_t K0, int64_t K1, const int32_t *len, pa_result **out) {
    return guard([&] {
        if (!s || !len || !out || !sim) fail(PA_EINVAL, "null argument");
        check_batch(s, B);
        check_master(s, La, Lb);
        std::vector<int32_t> ls, lt;
        lanes_from_len(len, B, La, Lb, ls, lt);
        Input in{};
        std::vector<uint8_t> cs, ct;
        std::vector<int64_t> os, ot;
        indexed_input(in, a, La, b, Lb, B, sim, K0, K1, s->opt.dir, cs, ct, os, ot);
        in.n = B;
        in.len_s = ls.data();
        in.len_t = lt.data();
        std::unique_ptr<pa_result> r;
        run(*s, in, &r, nullptr);
        *out = r.release();
    });
}

int pa_solve_binary(pa_solver *s, const uint32_t *a, int64_t La, const uint32_t *b, int64_t Lb, int64_t B,
                    float eq, float ne, const int32_t *len, pa_result **out) {
    return guard([&] {
        if (!s || !len || !out) fail(PA_EINVAL, "null argument");
        check_batch(s, B);
        check_master(s, La, Lb);
        std::vector<int32_t> ls, lt;
        lanes_from_len(len, B, La, Lb, ls, lt);
        std::vector<int64_t> os(static_cast<size_t>(B)), ot(static_cast<size_t>(B));
        for (int64_t k = 0; k < B; k++) {
            os[size_t(k)] = k * La;
            ot[size_t(k)] = k * Lb;
        }
        static const uint32_t zero = 0;
        Input in{};
        in.form = PA_FORM_BINARY;
        in.n = B;
        in.len_s = ls.data();
        in.len_t = lt.data();
        in.codes_s = B * La ? static_cast<const void *>(a) : &zero;
        in.codes_t = B * Lb ? static_cast<const void *>(b) : &zero;
        in.bytes_s = std::max<int64_t>(B * La, 1) * 4;
        in.bytes_t = std::max<int64_t>(B * Lb, 1) * 4;
        in.off_s = os.data();
        in.off_t = ot.data();
        in.eq = eq;
        in.ne = ne;
        std::unique_ptr<pa_result> r;
        run(*s, in, &r, nullptr);
        *out = r.release();
    });
}

static void ragged_input(const pa_solver *s, const pa_ragged *b, Input &in) {
    if (!b) fail(PA_EINVAL, "null batch");
    in.n = b->n;
    in.len_s = b->len_s;
    in.len_t = b->len_t;
    in.on_device = b->on_device != 0;
    in.off_s = b->off_s;
    in.off_t = b->off_t;
    in.form = b->form;
    int64_t end_s = 0, end_t = 0;
    if (b->form != PA_FORM_DENSE) {
        if (!b->off_s || !b->off_t) fail(PA_EINVAL, "null offsets");
        for (int64_t k = 0; k < b->n; k++) {
            end_s = std::max<int64_t>(end_s, b->off_s[k] + b->len_s[k]);
            end_t = std::max<int64_t>(end_t, b->off_t[k] + b->len_t[k]);
        }
    }
    switch (b->form) {
        case PA_FORM_INDEXED: {
            const int64_t K0 = b->K, K1 = b->K_t > 0 ? b->K_t : b->K;  // table rows, columns
            check_alphabet(K0, K1);
            if (!b->table) fail(PA_EINVAL, "null table");
            const int cb = b->code_bytes == 0 ? 1 : b->code_bytes;
            if (cb != 1 && cb != 2) fail(PA_EINVAL, "indexed codes are 1 or 2 bytes wide");
            // up to 128 symbols and uint8 codes: the table is staged in LDS;
            // otherwise read from global memory with uint16 codes
            in.table_g = std::max(K0, K1) + 1 > MAX_LDS_TABLE || cb == 2;
            if (b->raw_width) {
                // codes come from the device translation; its table must map
                // into the alphabet
                if (!b->lut || b->lut_len < 1) fail(PA_EINVAL, "raw symbols need a code table");
                for (int32_t q = 0; q < b->lut_len; q++)
                    if (b->lut[q] != 255 && (b->lut[q] >= K0 || b->lut[q] >= K1))
                        fail(PA_EINVAL, "code table entry beyond the alphabet");
                in.lut = b->lut;
                in.lut_len = b->lut_len;
            } else if (b->on_device) {
                if (in.table_g && cb != 2)
                    fail(PA_EINVAL, "device-resident codes for a table above 128 symbols must be uint16 (code_bytes 2)");
                in.bytes_s = end_s * cb;
                in.bytes_t = end_t * cb;
            } else if (cb == 1) {
                // host uint8 codes go through the device encoder with identity
                // tables (s codes below the table's rows, t codes below its
                // columns), which flags an out-of-range code before any DP
                // launch and widens the codes where the table stays in global
                // memory
                in.lut_store.assign(256, 255);
                in.lut_t_store.assign(256, 255);
                for (int64_t c = 0; c < K0 && c < 255; c++) in.lut_store[size_t(c)] = uint8_t(c);
                for (int64_t c = 0; c < K1 && c < 255; c++) in.lut_t_store[size_t(c)] = uint8_t(c);
                in.lut = in.lut_store.data();
                in.lut_len = 256;
                in.lut_t = in.lut_t_store.data();
                in.lut_t_len = 256;
                in.raw_width = 1;
                in.syms_s = end_s;
                in.syms_t = end_t;
                in.bad_s = "out of bounds index in a";
                in.bad_t = "out of bounds index in b";
            } else {
                // host uint16 codes: checked here, per problem, then uploaded
                const uint16_t *cs = static_cast<const uint16_t *>(b->codes_s);
                const uint16_t *ct = static_cast<const uint16_t *>(b->codes_t);
                for (int64_t k = 0; k < b->n; k++) {
                    for (int64_t q = 0; q < b->len_s[k]; q++)
                        if (cs[b->off_s[k] + q] >= K0) fail(PA_EINVAL, "out of bounds index in a");
                    for (int64_t q = 0; q < b->len_t[k]; q++)
                        if (ct[b->off_t[k] + q] >= K1) fail(PA_EINVAL, "out of bounds index in b");
                }
                in.bytes_s = end_s * 2;
                in.bytes_t = end_t * 2;
            }
            if (!b->raw_width && !b->on_device && cb == 1) {
                in.bytes_s = end_s;
                in.bytes_t = end_t;
            }
            bool nan_free = true;
            in.table = make_table(b->table, K0, K1, s->opt.dir, &nan_free);
            in.nan_free = nan_free;
            in.K1 = int(std::max(K0, K1) + 1);
            break;
        }
        case PA_FORM_BINARY:
            if (b->lut) fail(PA_EINVAL, "the binary form takes no code table");
            in.bytes_s = end_s * 4;
            in.bytes_t = end_t * 4;
            in.eq = b->eq;
            in.ne = b->ne;
            break;
        case PA_FORM_SERIES:
            in.dim = b->dim;
            in.bytes_s = end_s * 4 * b->dim;
            in.bytes_t = end_t * 4 * b->dim;
            break;
        case PA_FORM_DENSE: {
            if (!b->dense_off) fail(PA_EINVAL, "null dense offsets");
            int64_t end = 0;
            for (int64_t k = 0; k < b->n; k++) end = std::max<int64_t>(end, b->dense_off[k] + int64_t(b->len_s[k]) * b->len_t[k]);
            in.dense = b->dense;
            in.dense_elems = end;
            in.dense_off = b->dense_off;
            if (b->dense_ld > 0) {
                for (int64_t k = 0; k < b->n; k++) {
                    if (b->len_t[k] > b->dense_ld) fail(PA_EINVAL, "dense row stride below len_t");
                    end = std::max<int64_t>(end, b->dense_off[k] + int64_t(b->len_s[k]) * b->dense_ld);
                }
                in.dense_elems = end;
                in.dense_si_fixed = int32_t(b->dense_ld);
                in.dense_sj_fixed = 1;
            }
            break;
        }
        default: fail(PA_EINVAL, "unknown problem form");
    }
    if (b->raw_width) {
        if (b->raw_width != 1 && b->raw_width != 4) fail(PA_EINVAL, "raw symbols are 1 or 4 bytes wide");
        if (b->form == PA_FORM_INDEXED && b->code_bytes == 2) fail(PA_EINVAL, "raw symbols take no code width");
        if (b->form != PA_FORM_INDEXED && b->form != PA_FORM_BINARY)
            fail(PA_EINVAL, "raw symbols are for the indexed and binary forms");
        in.raw_width = b->raw_width;
        in.syms_s = end_s;
        in.syms_t = end_t;
        in.bytes_s = end_s * b->raw_width;
        in.bytes_t = end_t * b->raw_width;
    }
    in.codes_s = b->codes_s;
    in.codes_t = b->codes_t;
    static const uint32_t zero = 0;
    if (!in.codes_s) in.codes_s = &zero;
    if (!in.codes_t) in.codes_t = &zero;
    in.bytes_s = std::max<int64_t>(in.bytes_s, 1);
    in.bytes_t = std::max<int64_t>(in.bytes_t, 1);
}

int pa_solve_ragged(pa_solver *s, const pa_ragged *batch, int32_t schedule, pa_result **out) {
    return guard([&] {
        if (!s || !out) fail(PA_EINVAL, "null argument");
        Input in{};
        ragged_input(s, batch, in);
        in.schedule = schedule;
        std::unique_ptr<pa_result> r;
        run(*s, in, &r, nullptr);
        *out = r.release();
    });
}

int pa_solve_ragged_scores_device(pa_solver *s, const pa_ragged *batch, int32_t schedule, float *scores_device) {
    return guard([&] {
        if (!s || !scores_device) fail(PA_EINVAL, "null argument");
        if (s->opt.goal != PA_GOAL_SCORE && s->opt.type != PA_TYPE_DTW)
            fail(PA_EINVAL, "device score output needs a score-only solver");
        Input in{};
        ragged_input(s, batch, in);
        in.schedule = schedule;
        run(*s, in, nullptr, scores_device);
    });
}

int pa_solve_ragged_scores_noalloc(pa_solver *s, const pa_ragged *batch, int32_t schedule, float *scores_device,
                                   void *stream) {
    return guard([&] {
        if (!s || !batch || !scores_device) fail(PA_EINVAL, "null argument");
        const pa_options &o = s->opt;
        if (o.goal != PA_GOAL_SCORE && o.type != PA_TYPE_DTW)
            fail(PA_EINVAL, "device score output needs a score-only solver");
        if (o.gap == GAP_GEN) fail(PA_EINVAL, "general gap costs need the allocating entry points");
        if (!batch->on_device || batch->raw_width)
            fail(PA_EINVAL, "the allocation-free solve takes device-resident codes, series or similarities");
        Input in{};
        ragged_input(s, batch, in);
        in.schedule = schedule;
        const int64_t n = in.n;
        for (int64_t k = 0; k < n; k++) {
            if (in.len_s[k] < 0 || in.len_t[k] < 0) fail(PA_EINVAL, "negative sequence length");
            if (in.len_s[k] > s->lim_s || in.len_t[k] > s->lim_t) {
                std::ostringstream e;
                e << "sequence " << (in.len_s[k] > s->lim_s ? "s" : "t") << " of length "
                  << (in.len_s[k] > s->lim_s ? in.len_s[k] : in.len_t[k]) << " exceeds configured maximum length "
                  << (in.len_s[k] > s->lim_s ? s->lim_s : s->lim_t);
                fail(PA_ELENGTH, e.str());
            }
        }
        Job job;
        job.loc = o.type == PA_TYPE_DTW ? LOC_GLOBAL : o.loc;
        job.dir = o.dir == PA_MAXIMIZE ? DIR_MAX : DIR_MIN;
        job.gap = o.gap;
        job.tr = TR_SCORE;
        switch (in.form) {
            case PA_FORM_DENSE: job.form = F_DENSE; break;
            case PA_FORM_INDEXED: job.form = in.table_g ? F_TABLE_G : F_TABLE; break;
            case PA_FORM_BINARY: job.form = F_BINARY; break;
            case PA_FORM_SERIES: job.form = F_SERIES; break;
            default: fail(PA_EINVAL, "unknown problem form");
        }
        if (job.form == F_SERIES && o.type != PA_TYPE_DTW) fail(PA_EINVAL, "the series form is for DTW solvers");
        if (o.type == PA_TYPE_DTW && job.form != F_DENSE && job.form != F_SERIES)
            fail(PA_EINVAL, "DTW solvers take dense costs or series");
        if (job.form == F_SERIES && (in.dim < 1 || in.dim > 4)) fail(PA_EINVAL, "series dimension must be 1..4");
        job.fast = job.form == F_TABLE && job.loc == LOC_LOCAL && job.dir == DIR_MAX && o.gap != GAP_DTW &&
                   in.nan_free;

        void *base = nullptr;
        int cus = 0;
        HIPCHK(noalloc_scratch(&base, &cus));
        if (!s->has_device) s->cus = cus;
        // no band mode: its strips hand rows down through progress counters
        // whose failure flag the host would have to read back
        Plan plan = in.schedule == PA_SCHED_PLAIN ? plan_plain(job, in.len_s, in.len_t, n)
                                                  : plan_auto(job, in.len_s, in.len_t, n, cus * 4, nullptr, false);
        place_plan(plan, in, 0, false, 1);

        // the scratch layout: safe values, sink, queue tickets, borders,
        // table, problem and group descriptors, strip-boundary row buffer
        size_t at = 0;
        auto take = [&](size_t bytes) {
            const size_t here = at;
            at = (at + bytes + 255) / 256 * 256;
            return here;
        };
        const size_t o_safe = take(SAFE_FLOATS * 4);
        const size_t o_sink = take(SINK_WORDS * 4);
        const size_t o_queue = take(std::max<size_t>(plan.launches.size(), 1) * 4);
        const size_t o_bs = take(s->border_s.size() * 4);
        const size_t o_bt = take(s->border_t.size() * 4);
        const size_t o_table = take(in.table.size() * 4);
        const size_t o_probs = take(plan.probs.size() * sizeof(Prob));
        const size_t o_groups = take(plan.groups.size() * sizeof(Group));
        const size_t o_rowbuf = take(size_t(plan.rowbuf) * sizeof(float2));
        if (at > NOALLOC_SCRATCH_BYTES) {
            std::ostringstream e;
            e << "the allocation-free solve needs " << at << " bytes of device scratch for this batch; it holds "
              << NOALLOC_SCRATCH_BYTES << " (use pa_solve_ragged_scores_device, or smaller batches)";
            fail(PA_EINVAL, e.str());
        }
        char *dev = static_cast<char *>(base);
        hipStream_t st = static_cast<hipStream_t>(stream);
        float safe[SAFE_FLOATS] = {};
        safe[0] = -std::numeric_limits<float>::infinity();
        safe[1] = std::numeric_limits<float>::infinity();
        // pageable sources: each copy has read its source when the call returns
        auto put = [&](size_t off, const void *src, size_t bytes) {
            if (bytes) HIPCHK(hipMemcpyAsync(dev + off, src, bytes, hipMemcpyHostToDevice, st));
        };
        put(o_safe, safe, sizeof(safe));
        put(o_bs, s->border_s.data(), s->border_s.size() * 4);
        put(o_bt, s->border_t.data(), s->border_t.size() * 4);
        put(o_table, in.table.data(), in.table.size() * 4);
        put(o_probs, plan.probs.data(), plan.probs.size() * sizeof(Prob));
        put(o_groups, plan.groups.data(), plan.groups.size() * sizeof(Group));

        KArgs a;
        std::memset(&a, 0, sizeof(a));
        a.probs = reinterpret_cast<const Prob *>(dev + o_probs);
        a.groups = reinterpret_cast<const Group *>(dev + o_groups);
        a.codes_s = in.codes_s;
        a.codes_t = in.codes_t;
        a.table = reinterpret_cast<const float *>(dev + o_table);
        a.K1 = in.K1;
        a.dim = in.dim;
        a.eq = in.eq;
        a.ne = in.ne;
        a.dense = in.dense;
        a.rowbuf = reinterpret_cast<float2 *>(dev + o_rowbuf);
        a.border_s = reinterpret_cast<const float *>(dev + o_bs);
        a.border_t = reinterpret_cast<const float *>(dev + o_bt);
        a.safe = reinterpret_cast<const float *>(dev + o_safe);
        a.sink = reinterpret_cast<uint32_t *>(dev + o_sink);
        a.go_s = s->go_s;
        a.ge_s = s->ge_s;
        a.go_t = s->go_t;
        a.ge_t = s->ge_t;
        a.scores = scores_device;
        a.val_layers = 1;
        uint32_t *queue = reinterpret_cast<uint32_t *>(dev + o_queue);
        std::vector<LaunchSpec> specs;
        std::vector<KArgs> largs;
        build_launches(*s, job, in, plan, a, queue, specs, largs);
        // the queue tickets start at zero (a fill kernel: the last call's
        // counts are still there)
        hipError_t e = launch_zero_u32(queue, int64_t(std::max<size_t>(plan.launches.size(), 1)), st);
        for (size_t q = 0; q < specs.size() && e == hipSuccess; q++) e = launch_solve(specs[q], largs[q], st);
        if (e != hipSuccess) {
            std::ostringstream m;
            m << "kernel launch failed (allocation-free solve): " << hipGetErrorString(e);
            fail(PA_EHIP, m.str());
        }
    });
}

int pa_solver_sync(pa_solver *s) {
    return guard([&] {
        if (!s) fail(PA_EINVAL, "null solver");
        if (!s->stream) return;
        HIPCHK(hipStreamSynchronize(s->stream));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        s->last_ms = ms;
        check_band(*s);
    });
}

int pa_solver_stream(const pa_solver *s, void **stream) {
    return guard([&] {
        if (!s || !stream) fail(PA_EINVAL, "null argument");
        *stream = reinterpret_cast<void *>(s->stream);
    });
}

int pa_solver_last_kernel_ms(const pa_solver *s, double *ms, int32_t *launches, double *cells) {
    return guard([&] {
        if (!s) fail(PA_EINVAL, "null solver");
        if (ms) *ms = s->last_ms;
        if (launches) *launches = s->last_launches;
        if (cells) *cells = s->last_cells;
    });
}

int pa_solver_set_gap_costs(pa_solver *s, const float *w_s, int64_t n_s, const float *w_t, int64_t n_t) {
    return guard([&] {
        if (!s || !w_s || !w_t) fail(PA_EINVAL, "null argument");
        if (s->opt.gap != GAP_GEN) fail(PA_EINVAL, "the solver's gap costs are not general");
        auto check_len = [](int64_t got, int64_t want) {  // check_gap_tensor_shape, solver.h:3041-3045
            if (got != want) {
                std::ostringstream e;
                e << "expected gap cost tensor length of " << want << ", got " << got;
                fail(PA_EINVAL, e.str());
            }
        };
        check_len(n_s, s->lim_s + 1);
        check_len(n_t, s->lim_t + 1);
        const float sgn = s->opt.dir == PA_MINIMIZE ? 1.0f : -1.0f;
        s->ws_sgn.assign(size_t(n_s), 0.0f);
        s->wt_sgn.assign(size_t(n_t), 0.0f);
        for (int64_t k = 0; k < n_s; k++) s->ws_sgn[size_t(k)] = w_s[k] * sgn;
        for (int64_t k = 0; k < n_t; k++) s->wt_sgn[size_t(k)] = w_t[k] * sgn;
        s->gen_ready = true;
        make_borders(*s);
        if (s->has_device) {
            HIPCHK(hipMemcpy(s->b_border_s.p, s->border_s.data(), s->border_s.size() * 4, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(s->b_border_t.p, s->border_t.data(), s->border_t.size() * 4, hipMemcpyHostToDevice));
        }
        s->last_plan.valid = false;
    });
}

int pa_solver_set_gap_params(pa_solver *s, const double *gap_s, const double *gap_t) {
    return guard([&] {
        if (!s || !gap_s || !gap_t) fail(PA_EINVAL, "null argument");
        pa_options &o = s->opt;
        if (o.type != PA_TYPE_ALIGNMENT || (o.gap != GAP_LIN && o.gap != GAP_AFF))
            fail(PA_EINVAL, "the solver's gap costs are not linear or affine");
        const float sgn = o.dir == PA_MINIMIZE ? 1.0f : -1.0f;
        // the same float32 conversions as pa_create_options / pa_create_solver
        if (o.gap == GAP_LIN) {
            const float ls = float(gap_s[0]), lt = float(gap_t[0]);
            if (ls < 0 || lt < 0) fail(PA_EINVAL, "gap costs must be non-negative");
            if (ls == o.lin_s && lt == o.lin_t) return;
            o.lin_s = ls;
            o.lin_t = lt;
            s->go_s = ls * sgn;
            s->go_t = lt * sgn;
        } else {
            const float us = float(gap_s[1]), vs = float(gap_s[0]), ut = float(gap_t[1]), vt = float(gap_t[0]);
            if (us < 0 || vs < 0 || ut < 0 || vt < 0) fail(PA_EINVAL, "gap costs must be non-negative");
            if (us == o.aff_u_s && vs == o.aff_v_s && ut == o.aff_u_t && vt == o.aff_v_t) return;
            o.aff_u_s = us;
            o.aff_v_s = vs;
            o.aff_u_t = ut;
            o.aff_v_t = vt;
            s->go_s = (us + vs) * sgn;
            s->ge_s = us * sgn;
            s->go_t = (ut + vt) * sgn;
            s->ge_t = ut * sgn;
        }
        make_borders(*s);
        if (!s->has_device) return;
        // new borders on the solver's stream: after the work already queued
        // there (which read the old ones), before the next solve's kernels;
        // staged through pinned memory so the host does not wait
        const size_t ns = s->border_s.size(), nt = s->border_t.size();
        if (s->border_pending) HIPCHK(hipEventSynchronize(s->border_ev));
        s->border_pending = false;
        if (s->border_stage_cap < ns + nt) {
            if (s->border_stage) HIPCHK(hipHostFree(s->border_stage));
            s->border_stage = nullptr;
            s->border_stage_cap = 0;
            HIPCHK(hipHostMalloc(reinterpret_cast<void **>(&s->border_stage), (ns + nt) * 4, hipHostMallocDefault));
            s->border_stage_cap = ns + nt;
        }
        if (!s->border_ev) HIPCHK(hipEventCreateWithFlags(&s->border_ev, hipEventDisableTiming));
        std::memcpy(s->border_stage, s->border_s.data(), ns * 4);
        std::memcpy(s->border_stage + ns, s->border_t.data(), nt * 4);
        HIPCHK(hipMemcpyAsync(s->b_border_s.p, s->border_stage, ns * 4, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipMemcpyAsync(s->b_border_t.p, s->border_stage + ns, nt * 4, hipMemcpyHostToDevice, s->stream));
        HIPCHK(hipEventRecord(s->border_ev, s->stream));
        s->border_pending = true;
    });
}

int pa_solver_last_bad_symbol(const pa_solver *s, int64_t *pos_s, int64_t *pos_t) {
    return guard([&] {
        if (!s || !pos_s || !pos_t) fail(PA_EINVAL, "null argument");
        *pos_s = s->last_bad[0];
        *pos_t = s->last_bad[1];
    });
}

int pa_solver_set_stream(pa_solver *s, void *stream, int32_t external) {
    return guard([&] {
        if (!s) fail(PA_EINVAL, "null solver");
        ensure_device(*s);
        s->external_stream = external != 0;
        s->stream = s->external_stream ? static_cast<hipStream_t>(stream) : s->own_stream;
    });
}

int pa_solver_set_checkpoint(pa_solver *s, int64_t min_len) {
    return guard([&] {
        if (!s) fail(PA_EINVAL, "null solver");
        s->ckpt_min = min_len;
    });
}

int pa_solver_set_graphs(pa_solver *s, int32_t enable) {
    return guard([&] {
        if (!s) fail(PA_EINVAL, "null solver");
        s->use_graphs = enable != 0;
        s->prev_key.clear();
    });
}

int pa_solver_last_schedule(const pa_solver *s, int32_t *groups, int32_t *band_problems, int32_t *checkpointed) {
    return guard([&] {
        if (!s) fail(PA_EINVAL, "null solver");
        if (groups) *groups = s->last_groups;
        if (band_problems) *band_problems = s->last_band;
        if (checkpointed) *checkpointed = s->last_ckpt;
    });
}

// ---- differentiation -------------------------------------------------------

int pa_forward_ragged(pa_solver *s, const pa_ragged *batch, float *scores_device, pa_tape **tape) {
    return guard([&] {
        if (!s || !batch || !scores_device || !tape) fail(PA_EINVAL, "null argument");
        if ((batch->form == PA_FORM_DENSE || batch->form == PA_FORM_SERIES) && !batch->on_device)
            fail(PA_EINVAL, "differentiation needs the dense or series payload resident on the device");
        if (s->opt.type == PA_TYPE_ALIGNMENT && s->opt.gap != GAP_LIN && s->opt.gap != GAP_AFF)
            fail(PA_EINVAL, "unsupported gap cost for differentiation");
        Input in{};
        ragged_input(s, batch, in);
        auto t = std::make_unique<pa_tape>();
        run(*s, in, nullptr, scores_device, t.get());
        *tape = t.release();
    });
}

}  // extern "C"

namespace {

// Backward over the checkpointed problems of a tape.  Their forward kept
// only the bottom row of every strip (in the tape's row buffer), so each
// strip's traceback can be recomputed on its own from the row above (the
// band kernel with no waits: every row above is complete).  Problems whose
// whole recomputed traceback fits the scratch budget are done in chunks:
// every strip of the chunk recomputed in one launch, then the ordinary
// backward walk over them.  A problem too large for that walks one strip per
// pass, from the strip of its seed (the path's end the forward found) up:
// each pass recomputes that strip into the scratch buffer and
// pa_backward_strip_kernel continues the walk through it.  The
// scratch buffer belongs to the solver and is reused by its tapes.
void backward_checkpointed(pa_tape &t, const BwArgs &base, int form, hipStream_t st) {
    pa_solver &s = *t.solver;
    const int bits = tb_bits(t.gap, TR_ONE);
    // scratch traceback words: a quarter of the free device memory (plus
    // what the solver's scratch already holds), at least 256 MiB
    size_t free_b = 0, total_b = 0;
    HIPCHK(hipMemGetInfo(&free_b, &total_b));
    const int64_t budget = s.ckpt_words > 0 ? s.ckpt_words
                                            : std::max<int64_t>(int64_t(64) << 20,
                                                                int64_t((free_b / 4 + s.b_ckpt_tb.cap) / 4));
    LaunchSpec spec;
    std::memset(&spec, 0, sizeof(spec));
    spec.loc = t.loc;
    spec.dir = t.dir;
    spec.gap = t.gap;
    spec.form = t.form;
    spec.tr = TR_ONE;
    spec.R = BAND_R;
    const int waves_per_cu = std::max(1, occupancy_waves(spec, t.dim));
    t.s_progress.ensure(size_t(std::max<int64_t>(t.band_slots, 1)) * 4);
    t.s_partials.ensure(size_t(std::max<int64_t>(t.band_slots, 1)) * sizeof(BandPartial));
    t.s_err.ensure(4);
    t.s_queue.ensure(4);
    HIPCHK(launch_fill_u32(t.s_progress.as<uint32_t>(), t.band_slots, 0xffffffffu, st));
    HIPCHK(launch_zero_u32(t.s_err.as<uint32_t>(), 1, st));
    auto words_of = [&](const Prob &pr) {
        return (int64_t(pr.len_t) + pr.L - 1) * pr.L * ((bits * pr.R + 31) / 32);
    };
    auto strips_of = [&](const Prob &pr) { return (pr.len_s + pr.L * pr.R - 1) / (pr.L * pr.R); };
    t.s_scores.ensure(size_t(std::max<int64_t>(t.n, 1)) * 4);
    std::vector<Prob> whole, by_strip;
    for (const Prob &pr : t.ckpt) (words_of(pr) * strips_of(pr) <= budget ? whole : by_strip).push_back(pr);
    // longest walks first: the walk kernel takes problems in this order, a
    // wavefront each, more of them than fit the GPU at once
    std::stable_sort(whole.begin(), whole.end(), [](const Prob &x, const Prob &y) {
        return int64_t(x.len_s) + x.len_t > int64_t(y.len_s) + y.len_t;
    });
    auto recompute = [&](const std::vector<Prob> &probs, const std::vector<Group> &groups) {
        upload(st, t.s_probs, probs.data(), probs.size());
        upload(st, t.s_groups, groups.data(), groups.size());
        HIPCHK(launch_zero_u32(t.s_queue.as<uint32_t>(), 1, st));
        KArgs k = t.kargs;
        k.probs = t.s_probs.as<Prob>();
        k.groups = t.s_groups.as<Group>();
        k.n_groups = int32_t(groups.size());
        k.queue = t.s_queue.as<uint32_t>();
        k.tb = s.b_ckpt_tb.as<uint32_t>();
        k.seeds = nullptr;
        k.scores = t.s_scores.as<float>();  // a throwaway copy of the scores (whole-path groups write them)
        k.dvals = nullptr;
        k.val_layers = 1;
        k.progress = t.s_progress.as<uint32_t>();
        k.partials = t.s_partials.as<BandPartial>();
        k.band_err = t.s_err.as<uint32_t>();
        k.band_sync = 0;  // the rows above were all kept by the forward
        const int64_t wg = (int64_t(groups.size()) + 3) / 4;
        spec.grid = int(std::max<int64_t>(1, std::min<int64_t>(wg, int64_t(s.cus) * std::max(1, waves_per_cu / 4))));
        HIPCHK(launch_solve(spec, k, st));
    };
    auto one_group = [&](int q, const Prob &pr, int strip) {
        Group g;
        std::memset(&g, 0, sizeof(g));
        g.first = q;
        g.count = 1;
        g.L = pr.L;
        g.lgL = ilog2(pr.L);
        g.max_lt = pr.len_t;
        g.max_nst = strips_of(pr);
        g.R = pr.R;
        g.strip = strip;
        return g;
    };
    // whole paths.  All of them are recomputed in one launch on st where
    // the budget allows (one launch keeps the longest problem's serial time
    // off the others'), else in chunks of at most half the budget, into
    // alternating halves of the scratch buffer; each chunk's walk runs on a
    // second stream while the next chunk is recomputed (a half is reused
    // only after the walk two chunks back).  The problems that kept their
    // traceback are walked on the second stream first, during the first
    // recomputation.
    if (!s.aux_stream) HIPCHK(hipStreamCreateWithFlags(&s.aux_stream, hipStreamNonBlocking));
    hipStream_t aux = s.aux_stream;
    struct Chunk {
        size_t p0, np, g0, ng;
        int64_t words;
    };
    std::vector<Chunk> chunks;
    std::vector<Prob> allp;
    std::vector<Group> allg;
    {
        int64_t total = 0;
        for (const Prob &pr : whole) total += words_of(pr) * strips_of(pr);
        const int64_t limit = total <= budget ? budget : budget / 2;
        size_t w0 = 0;
        while (w0 < whole.size()) {
            Chunk c{allp.size(), 0, allg.size(), 0, 0};
            while (w0 < whole.size() &&
                   (c.np == 0 || c.words + words_of(whole[w0]) * strips_of(whole[w0]) <= limit)) {
                Prob pr = whole[w0++];
                pr.tb_off = c.words;  // the forward's layout, at the start of the chunk's half
                c.words += words_of(pr) * strips_of(pr);
                // one wavefront sweeps all strips in order, as in a forward that
                // keeps the traceback (the kept rows are rewritten unchanged)
                allg.push_back(one_group(int(c.np), pr, -1));
                allp.push_back(pr);
                c.np++;
            }
            c.ng = allg.size() - c.g0;
            // longest problems first
            std::stable_sort(allg.begin() + int64_t(c.g0), allg.end(), [&](const Group &x, const Group &y) {
                return int64_t(x.max_nst) * (x.max_lt + 63) > int64_t(y.max_nst) * (y.max_lt + 63);
            });
            chunks.push_back(c);
        }
    }
    const size_t nch = chunks.size();
    int64_t half = 0;
    for (const Chunk &c : chunks) half = std::max(half, c.words);
    s.b_ckpt_tb.ensure(size_t(std::max<int64_t>(nch > 1 ? 2 * half : half, 1)) * 4);
    while (s.ck_events.size() < 2 * nch + 2) {
        hipEvent_t e;
        HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        s.ck_events.push_back(e);
    }
    hipEvent_t *ev_rec = s.ck_events.data() + 2, *ev_walk = s.ck_events.data() + 2 + nch;
    if (!allp.empty()) {
        upload(st, t.s_probs, allp.data(), allp.size());
        upload(st, t.s_groups, allg.data(), allg.size());
    }
    t.s_queue.ensure(4 * (nch + 1));
    HIPCHK(launch_zero_u32(t.s_queue.as<uint32_t>(), int64_t(nch + 1), st));
    HIPCHK(hipEventRecord(s.ck_events[0], st));
    HIPCHK(hipStreamWaitEvent(aux, s.ck_events[0], 0));
    HIPCHK(launch_backward(t.gap, form, t.dim, base, aux));  // skips the checkpointed problems
    for (size_t c = 0; c < nch; c++) {
        const Chunk &ch = chunks[c];
        if (c >= 2) HIPCHK(hipStreamWaitEvent(st, ev_walk[c - 2], 0));
        uint32_t *tbh = s.b_ckpt_tb.as<uint32_t>() + (c & 1) * half;
        KArgs k = t.kargs;
        k.probs = t.s_probs.as<Prob>() + ch.p0;
        k.groups = t.s_groups.as<Group>() + ch.g0;
        k.n_groups = int32_t(ch.ng);
        k.queue = t.s_queue.as<uint32_t>() + c;
        k.tb = tbh;
        k.seeds = nullptr;
        k.scores = t.s_scores.as<float>();  // a throwaway copy of the scores
        k.dvals = nullptr;
        k.val_layers = 1;
        k.progress = t.s_progress.as<uint32_t>();
        k.partials = t.s_partials.as<BandPartial>();
        k.band_err = t.s_err.as<uint32_t>();
        k.band_sync = 0;
        const int64_t wg = (int64_t(ch.ng) + 3) / 4;
        spec.grid = int(std::max<int64_t>(1, std::min<int64_t>(wg, int64_t(s.cus) * std::max(1, waves_per_cu / 4))));
        HIPCHK(launch_solve(spec, k, st));
        HIPCHK(hipEventRecord(ev_rec[c], st));
        HIPCHK(hipStreamWaitEvent(aux, ev_rec[c], 0));
        BwArgs bw = base;
        bw.probs = t.s_probs.as<Prob>() + ch.p0;
        bw.n = int64_t(ch.np);
        bw.tb = tbh;
        HIPCHK(launch_backward(t.gap, form, t.dim, bw, aux));
        HIPCHK(hipEventRecord(ev_walk[c], aux));
    }
    HIPCHK(hipEventRecord(s.ck_events[1], aux));
    HIPCHK(hipStreamWaitEvent(st, s.ck_events[1], 0));  // everything below is ordered after the walks
    // too large for a whole path: one strip per pass
    size_t c0 = 0;
    while (c0 < by_strip.size()) {
        size_t c1 = c0;
        int64_t words = 0;
        while (c1 < by_strip.size() && (c1 == c0 || words + words_of(by_strip[c1]) <= budget)) words += words_of(by_strip[c1++]);
        const int nc = int(c1 - c0);
        std::vector<int64_t> base_off(static_cast<size_t>(nc));
        std::vector<StripWalk> init(static_cast<size_t>(nc));
        std::vector<int> first(static_cast<size_t>(nc));  // the seed's strip: the walk's first
        int passes = 0;
        int64_t acc = 0;
        for (int q = 0; q < nc; q++) {
            const Prob &pr = by_strip[c0 + size_t(q)];
            base_off[size_t(q)] = acc;
            acc += words_of(pr);
            int32_t seed[2];  // the forward's (global: the last cell)
            HIPCHK(hipMemcpyAsync(seed, base.seeds + 2 * pr.out, 8, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            StripWalk w;
            std::memset(&w, 0, sizeof(w));
            w.i = seed[0];
            w.j = seed[1];
            w.cur_i = w.cur_j = -1;
            init[size_t(q)] = w;
            // no seed (a local problem with no cell above zero): one pass
            // that finds the walk finished and adds nothing
            first[size_t(q)] = seed[0] >= 0 && seed[1] >= 0 ? seed[0] / (pr.L * pr.R) : 0;
            passes = std::max(passes, first[size_t(q)] + 1);
        }
        s.b_ckpt_tb.ensure(size_t(std::max<int64_t>(acc, 1)) * 4);
        upload(st, t.s_state, init.data(), init.size());
        for (int it = 0; it < passes; it++) {
            std::vector<Prob> probs(by_strip.begin() + int64_t(c0), by_strip.begin() + int64_t(c1));
            std::vector<int32_t> strip_of(static_cast<size_t>(nc), -1);
            std::vector<Group> groups;
            for (int q = 0; q < nc; q++) {
                Prob &pr = probs[size_t(q)];
                const int sp = first[size_t(q)] - it;
                if (sp < 0) continue;
                // the strip's words land at this problem's scratch slot
                pr.tb_off = base_off[size_t(q)] - int64_t(sp) * words_of(pr);
                strip_of[size_t(q)] = sp;
                groups.push_back(one_group(q, pr, sp));
            }
            upload(st, t.s_strip, strip_of.data(), strip_of.size());
            recompute(probs, groups);
            BwArgs bw = base;
            bw.probs = t.s_probs.as<Prob>();
            bw.n = nc;
            bw.tb = s.b_ckpt_tb.as<uint32_t>();
            HIPCHK(launch_backward_strips(t.gap, t.form, t.dim, bw, t.s_state.as<StripWalk>(),
                                          t.s_strip.as<int32_t>(), st));
            HIPCHK(hipStreamSynchronize(st));  // the host arrays above are reused
        }
        c0 = c1;
    }
    uint32_t err = 0;
    HIPCHK(hipMemcpyAsync(&err, t.s_err.p, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (err) fail(PA_ERUNTIME, "checkpointed backward: a strip recomputation waited too long");
}

}  // namespace

extern "C" {

int pa_backward(pa_tape *t, const float *grad_scores_device, float *grad_dense_device,
                float *grad_series_s_device, float *grad_series_t_device, float *grad_gap_device) {
    return guard([&] {
        if (!t || !grad_scores_device) fail(PA_EINVAL, "null argument");
        if (!t->solver) fail(PA_EINVAL, "the tape's solver was freed");
        BwArgs a{};
        a.probs = t->probs.as<Prob>();
        a.n = t->n_probs;
        a.tb = t->tb.as<uint32_t>();
        a.seeds = t->seeds.as<int32_t>();
        a.global = t->loc == LOC_GLOBAL ? 1 : 0;
        a.gsign = t->dir == DIR_MAX ? -1.0f : 1.0f;
        a.grad_scores = grad_scores_device;
        a.grad_dense = t->form == F_DENSE ? grad_dense_device : nullptr;
        a.xs = t->xs;
        a.xt = t->xt;
        a.grad_xs = t->form == F_SERIES ? grad_series_s_device : nullptr;
        a.grad_xt = t->form == F_SERIES ? grad_series_t_device : nullptr;
        a.grad_gap = t->gap == GAP_DTW ? nullptr : grad_gap_device;
        const int form = (t->form == F_TABLE || t->form == F_BINARY) ? F_DENSE : t->form;
        if (t->form == F_TABLE || t->form == F_BINARY) a.grad_dense = nullptr;  // no similarity input to differentiate
        hipStream_t st = t->solver->stream;
        if (t->ckpt.empty()) HIPCHK(launch_backward(t->gap, form, t->dim, a, st));
        else backward_checkpointed(*t, a, form, st);  // also walks the other problems
        // on the caller's stream the gradients are ordered there; the
        // solver's own stream is waited for
        if (!t->solver->external_stream) HIPCHK(hipStreamSynchronize(st));
    });
}

void pa_tape_free(pa_tape *t) {
    if (!t) return;
    if (pa_solver *s = t->solver) {
        s->tapes.erase(std::remove(s->tapes.begin(), s->tapes.end(), t), s->tapes.end());
        // the larger of each pair stays with the solver: the solver's later
        // work is ordered after the tape's on the same stream, and hipFree
        // of the smaller one waits for the device
        auto give_back = [](DevBuf &from, DevBuf &to) {
            if (from.cap <= to.cap) return false;
            std::swap(from.p, to.p);
            std::swap(from.cap, to.cap);
            return true;
        };
        // the cached plan's descriptors live in b_probs
        if (give_back(t->probs, s->b_probs)) s->last_plan.uploaded = false;
        give_back(t->seeds, s->b_seeds);
        give_back(t->tb, s->b_tb);
        give_back(t->rowbuf, s->b_rowbuf);
        give_back(t->borders, s->b_tape_borders);
    }
    delete t;
}

// ---- results ---------------------------------------------------------------

int64_t pa_result_count(const pa_result *r) { return r ? r->n : 0; }

int pa_result_scores(const pa_result *r, float *out) {
    return guard([&] {
        if (!r || !out) fail(PA_EINVAL, "null argument");
        std::copy(r->scores.begin(), r->scores.end(), out);
    });
}

int pa_result_has_path(const pa_result *r, int64_t k) {
    if (!r || k < 0 || k >= r->n) return 0;
    if (r->goal == PA_GOAL_ONE) return r->path_len[size_t(k)] >= 0 ? 1 : 0;
    return 0;
}

int64_t pa_result_path_len(const pa_result *r, int64_t k) {
    if (!r || k < 0 || k >= r->n || r->goal != PA_GOAL_ONE) return -1;
    return r->path_len[size_t(k)];
}

int pa_result_path(const pa_result *r, int64_t k, int32_t *uv) {
    return guard([&] {
        if (!r || !uv || k < 0 || k >= r->n || r->goal != PA_GOAL_ONE) fail(PA_EINVAL, "no path for this problem");
        const int32_t m = r->path_len[size_t(k)];
        if (m < 0) fail(PA_EINVAL, "no path for this problem");
        std::memcpy(uv, r->paths.data() + 2 * r->path_off[size_t(k)], size_t(m) * 8);
    });
}

int pa_result_paths(const pa_result *r, int64_t *lens, int32_t *uv) {
    return guard([&] {
        if (!r || !lens || r->goal != PA_GOAL_ONE) fail(PA_EINVAL, "not a one-optimal result");
        int64_t at = 0;
        for (int64_t k = 0; k < r->n; k++) {
            const int32_t m = r->path_len[size_t(k)];
            lens[k] = m;
            if (m <= 0) continue;
            if (uv) std::memcpy(uv + 2 * at, r->paths.data() + 2 * r->path_off[size_t(k)], size_t(m) * 8);
            at += m;
        }
    });
}

// path_compressor (solver.h:809-846) + Alignment::add_edge (factory.h:164-170)
static void alignment_from_path(const int32_t *uv, int64_t m, int32_t ls, int32_t lt, int32_t *s_to_t,
                                int32_t *t_to_s) {
    std::fill(s_to_t, s_to_t + ls, -1);
    std::fill(t_to_s, t_to_s + lt, -1);
    const bool by_s = ls <= lt;
    std::vector<int32_t> primary(size_t(by_s ? ls : lt), -1);
    for (int64_t q = 1; q < m; q++) {
        const int32_t lu = uv[2 * (q - 1)], lv = uv[2 * (q - 1) + 1];
        const int32_t u = uv[2 * q], v = uv[2 * q + 1];
        if (u != lu && v != lv && lu >= 0 && lv >= 0) {
            if (by_s) primary[size_t(lu)] = lv;
            else primary[size_t(lv)] = lu;
        }
    }
    // the other direction comes from inversion (factory.h:110-128)
    if (by_s) {
        for (int32_t i = 0; i < ls; i++) s_to_t[i] = primary[size_t(i)];
        for (int32_t i = 0; i < ls; i++)
            if (s_to_t[i] >= 0) t_to_s[s_to_t[i]] = i;
    } else {
        for (int32_t j = 0; j < lt; j++) t_to_s[j] = primary[size_t(j)];
        for (int32_t j = 0; j < lt; j++)
            if (t_to_s[j] >= 0) s_to_t[t_to_s[j]] = j;
    }
}

int pa_result_alignments(const pa_result *r, int32_t *s_to_t, int32_t *t_to_s) {
    return guard([&] {
        if (!r || !s_to_t || !t_to_s || r->goal != PA_GOAL_ONE) fail(PA_EINVAL, "not a one-optimal result");
        int64_t as = 0, at = 0;
        for (int64_t k = 0; k < r->n; k++) {
            const int32_t ls = r->len_s[size_t(k)], lt = r->len_t[size_t(k)];
            const int32_t m = r->path_len[size_t(k)];
            if (m < 0) {
                std::fill(s_to_t + as, s_to_t + as + ls, -1);
                std::fill(t_to_s + at, t_to_s + at + lt, -1);
            } else {
                alignment_from_path(r->paths.data() + 2 * r->path_off[size_t(k)], m, ls, lt, s_to_t + as,
                                    t_to_s + at);
            }
            as += ls;
            at += lt;
        }
    });
}

int pa_result_alignment(const pa_result *r, int64_t k, int32_t *s_to_t, int32_t *t_to_s) {
    return guard([&] {
        if (!r || k < 0 || k >= r->n) fail(PA_EINVAL, "bad problem index");
        const int32_t ls = r->len_s[size_t(k)], lt = r->len_t[size_t(k)];
        if (r->goal == PA_GOAL_ONE) {
            const int32_t m = r->path_len[size_t(k)];
            if (m < 0) {
                std::fill(s_to_t, s_to_t + ls, -1);
                std::fill(t_to_s, t_to_s + lt, -1);
                return;
            }
            alignment_from_path(r->paths.data() + 2 * r->path_off[size_t(k)], m, ls, lt, s_to_t, t_to_s);
        } else if (r->goal == PA_GOAL_ALL) {
            const IterState *it = r->iters[size_t(k)].get();
            if (!it || it->cur_path.empty()) fail(PA_EINVAL, "no current alignment");
            std::vector<int32_t> uv;
            for (auto &p : it->cur_path) {
                uv.push_back(p.first);
                uv.push_back(p.second);
            }
            alignment_from_path(uv.data(), int64_t(it->cur_path.size()), ls, lt, s_to_t, t_to_s);
        } else {
            fail(PA_EINVAL, "score-only results carry no alignment");
        }
    });
}

// ---- all-optimal enumeration over the GPU's 8-bit tie sets ----------------

// the traceback code of cell (i, j): bits per cell 2 (linear, DTW), 4
// (affine) or 8 (all-optimal tie sets); layout as written by the kernels
static uint32_t cell_bits(const pa_result *r, const Prob &pr, int32_t i, int32_t j, int bits) {
    const int L = pr.L, R = pr.R;
    const int rps = L * R;
    const int wpc = (bits * R + 31) / 32;
    const int st = i / rps;
    const int rem = i - st * rps;
    const int kk = rem / R;
    const int rr = rem - kk * R;
    const int64_t s = j + kk;
    const int64_t nsteps = pr.len_t + L - 1;
    const uint32_t w = r->tb[size_t(pr.tb_off + ((int64_t(st) * nsteps + s) * L + kk) * wpc + (rr * bits) / 32)];
    return (w >> ((rr * bits) % 32)) & ((1u << bits) - 1u);
}

static uint32_t tie_bits(const pa_result *r, const Prob &pr, int32_t i, int32_t j) {
    return cell_bits(r, pr, i, j, 8);
}

// the edge list of a traceback_n cell: candidates equal to the cell's value
// in push order (solver.h:1324-1427)
static void push_gap_list(const pa_result *r, const Prob &pr, int32_t i, int32_t j, bool vertical,
                          std::vector<std::pair<int32_t, int32_t>> &out) {
    while (true) {
        if (i < 0 || j < 0) return;  // border P/Q traceback is cleared (solver.h:2953-2966)
        const uint32_t b = tie_bits(r, pr, i, j);
        const bool open = vertical ? (b & 16u) : (b & 64u);
        const bool ext = vertical ? (b & 32u) : (b & 128u);
        if (open) out.emplace_back(vertical ? i - 1 : i, vertical ? j : j - 1);
        if (!ext) return;
        if (vertical) i -= 1;
        else j -= 1;
    }
}

// General gap costs: the cell's value against every candidate, recomputed
// from the kernel's D and S planes in the same float32 arithmetic
// (solver.h:3161-3180).  A NaN cell came from a NaN first candidate, which
// no later candidate replaces or joins.
static float wsb_D(const pa_result *r, const Prob &pr, int32_t i, int32_t j) {
    if (i < 0) return r->border_t[size_t(j + 1)];
    if (j < 0) return r->border_s[size_t(i + 1)];
    return r->dvals[size_t(pr.dv_off + int64_t(i) * pr.len_t + j)];
}

static void wsb_cell_edges(const pa_result *r, const Prob &pr, int32_t u, int32_t v, int32_t no_tb,
                           std::vector<std::pair<int32_t, int32_t>> &out) {
    const float x = wsb_D(r, pr, u, v);
    const int64_t cells = int64_t(pr.len_s) * pr.len_t;
    const float S = r->dvals[size_t(pr.dv_off + 2 * cells + int64_t(u) * pr.len_t + v)];
    if (std::isnan(x)) {
        out.emplace_back(u - 1, v - 1);
        return;
    }
    if (r->loc == PA_LOCAL && x == 0.0f) out.emplace_back(no_tb, no_tb);
    if (wsb_D(r, pr, u - 1, v - 1) + S == x) out.emplace_back(u - 1, v - 1);
    for (int32_t k = -1; k < u; k++)
        if (wsb_D(r, pr, k, v) + r->ws_sgn[size_t(u - k)] == x) out.emplace_back(k, v);
    for (int32_t k = -1; k < v; k++)
        if (wsb_D(r, pr, u, k) + r->wt_sgn[size_t(v - k)] == x) out.emplace_back(u, k);
}

static void cell_edges(const pa_result *r, const Prob &pr, int32_t u, int32_t v, int32_t no_tb,
                       std::vector<std::pair<int32_t, int32_t>> &out) {
    out.clear();
    if (r->gap == GAP_GEN) {
        wsb_cell_edges(r, pr, u, v, no_tb, out);
        return;
    }
    const uint32_t b = tie_bits(r, pr, u, v);
    if (b & 1u) out.emplace_back(no_tb, no_tb);
    if (b & 2u) out.emplace_back(u - 1, v - 1);
    if (r->gap == GAP_AFF) {
        if (b & 4u) push_gap_list(r, pr, u, v, true, out);
        if (b & 8u) push_gap_list(r, pr, u, v, false, out);
    } else {
        if (b & 4u) out.emplace_back(u - 1, v);
        if (b & 8u) out.emplace_back(u, v - 1);
    }
}

int pa_result_iter_next(pa_result *r, int64_t k, float *score) {
    int found = 0;
    const int status = guard([&] {
        if (!r || k < 0 || k >= r->n) fail(PA_EINVAL, "bad problem index");
        if (r->goal != PA_GOAL_ALL) fail(PA_EINVAL, "not an all-optimal result");
        auto &slot = r->iters[size_t(k)];
        if (!slot) slot = std::make_unique<IterState>();
        IterState &it = *slot;
        const Prob &pr = r->probs[size_t(k)];
        const int32_t ls = r->len_s[size_t(k)], lt = r->len_t[size_t(k)];
        const int32_t no_tb = r->index_bits == 16 ? int32_t(INT16_MIN) : INT32_MIN;
        const bool local = r->loc == PA_LOCAL;
        const bool mx = r->dir == PA_MAXIMIZE;
        auto Dval = [&](int32_t i, int32_t j) { return r->dvals[size_t(pr.dv_off + int64_t(i) * lt + j)]; };
        auto cont = [&](int32_t u, int32_t v) {
            if (!(u >= 0 && v >= 0)) return false;
            if (local && r->gap == GAP_GEN) return mx ? Dval(u, v) > 0.0f : Dval(u, v) < 0.0f;
            if (local) return (tie_bits(r, pr, u, v) & 1u) == 0;  // D(u, v) is not zero
            return true;
        };
        if (!it.started) {
            it.started = true;
            // seeds (solver.h:2060-2092, 2179-2187, 2273-2293)
            if (local) {
                if (ls > 0 && lt > 0) {
                    float best = 0.0f;
                    for (int32_t i = ls - 1; i >= 0; i--)
                        for (int32_t j = lt - 1; j >= 0; j--) {
                            const float x = Dval(i, j);
                            if (mx ? x > best : x < best) best = x;
                        }
                    for (int32_t i = ls - 1; i >= 0; i--)
                        for (int32_t j = lt - 1; j >= 0; j--) {
                            const float x = Dval(i, j);
                            if (x == best && (mx ? x > 0.0f : x < 0.0f))
                                it.stack.push_back(IterState::Entry{x, i, j, 0});
                        }
                }
            } else {
                const int32_t si = r->seeds[size_t(2 * k)], sj = r->seeds[size_t(2 * k + 1)];
                if (si != NO_SEED) {
                    const float val = r->loc == PA_GLOBAL ? r->scores[size_t(k)] : r->scores[size_t(k)];
                    it.stack.push_back(IterState::Entry{val, si, sj, 0});
                }
            }
        }
        std::vector<std::pair<int32_t, int32_t>> edges;
        while (!it.stack.empty()) {
            const IterState::Entry e = it.stack.back();
            it.stack.pop_back();
            if (e.path_len == 0) it.path.clear();
            else it.path.resize(size_t(e.path_len));
            it.path.emplace_back(e.u, e.v);
            if (cont(e.u, e.v)) {
                cell_edges(r, pr, e.u, e.v, no_tb, edges);
                const int64_t ps = int64_t(it.path.size());
                if (!edges.empty()) {
                    for (auto &q : edges) it.stack.push_back(IterState::Entry{e.val, q.first, q.second, ps});
                } else {
                    it.stack.push_back(IterState::Entry{e.val, no_tb, no_tb, ps});
                }
            } else {
                if (r->remove_dup) {
                    std::vector<std::pair<int32_t, int32_t>> key;
                    for (size_t q = 1; q < it.path.size(); q++) {
                        const auto &a0 = it.path[q - 1], &a1 = it.path[q];
                        if (a1.first != a0.first && a1.second != a0.second && a0.first >= 0 && a0.second >= 0)
                            key.push_back(a0);
                    }
                    if (!it.emitted.insert(key).second) continue;
                }
                it.cur_path = it.path;
                if (score) *score = e.val;
                found = 1;
                return;
            }
        }
        it.cur_path.clear();
    });
    return status == PA_OK ? found : -status;
}

int64_t pa_result_iter_path_len(const pa_result *r, int64_t k) {
    if (!r || k < 0 || k >= r->n || r->goal != PA_GOAL_ALL || !r->iters[size_t(k)]) return -1;
    return int64_t(r->iters[size_t(k)]->cur_path.size());
}

int pa_result_iter_path(const pa_result *r, int64_t k, int32_t *uv) {
    return guard([&] {
        if (!r || !uv || k < 0 || k >= r->n || r->goal != PA_GOAL_ALL || !r->iters[size_t(k)])
            fail(PA_EINVAL, "no current alignment");
        const auto &p = r->iters[size_t(k)]->cur_path;
        for (size_t q = 0; q < p.size(); q++) {
            uv[2 * q] = p[q].first;
            uv[2 * q + 1] = p[q].second;
        }
    });
}

void pa_result_free(pa_result *r) { delete r; }

}  // extern "C"

// ---- Solution matrices (Solution::set_values / traceback_as_matrix /
// traceback_as_edges, solver.h:2355-2471) ------------------------------------
//
// Values: interior cells are the kernel's per-layer cell values; the border
// row and column hold the solver's initial values (affine: D and P share the
// column, D and Q the row, P's row and Q's column hold the worst value,
// solver.h:2926-2951; DTW: worst except D(-1,-1) = 0, solver.h:3205-3210).
// Tracebacks: border cells are cleared (solver.h:2953-2966); interior cells
// are rebuilt from the traceback codes.  One optimal path (traceback_1,
// solver.h:196-278): the single predecessor, which P and Q copy along
// extended gaps (solver.h:2993-3016).  All optimal paths (traceback_n,
// solver.h:280-360): the predecessors of every candidate equal to the cell's
// value, in push order.

static const Prob &matrices_prob(const pa_result *r, int64_t k, int32_t layer) {
    if (!r || k < 0 || k >= r->n) fail(PA_EINVAL, "bad problem index");
    if (!r->matrices) fail(PA_EINVAL, "the result keeps no solution matrices");
    if (layer < 0 || layer >= r->layers) fail(PA_EINVAL, "bad layer");
    return r->probs[size_t(k)];
}

static int32_t no_traceback_of(const pa_result *r) {  // solver.h:182-185
    return r->index_bits == 16 ? int32_t(INT16_MIN) : INT32_MIN;
}

// first predecessor per cell and layer, [layers][ls+1][lt+1][2]
static void fill_traceback(const pa_result *r, const Prob &pr, int32_t *out) {
    const int64_t ls = pr.len_s, lt = pr.len_t, W = lt + 1, plane = (ls + 1) * W;
    const int32_t no_tb = no_traceback_of(r);
    for (int64_t q = 0; q < r->layers * plane * 2; q++) out[q] = no_tb;
    if (r->gap == GAP_GEN) {
        // the kernel's traceback codes: the first candidate holding the value,
        // which is also the first entry of an all-optimal list
        const int32_t *code = reinterpret_cast<const int32_t *>(r->dvals.data() + pr.dv_off + 3 * ls * lt);
        for (int64_t u = 0; u < ls; u++) {
            for (int64_t v = 0; v < lt; v++) {
                const int32_t c = code[u * lt + v];
                int32_t *d = out + 2 * ((u + 1) * W + v + 1);
                if (c == 0) continue;  // local zero: (no_tb, no_tb)
                if (c == 1) {
                    d[0] = int32_t(u - 1);
                    d[1] = int32_t(v - 1);
                } else if (c <= u + 2) {
                    d[0] = c - 3;
                    d[1] = int32_t(v);
                } else {
                    d[0] = int32_t(u);
                    d[1] = int32_t(c - u - 4);
                }
            }
        }
        return;
    }
    const bool aff = r->gap == GAP_AFF;
    const bool all = r->tr == TR_ALL;
    const int bits = all ? 8 : (aff ? 4 : 2);
    // all-optimal P/Q lists may be empty (a cleared border copied along)
    std::vector<uint8_t> ne_p, ne_q;
    if (all && aff) {
        ne_p.assign(size_t(plane), 0);
        ne_q.assign(size_t(plane), 0);
    }
    auto at = [&](int layer, int64_t i, int64_t j) { return out + 2 * (layer * plane + i * W + j); };
    auto put = [](int32_t *c, int64_t u, int64_t v) {
        c[0] = int32_t(u);
        c[1] = int32_t(v);
    };
    for (int64_t i = 1; i <= ls; i++) {
        for (int64_t j = 1; j <= lt; j++) {
            const int64_t u = i - 1, v = j - 1;  // DP cell
            const uint32_t b = cell_bits(r, pr, int32_t(u), int32_t(v), bits);
            int32_t *d = at(0, i, j);
            if (!all) {
                if (aff) {
                    int32_t *p = at(1, i, j), *q = at(2, i, j);
                    if (b & 4u) std::memcpy(p, at(1, i - 1, j), 8);
                    else put(p, u - 1, v);
                    if (b & 8u) std::memcpy(q, at(2, i, j - 1), 8);
                    else put(q, u, v - 1);
                    switch (b & 3u) {
                        case 0: break;  // local zero: (no_tb, no_tb)
                        case 1: put(d, u - 1, v - 1); break;
                        case 2: std::memcpy(d, p, 8); break;
                        default: std::memcpy(d, q, 8); break;
                    }
                } else {
                    switch (b & 3u) {
                        case 0: break;
                        case 1: put(d, u - 1, v - 1); break;
                        case 2: put(d, u - 1, v); break;
                        default: put(d, u, v - 1); break;
                    }
                }
                continue;
            }
            bool p_ne = false, q_ne = false;
            if (aff) {
                int32_t *p = at(1, i, j), *q = at(2, i, j);
                const int64_t c = i * W + j;
                if (b & 16u) { put(p, u - 1, v); p_ne = true; }
                else if ((b & 32u) && ne_p[size_t(c - W)]) { std::memcpy(p, at(1, i - 1, j), 8); p_ne = true; }
                if (b & 64u) { put(q, u, v - 1); q_ne = true; }
                else if ((b & 128u) && ne_q[size_t(c - 1)]) { std::memcpy(q, at(2, i, j - 1), 8); q_ne = true; }
                ne_p[size_t(c)] = p_ne;
                ne_q[size_t(c)] = q_ne;
            }
            if (b & 1u) continue;  // the local zero's (no_tb, no_tb) comes first
            if (b & 2u) put(d, u - 1, v - 1);
            else if (b & 4u) {
                if (!aff) put(d, u - 1, v);
                else if (p_ne) std::memcpy(d, at(1, i, j), 8);
                else if ((b & 8u) && q_ne) std::memcpy(d, at(2, i, j), 8);
            } else if (b & 8u) {
                if (!aff) put(d, u, v - 1);
                else if (q_ne) std::memcpy(d, at(2, i, j), 8);
            }
        }
    }
}

// every traceback edge of one layer in the reference's order: cells row by
// row, each cell's predecessors in push order; f(from_u, from_v, to_u, to_v)
template <class F>
static void for_each_edge(const pa_result *r, const Prob &pr, int32_t layer, F &&f) {
    const int64_t ls = pr.len_s, lt = pr.len_t, W = lt + 1;
    if (r->tr != TR_ALL) {
        // traceback_1 cells always hold one entry (Single::size() == 1)
        std::vector<int32_t> tb(size_t(r->layers * (ls + 1) * W * 2));
        fill_traceback(r, pr, tb.data());
        const int32_t *t = tb.data() + 2 * layer * (ls + 1) * W;
        for (int64_t i = 0; i <= ls; i++)
            for (int64_t j = 0; j <= lt; j++)
                f(int32_t(i - 1), int32_t(j - 1), t[2 * (i * W + j)], t[2 * (i * W + j) + 1]);
        return;
    }
    const int32_t no_tb = no_traceback_of(r);
    std::vector<std::pair<int32_t, int32_t>> lst;
    for (int64_t i = 1; i <= ls; i++) {  // border cells hold empty lists
        for (int64_t j = 1; j <= lt; j++) {
            const int32_t u = int32_t(i - 1), v = int32_t(j - 1);
            lst.clear();
            if (layer == 0) cell_edges(r, pr, u, v, no_tb, lst);
            else push_gap_list(r, pr, u, v, layer == 1, lst);
            for (const auto &e : lst) f(u, v, e.first, e.second);
        }
    }
}

extern "C" {

int pa_solve_ragged_matrices(pa_solver *s, const pa_ragged *batch, pa_result **out) {
    return guard([&] {
        if (!s || !out) fail(PA_EINVAL, "null argument");
        Input in{};
        ragged_input(s, batch, in);
        in.matrices = true;
        std::unique_ptr<pa_result> r;
        run(*s, in, &r, nullptr);
        *out = r.release();
    });
}

int32_t pa_result_layers(const pa_result *r) { return r && r->matrices ? r->layers : 0; }

int pa_result_values(const pa_result *r, int64_t k, float *out) {
    return guard([&] {
        const Prob &pr = matrices_prob(r, k, 0);
        if (!out) fail(PA_EINVAL, "null argument");
        const int64_t ls = pr.len_s, lt = pr.len_t, W = lt + 1, plane = (ls + 1) * W;
        const float wv = worst(r->dir);
        for (int layer = 0; layer < r->layers; layer++) {
            float *o = out + layer * plane;
            for (int64_t j = 0; j <= lt; j++) o[j] = (layer == 1 && j > 0) ? wv : r->border_t[size_t(j)];
            for (int64_t i = 1; i <= ls; i++) o[i * W] = layer == 2 ? wv : r->border_s[size_t(i)];
            const float *src = r->dvals.data() + pr.dv_off + int64_t(layer) * ls * lt;
            for (int64_t i = 1; i <= ls; i++)
                std::memcpy(o + i * W + 1, src + (i - 1) * lt, size_t(lt) * 4);
        }
    });
}

int pa_result_traceback(const pa_result *r, int64_t k, int32_t *out) {
    return guard([&] {
        const Prob &pr = matrices_prob(r, k, 0);
        if (!out) fail(PA_EINVAL, "null argument");
        fill_traceback(r, pr, out);
    });
}

int64_t pa_result_edges_len(const pa_result *r, int64_t k, int32_t layer) {
    int64_t n = 0;
    const int status = guard([&] {
        const Prob &pr = matrices_prob(r, k, layer);
        for_each_edge(r, pr, layer, [&](int32_t, int32_t, int32_t, int32_t) { n++; });
    });
    return status == PA_OK ? n : -1;
}

int pa_result_edges(const pa_result *r, int64_t k, int32_t layer, int32_t *out) {
    return guard([&] {
        const Prob &pr = matrices_prob(r, k, layer);
        if (!out) fail(PA_EINVAL, "null argument");
        int64_t q = 0;
        for_each_edge(r, pr, layer, [&](int32_t a, int32_t b, int32_t c, int32_t d) {
            out[4 * q] = a;
            out[4 * q + 1] = b;
            out[4 * q + 2] = c;
            out[4 * q + 3] = d;
            q++;
        });
    });
}

}  // extern "C"
