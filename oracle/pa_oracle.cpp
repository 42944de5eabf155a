// pa_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of the reference dynamic-programming solver
// (pyalign/algorithm/solver.h and the per-lane input forms of
// pyalign/algorithm/factory.h).  It is the checker that the HIP path is
// compared against, and the CPU baseline timed beside it.  It is never
// linked into, loaded by or called from the product library.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
//
// Each function names the reference lines it restates.  Lanes are a
// fixed-width array (template parameter B) laid out innermost, exactly as
// the reference's cell_type/tiny_vector (solver.h:376-422) does, so B=1 is
// the unbatched path and B=16 the AVX-512 batched path.
//
// Pinning: the unit tests in tests/test_oracle_*.py check every function
// here against the reference's own known-answer tests
// (pyalign/tests/*.py, README.md:94-110), against brute-force enumeration
// of all alignments on tiny inputs, and against closed forms.  The one
// part with no such check -- xtensor 0.23.10 linspace rounding for global
// affine borders when the solver limit is >= 1024 and not a power of two
// (solver.h:2876-2880) -- is marked "parity unpinned" below.
//
// Build: see oracle/Makefile (g++ -O3 -march=native -ffp-contract=off).

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace {

// enum values follow pyalign/algorithm/enum.h:6-31
enum { T_ALIGNMENT = 0, T_DTW = 1 };
enum { D_MINIMIZE = 0, D_MAXIMIZE = 1 };
enum { L_LOCAL = 0, L_GLOBAL = 1, L_SEMIGLOBAL = 2 };
enum { G_SCORE = 0, G_ONE = 1, G_ALL = 2 };
enum { GAP_LINEAR = 0, GAP_AFFINE = 1, GAP_GENERAL = 2 };

}  // namespace

extern "C" {

// Configuration of one reference solver instance (options.h:16-44 and
// factory.h:33-108 after the gap-cost special-case mapping).
typedef struct oc_config {
    int32_t type;        // T_ALIGNMENT / T_DTW
    int32_t direction;   // D_MINIMIZE / D_MAXIMIZE
    int32_t locality;    // L_LOCAL / L_GLOBAL / L_SEMIGLOBAL
    int32_t goal;        // G_SCORE / G_ONE / G_ALL
    int32_t gap_kind;    // GAP_LINEAR / GAP_AFFINE / GAP_GENERAL
    int32_t index_bits;  // 16 (m16) or 32 (m32)
    int32_t remove_dup;  // !return_dup (options.h:21)
    int32_t reserved;
    float lin_s, lin_t;                  // LinearGapCost u, as float32
    float aff_u_s, aff_v_s;              // AffineCost(u=extend, v=open-extend)
    float aff_u_t, aff_v_t;              //   for s and t (factory.h:48-52)
    int64_t lim_s, lim_t;                // solver limits (max_len_s/t)
    const float *gen_s, *gen_t;          // general: w(k), k = 0 .. lim (lim + 1 floats)
} oc_config;

}  // extern "C"

namespace {

struct Pt {
    int32_t u, v;
};

inline bool is_opt(int dir, float a, float b) {  // solver.h:43-85
    return dir == D_MAXIMIZE ? (a > b) : (a < b);
}

inline float worst_val(int dir) {  // solver.h:61-64, 83-86
    return dir == D_MAXIMIZE ? -std::numeric_limits<float>::infinity()
                             : std::numeric_limits<float>::infinity();
}

// AffineCost::vector (solver.h:2876-2880):
//   w = xt::linspace<float>(0, (n-1)*u, n) + v;  w[0] = 0
// xtensor 0.23.10 linspace evaluates start + k*step in double with
// step = double(stop - start) / (n - 1) and returns stop at k = n-1.
// When (n-1)*u is exact in float32 (every power-of-two limit, every
// integer extend) this equals fl32(fl32(k*u) + v).  Otherwise it rests on
// xtensor's generator rounding: parity unpinned for that case.
std::vector<float> affine_vector(float u, float v, int64_t n) {
    std::vector<float> w(n);
    const float stop = float(n - 1) * u;
    const double step = double(stop - 0.0f) / std::max(1.0, double(n - 1));
    for (int64_t k = 0; k < n; k++) {
        float lin;
        if (n > 1 && k == n - 1) {
            lin = stop;
        } else {
            lin = float(0.0 + double(k) * step);
        }
        w[k] = lin + v;
    }
    w[0] = 0.0f;
    return w;
}

// The per-problem output of one lane.
struct LaneResult {
    bool has_seed = false;
    float score = 0.0f;
    std::vector<Pt> path;         // seed first (build_path, solver.h:858-935)
    bool has_alignment = false;
    std::vector<Pt> edges;        // compressed path (path_compressor, solver.h:809-846)
};

// A full reference solver instance over B lanes (solver.h:2509-2745 plus
// the concrete solvers 2746-3045 and 3186-3260).
template <int B>
class LaneSolver {
public:
    oc_config cfg;
    int layers;
    int64_t ls1, lt1;                 // lim + 1
    int32_t no_tb;                    // no_traceback<Index>() (solver.h:182-185)
    float sgn;                        // gap_sgn
    float w1_s = 0, w1_t = 0, ext_s = 0, ext_t = 0;
    std::vector<float> gen_s, gen_t;  // general gap costs w(k)

    std::unique_ptr<float[]> val;     // [layer][i+1][j+1][lane]
    std::unique_ptr<int32_t[]> tbu, tbv;            // traceback_1
    std::vector<std::vector<Pt>> tbn;               // traceback_n
    bool trace;                                     // goal != score
    bool all;                                       // goal == all

    int64_t cur_len_s[B], cur_len_t[B];

    inline int64_t idx(int k, int64_t i, int64_t j) const {  // shifted by +1
        return ((int64_t(k) * ls1 + (i + 1)) * lt1 + (j + 1));
    }

    LaneSolver(const oc_config &c) : cfg(c) {
        if (cfg.type == T_DTW) {
            cfg.goal = G_ONE;  // factory.h:1072-1099: DTW is always one-optimal
            cfg.locality = L_GLOBAL;
        }
        trace = cfg.goal != G_SCORE;
        all = cfg.goal == G_ALL;
        layers = (cfg.type == T_ALIGNMENT && cfg.gap_kind == GAP_AFFINE) ? 3 : 1;
        const int64_t idx_max = (cfg.index_bits == 16 ? int64_t(INT16_MAX) : int64_t(INT32_MAX)) >> 1;
        if (cfg.lim_s > idx_max || cfg.lim_t > idx_max)
            throw std::runtime_error("exceeded_implementation_length");  // solver.h:496-503
        ls1 = cfg.lim_s + 1;
        lt1 = cfg.lim_t + 1;
        no_tb = cfg.index_bits == 16 ? int32_t(INT16_MIN) : INT32_MIN;
        sgn = cfg.direction == D_MINIMIZE ? 1.0f : -1.0f;
        const int64_t ncell = int64_t(layers) * ls1 * lt1;
        val.reset(new float[ncell * B]);
        if (trace && !all) {
            tbu.reset(new int32_t[ncell * B]);
            tbv.reset(new int32_t[ncell * B]);
        }
        if (all) tbn.resize(ncell * B);
        for (int b = 0; b < B; b++) cur_len_s[b] = cur_len_t[b] = 0;
        init_borders();
    }

    void set_val(int k, int64_t i, int64_t j, float x) {
        for (int b = 0; b < B; b++) val[idx(k, i, j) * B + b] = x;
    }
    void clear_tb(int k, int64_t i, int64_t j) {
        const int64_t c = idx(k, i, j) * B;
        for (int b = 0; b < B; b++) {
            if (tbu) { tbu[c + b] = no_tb; tbv[c + b] = no_tb; }
            if (all) tbn[c + b].clear();
        }
    }

    // Border setup: MatrixFactory ctor (solver.h:506-538) clears the
    // traceback of row 0 / column 0; the solver ctors fill border values:
    // linear 2806-2814, affine 2918-2967, DTW 3205-3210; localities'
    // init_border_case: Local 2021-2029, Global 2143-2155, Semiglobal 2253-2261.
    void init_borders() {
        for (int k = 0; k < layers; k++) {
            for (int64_t i = -1; i < cfg.lim_s; i++) clear_tb(k, i, -1);
            for (int64_t j = -1; j < cfg.lim_t; j++) clear_tb(k, -1, j);
        }
        const bool global = cfg.locality == L_GLOBAL;
        if (cfg.type == T_DTW) {
            const float w = worst_val(cfg.direction);
            for (int64_t i = -1; i < cfg.lim_s; i++)
                for (int64_t j = -1; j < cfg.lim_t; j++) set_val(0, i, j, w);
            set_val(0, -1, -1, 0.0f);
            return;
        }
        if (cfg.gap_kind == GAP_GENERAL) {
            // GeneralGapCostSolver ctor (solver.h:3114-3123): the column from
            // w_s * sgn, then the row from w_t * sgn (its corner last)
            gen_s.assign(cfg.gen_s, cfg.gen_s + ls1);
            gen_t.assign(cfg.gen_t, cfg.gen_t + lt1);
            for (int64_t i = 0; i < ls1; i++) set_val(0, i - 1, -1, global ? gen_s[i] * sgn : 0.0f);
            for (int64_t j = 0; j < lt1; j++) set_val(0, -1, j - 1, global ? gen_t[j] * sgn : 0.0f);
            return;
        }
        if (cfg.gap_kind == GAP_LINEAR) {
            for (int64_t i = 0; i < ls1; i++) {
                float x = global ? (float(i) * cfg.lin_s) * sgn : 0.0f;
                set_val(0, i - 1, -1, x);
            }
            for (int64_t j = 0; j < lt1; j++) {
                float x = global ? (float(j) * cfg.lin_t) * sgn : 0.0f;
                set_val(0, -1, j - 1, x);
            }
            return;
        }
        // affine: D=0, P=1, Q=2 (solver.h:2509-2513)
        const float inf = worst_val(cfg.direction);
        for (int64_t i = 0; i < ls1; i++) set_val(2, i - 1, -1, inf);  // Q column 0
        for (int64_t j = 0; j < lt1; j++) set_val(1, -1, j - 1, inf);  // P row 0
        std::vector<float> ws = affine_vector(cfg.aff_u_s, cfg.aff_v_s, ls1);
        std::vector<float> wt = affine_vector(cfg.aff_u_t, cfg.aff_v_t, lt1);
        for (int64_t i = 0; i < ls1; i++) {
            float x = global ? ws[i] * sgn : 0.0f;
            set_val(0, i - 1, -1, x);
            set_val(1, i - 1, -1, x);
        }
        for (int64_t j = 0; j < lt1; j++) {
            float x = global ? wt[j] * sgn : 0.0f;
            set_val(0, -1, j - 1, x);
            set_val(2, -1, j - 1, x);
        }
        w1_s = cfg.aff_u_s + cfg.aff_v_s;  // AffineCost::w1 (solver.h:2872-2874)
        w1_t = cfg.aff_u_t + cfg.aff_v_t;
        ext_s = cfg.aff_u_s;
        ext_t = cfg.aff_u_t;
    }

    // ---- the accumulator (solver.h:1226-1427) ------------------------
    // A cell accumulator over B lanes: first push initialises, later pushes
    // replace on strict improvement (one) or append on equality (all).
    struct Acc {
        LaneSolver *s;
        int64_t c;  // cell index * B
        bool started = false;

        void first(const float *v, int32_t u0, int32_t v0) {
            for (int b = 0; b < B; b++) {
                s->val[c + b] = v[b];
                if (s->tbu) { s->tbu[c + b] = u0; s->tbv[c + b] = v0; }
                if (s->all) { s->tbn[c + b].clear(); s->tbn[c + b].push_back(Pt{u0, v0}); }
            }
            started = true;
        }
        void first_copy(const float *v, int64_t src) {
            for (int b = 0; b < B; b++) {
                s->val[c + b] = v[b];
                if (s->tbu) { s->tbu[c + b] = s->tbu[src + b]; s->tbv[c + b] = s->tbv[src + b]; }
                if (s->all) s->tbn[c + b] = s->tbn[src + b];
            }
            started = true;
        }
        void push(const float *v, int32_t u0, int32_t v0) {
            if (!started) { first(v, u0, v0); return; }
            for (int b = 0; b < B; b++) {
                const float cur = s->val[c + b];
                if (is_opt(s->cfg.direction, v[b], cur)) {
                    s->val[c + b] = v[b];
                    if (s->tbu) { s->tbu[c + b] = u0; s->tbv[c + b] = v0; }
                    if (s->all) { s->tbn[c + b].clear(); s->tbn[c + b].push_back(Pt{u0, v0}); }
                } else if (s->all && v[b] == cur) {
                    s->tbn[c + b].push_back(Pt{u0, v0});
                }
            }
        }
        void push_copy(const float *v, int64_t src) {
            if (!started) { first_copy(v, src); return; }
            for (int b = 0; b < B; b++) {
                const float cur = s->val[c + b];
                if (is_opt(s->cfg.direction, v[b], cur)) {
                    s->val[c + b] = v[b];
                    if (s->tbu) { s->tbu[c + b] = s->tbu[src + b]; s->tbv[c + b] = s->tbv[src + b]; }
                    if (s->all) s->tbn[c + b] = s->tbn[src + b];
                } else if (s->all && v[b] == cur) {
                    const std::vector<Pt> add = s->tbn[src + b];
                    for (const Pt &p : add) s->tbn[c + b].push_back(p);
                }
            }
        }
        // Value-only push used when no traceback is kept (Accumulator,
        // solver.h:1226-1309): xt::maximum/minimum keep the old value on ties.
    };

    // ---- DP solve over the batch's master shape -----------------------
    // pairwise(i, j, out[B]) gives S(i,j) per lane (factory.h:573-576,
    // 624-630, 673-679).
    // Value-only accumulation (Accumulator, solver.h:1226-1309): lane-wise
    // cur = opt(x, cur) keeping the old value on ties.
    // Lane vectors as in the reference's tiny_vector/xsimd batches
    // (solver.h:376-422): one SIMD register of B float32 lanes.
    typedef float VecB __attribute__((vector_size(sizeof(float) * B)));
    static inline VecB ld(const float *p) { VecB v; std::memcpy(&v, p, sizeof(VecB)); return v; }
    static inline void st(float *p, VecB v) { std::memcpy(p, &v, sizeof(VecB)); }
    static inline VecB vopt(int dir, VecB x, VecB cur) {
        // strict improvement, old value kept on ties (xt::maximum / minimum)
        return dir == D_MAXIMIZE ? (x > cur ? x : cur) : (x < cur ? x : cur);
    }

    template <typename Pairwise>
    void solve_values(const Pairwise &pairwise, int64_t mls, int64_t mlt) {
        alignas(64) float S[B];
        const int dir = cfg.direction;
        const bool local = cfg.locality == L_LOCAL;
        const VecB zero = VecB{} * 0.0f;
        if (cfg.gap_kind == GAP_LINEAR) {
            const VecB gs = zero + cfg.lin_s * sgn, gt = zero + cfg.lin_t * sgn;
            for (int64_t u = 0; u < mls; u++) {
                for (int64_t v = 0; v < mlt; v++) {
                    pairwise(u, v, S);
                    const VecB x = ld(&val[idx(0, u - 1, v - 1) * B]) + ld(S);
                    VecB m = local ? vopt(dir, x, zero) : x;
                    m = vopt(dir, ld(&val[idx(0, u - 1, v) * B]) + gs, m);
                    m = vopt(dir, ld(&val[idx(0, u, v - 1) * B]) + gt, m);
                    st(&val[idx(0, u, v) * B], m);
                }
            }
            return;
        }
        const VecB os = zero + w1_s * sgn, es = zero + ext_s * sgn;
        const VecB ot = zero + w1_t * sgn, et = zero + ext_t * sgn;
        for (int64_t i = 0; i < mls; i++) {
            for (int64_t j = 0; j < mlt; j++) {
                pairwise(i, j, S);
                const VecB P = vopt(dir, ld(&val[idx(1, i - 1, j) * B]) + es, ld(&val[idx(0, i - 1, j) * B]) + os);
                const VecB Q = vopt(dir, ld(&val[idx(2, i, j - 1) * B]) + et, ld(&val[idx(0, i, j - 1) * B]) + ot);
                const VecB x = ld(&val[idx(0, i - 1, j - 1) * B]) + ld(S);
                VecB m = local ? vopt(dir, x, zero) : x;
                m = vopt(dir, P, m);
                m = vopt(dir, Q, m);
                st(&val[idx(1, i, j) * B], P);
                st(&val[idx(2, i, j) * B], Q);
                st(&val[idx(0, i, j) * B], m);
            }
        }
    }

    template <typename Pairwise>
    void solve(const Pairwise &pairwise, int64_t mls, int64_t mlt,
               const int64_t *len_s, const int64_t *len_t) {
        if (mls > cfg.lim_s || mlt > cfg.lim_t)
            throw std::runtime_error("exceeded_configured_length");  // solver.h:134-152
        for (int b = 0; b < B; b++) { cur_len_s[b] = len_s[b]; cur_len_t[b] = len_t[b]; }
        if (!trace && cfg.type == T_ALIGNMENT && cfg.gap_kind != GAP_GENERAL) {
            solve_values(pairwise, mls, mlt);
            return;
        }
        float S[B], x[B];
        const int dir = cfg.direction;
        if (cfg.type == T_DTW) {
            // DynamicTimeSolver::solve (solver.h:3213-3251)
            for (int64_t u = 0; u < mls; u++) {
                for (int64_t v = 0; v < mlt; v++) {
                    pairwise(u, v, S);
                    Acc acc{this, idx(0, u, v) * B};
                    const float *d = &val[idx(0, u - 1, v - 1) * B];
                    acc.push(d, int32_t(u - 1), int32_t(v - 1));
                    const float *up = &val[idx(0, u - 1, v) * B];
                    acc.push(up, int32_t(u - 1), int32_t(v));
                    const float *lf = &val[idx(0, u, v - 1) * B];
                    acc.push(lf, int32_t(u), int32_t(v - 1));
                    float *m = &val[idx(0, u, v) * B];
                    for (int b = 0; b < B; b++) m[b] += S[b];
                }
            }
            return;
        }
        const bool local = cfg.locality == L_LOCAL;
        const float zero[B] = {};
        if (cfg.gap_kind == GAP_LINEAR) {
            // LinearGapCostSolver::solve (solver.h:2825-2855)
            const float gs = cfg.lin_s * sgn, gt = cfg.lin_t * sgn;
            for (int64_t u = 0; u < mls; u++) {
                for (int64_t v = 0; v < mlt; v++) {
                    pairwise(u, v, S);
                    Acc acc{this, idx(0, u, v) * B};
                    if (local) acc.push(zero, no_tb, no_tb);  // Local::accumulate_to 2008-2011
                    const float *d = &val[idx(0, u - 1, v - 1) * B];
                    for (int b = 0; b < B; b++) x[b] = d[b] + S[b];
                    acc.push(x, int32_t(u - 1), int32_t(v - 1));
                    const float *up = &val[idx(0, u - 1, v) * B];
                    for (int b = 0; b < B; b++) x[b] = up[b] + gs;
                    acc.push(x, int32_t(u - 1), int32_t(v));
                    const float *lf = &val[idx(0, u, v - 1) * B];
                    for (int b = 0; b < B; b++) x[b] = lf[b] + gt;
                    acc.push(x, int32_t(u), int32_t(v - 1));
                }
            }
            return;
        }
        if (cfg.gap_kind == GAP_GENERAL) {
            // GeneralGapCostSolver::solve (solver.h:3145-3183)
            for (int64_t u = 0; u < mls; u++) {
                for (int64_t v = 0; v < mlt; v++) {
                    pairwise(u, v, S);
                    Acc acc{this, idx(0, u, v) * B};
                    if (local) acc.push(zero, no_tb, no_tb);
                    const float *d = &val[idx(0, u - 1, v - 1) * B];
                    for (int b = 0; b < B; b++) x[b] = d[b] + S[b];
                    acc.push(x, int32_t(u - 1), int32_t(v - 1));
                    for (int64_t k = -1; k < u; k++) {
                        const float *c = &val[idx(0, k, v) * B];
                        const float w = gen_s[size_t(u - k)] * sgn;
                        for (int b = 0; b < B; b++) x[b] = c[b] + w;
                        acc.push(x, int32_t(k), int32_t(v));
                    }
                    for (int64_t k = -1; k < v; k++) {
                        const float *c = &val[idx(0, u, k) * B];
                        const float w = gen_t[size_t(v - k)] * sgn;
                        for (int b = 0; b < B; b++) x[b] = c[b] + w;
                        acc.push(x, int32_t(u), int32_t(k));
                    }
                }
            }
            return;
        }
        // AffineGapCostSolver::solve (solver.h:2970-3036)
        const float os = w1_s * sgn, es = ext_s * sgn, ot = w1_t * sgn, et = ext_t * sgn;
        for (int64_t i = 0; i < mls; i++) {
            for (int64_t j = 0; j < mlt; j++) {
                pairwise(i, j, S);
                {   // Gotoh (4): P
                    Acc acc{this, idx(1, i, j) * B};
                    const float *d = &val[idx(0, i - 1, j) * B];
                    for (int b = 0; b < B; b++) x[b] = d[b] + os;
                    acc.push(x, int32_t(i - 1), int32_t(j));
                    const int64_t src = idx(1, i - 1, j) * B;
                    for (int b = 0; b < B; b++) x[b] = val[src + b] + es;
                    acc.push_copy(x, src);
                }
                {   // Gotoh (5): Q
                    Acc acc{this, idx(2, i, j) * B};
                    const float *d = &val[idx(0, i, j - 1) * B];
                    for (int b = 0; b < B; b++) x[b] = d[b] + ot;
                    acc.push(x, int32_t(i), int32_t(j - 1));
                    const int64_t src = idx(2, i, j - 1) * B;
                    for (int b = 0; b < B; b++) x[b] = val[src + b] + et;
                    acc.push_copy(x, src);
                }
                {   // Gotoh (1): D
                    Acc acc{this, idx(0, i, j) * B};
                    if (local) acc.push(zero, no_tb, no_tb);
                    const float *d = &val[idx(0, i - 1, j - 1) * B];
                    for (int b = 0; b < B; b++) x[b] = d[b] + S[b];
                    acc.push(x, int32_t(i - 1), int32_t(j - 1));
                    const int64_t sp = idx(1, i, j) * B;
                    for (int b = 0; b < B; b++) x[b] = val[sp + b];
                    acc.push_copy(x, sp);
                    const int64_t sq = idx(2, i, j) * B;
                    for (int b = 0; b < B; b++) x[b] = val[sq + b];
                    acc.push_copy(x, sq);
                }
            }
        }
    }

    float D(int lane, int64_t i, int64_t j) const { return val[idx(0, i, j) * B + lane]; }

    // ---- seeds (solver.h:1937-2327) -----------------------------------
    // Returns the seed stack in push order (the iterator pops from the back).
    std::vector<Pt> seeds(int lane) const {
        const int64_t n = cur_len_s[lane], m = cur_len_t[lane];
        const int dir = cfg.direction;
        std::vector<Pt> r;
        if (cfg.locality == L_GLOBAL) {  // Global::TracebackSeeds 2179-2187
            r.push_back(Pt{int32_t(n - 1), int32_t(m - 1)});
            return r;
        }
        auto optima = [&](auto &&visit) {  // Optima (1962-1974)
            float best = worst_val(dir);
            int32_t bi = 0, bj = 0;
            visit([&](int64_t i, int64_t j) {
                const float v = D(lane, i, j);
                if (is_opt(dir, v, best)) { best = v; bi = int32_t(i); bj = int32_t(j); }
            });
            if (is_opt(dir, best, worst_val(dir))) r.push_back(Pt{bi, bj});
        };
        if (cfg.locality == L_SEMIGLOBAL) {  // Semiglobal::TracebackSeeds 2273-2293
            optima([&](auto &&add) {
                for (int64_t j = 0; j < m; j++) add(n - 1, j);
                for (int64_t i = 0; i < n; i++) add(i, m - 1);
            });
            return r;
        }
        if (!all) {  // Local one-optimal 2042-2057
            optima([&](auto &&add) {
                for (int64_t i = n - 1; i >= 0; i--)
                    for (int64_t j = m - 1; j >= 0; j--) add(i, j);
            });
            return r;
        }
        // Local all-optimal 2060-2092
        float best = 0.0f;
        for (int64_t i = n - 1; i >= 0; i--)
            for (int64_t j = m - 1; j >= 0; j--) {
                const float v = D(lane, i, j);
                if (is_opt(dir, v, best)) best = v;
            }
        for (int64_t i = n - 1; i >= 0; i--)
            for (int64_t j = m - 1; j >= 0; j--) {
                const float v = D(lane, i, j);
                if (v == best && is_opt(dir, v, 0.0f)) r.push_back(Pt{int32_t(i), int32_t(j)});
            }
        return r;
    }

    // Strategy::continue_traceback_1/2 (Local 2117-2127, Global 2207-2217).
    bool cont(int lane, int32_t u, int32_t v) const {
        if (!(u >= 0 && v >= 0)) return false;
        if (cfg.locality == L_LOCAL) return is_opt(cfg.direction, D(lane, u, v), 0.0f);
        return true;
    }

    static void compress(const std::vector<Pt> &path, std::vector<Pt> &edges) {
        // path_compressor (solver.h:809-846)
        edges.clear();
        bool empty = true;
        int32_t lu = 0, lv = 0;
        for (const Pt &p : path) {
            if (empty) {
                empty = false;
            } else if (p.u != lu && p.v != lv) {
                if (lu >= 0 && lv >= 0) edges.push_back(Pt{lu, lv});
            }
            lu = p.u;
            lv = p.v;
        }
    }

    // One-optimal iterator (TracebackIterators<false>, solver.h:1466-1536).
    LaneResult one(int lane) const {
        LaneResult res;
        std::vector<Pt> sd = seeds(lane);
        if (sd.empty()) return res;
        Pt p = sd.back();
        res.has_seed = true;
        res.score = D(lane, p.u, p.v);
        if (!trace) return res;
        int32_t u = p.u, v = p.v;
        res.path.push_back(Pt{u, v});
        while (cont(lane, u, v)) {
            const int64_t c = idx(0, u, v) * B + lane;
            const int32_t nu = tbu[c], nv = tbv[c];
            u = nu;
            v = nv;
            res.path.push_back(Pt{u, v});
        }
        res.has_alignment = true;
        compress(res.path, res.edges);
        return res;
    }
};

// All-optimal DFS with optional dedupe (TracebackIterators<true>,
// solver.h:1568-1743), over lane `lane` of a solver with goal == all.
template <int B>
struct AllIterator {
    const LaneSolver<B> *s;
    int lane;
    struct Entry {
        float path_val;
        Pt cur;
        int64_t path_len;
    };
    std::vector<Entry> stack;
    std::vector<Pt> path;
    std::set<std::vector<std::pair<int32_t, int32_t>>> emitted;

    AllIterator(const LaneSolver<B> *solver, int l) : s(solver), lane(l) {
        for (const Pt &p : s->seeds(lane)) stack.push_back(Entry{s->D(lane, p.u, p.v), p, 0});
    }

    bool next(LaneResult &out) {
        while (!stack.empty()) {
            Entry e = stack.back();
            stack.pop_back();
            if (e.path_len == 0) path.clear();
            else path.resize(size_t(e.path_len));
            path.push_back(e.cur);
            if (s->cont(lane, e.cur.u, e.cur.v)) {
                const auto &t = s->tbn[s->idx(0, e.cur.u, e.cur.v) * B + lane];
                const int64_t ps = int64_t(path.size());
                if (!t.empty()) {
                    for (const Pt &q : t) stack.push_back(Entry{e.path_val, q, ps});
                } else {
                    stack.push_back(Entry{e.path_val, Pt{s->no_tb, s->no_tb}, ps});
                }
            } else {
                std::vector<Pt> edges;
                LaneSolver<B>::compress(path, edges);
                if (s->cfg.remove_dup) {
                    std::vector<std::pair<int32_t, int32_t>> key;
                    for (const Pt &q : edges) key.emplace_back(q.u, q.v);
                    if (!emitted.insert(key).second) continue;
                }
                out = LaneResult();
                out.has_seed = true;
                out.score = e.path_val;
                out.path = path;
                out.has_alignment = true;
                out.edges = edges;
                return true;
            }
        }
        return false;
    }
};

thread_local std::string g_err;

}  // namespace

// ---------------------------------------------------------------------------
// C interface used by the tests (ctypes): one problem per call, dense S.
// ---------------------------------------------------------------------------

struct oc_handle {
    std::unique_ptr<LaneSolver<1>> solver;
    LaneResult one;
    std::unique_ptr<AllIterator<1>> it;
    int64_t len_s = 0, len_t = 0;
};

extern "C" {

const char *oc_last_error(void) { return g_err.c_str(); }

// Build a solver with the configured limits, run the DP over a dense
// similarity matrix sim[len_s][len_t], and prepare the one-optimal result
// (or the all-optimal iterator).  Returns NULL on error.
oc_handle *oc_solve_dense(const oc_config *cfg, const float *sim, int64_t len_s, int64_t len_t) {
    try {
        auto h = std::make_unique<oc_handle>();
        h->solver = std::make_unique<LaneSolver<1>>(*cfg);
        h->len_s = len_s;
        h->len_t = len_t;
        const int64_t ls[1] = {len_s}, lt[1] = {len_t};
        h->solver->solve(
            [&](int64_t i, int64_t j, float *out) { out[0] = sim[i * len_t + j]; },
            len_s, len_t, ls, lt);
        if (h->solver->all) h->it = std::make_unique<AllIterator<1>>(h->solver.get(), 0);
        else h->one = h->solver->one(0);
        return h.release();
    } catch (const std::exception &e) {
        g_err = e.what();
        return nullptr;
    }
}

void oc_free(oc_handle *h) { delete h; }

// Score (Solver::score, solver.h:2593-2608): worst value when no seed.
float oc_score(const oc_handle *h) {
    return h->one.has_seed ? h->one.score : worst_val(h->solver->cfg.direction);
}
int32_t oc_has_seed(const oc_handle *h) { return h->one.has_seed ? 1 : 0; }

int64_t oc_path_len(const oc_handle *h) { return int64_t(h->one.path.size()); }
void oc_path(const oc_handle *h, int32_t *out) {
    for (size_t k = 0; k < h->one.path.size(); k++) {
        out[2 * k] = h->one.path[k].u;
        out[2 * k + 1] = h->one.path[k].v;
    }
}
int64_t oc_edges_len(const oc_handle *h) { return int64_t(h->one.edges.size()); }
void oc_edges(const oc_handle *h, int32_t *out) {
    for (size_t k = 0; k < h->one.edges.size(); k++) {
        out[2 * k] = h->one.edges[k].u;
        out[2 * k + 1] = h->one.edges[k].v;
    }
}

// All-optimal iteration: returns 1 and fills the current item, 0 at the end.
int32_t oc_all_next(oc_handle *h) {
    if (!h->it) return 0;
    LaneResult r;
    if (!h->it->next(r)) return 0;
    h->one = r;
    return 1;
}

// Solution data (Solution::set_values/set_traceback, solver.h:2355-2395):
// values [layers, len_s+1, len_t+1] and first-edge traceback (u, v).
int32_t oc_layers(const oc_handle *h) { return h->solver->layers; }
void oc_values(const oc_handle *h, int32_t layer, float *out) {
    const auto &s = *h->solver;
    for (int64_t i = 0; i <= h->len_s; i++)
        for (int64_t j = 0; j <= h->len_t; j++)
            out[i * (h->len_t + 1) + j] = s.val[s.idx(layer, i - 1, j - 1)];
}
void oc_traceback(const oc_handle *h, int32_t layer, int32_t *out) {
    const auto &s = *h->solver;
    for (int64_t i = 0; i <= h->len_s; i++)
        for (int64_t j = 0; j <= h->len_t; j++) {
            const int64_t c = s.idx(layer, i - 1, j - 1);
            int32_t u = s.no_tb, v = s.no_tb;
            if (s.tbu) { u = s.tbu[c]; v = s.tbv[c]; }
            else if (s.all && !s.tbn[c].empty()) { u = s.tbn[c][0].u; v = s.tbn[c][0].v; }
            out[2 * (i * (h->len_t + 1) + j)] = u;
            out[2 * (i * (h->len_t + 1) + j) + 1] = v;
        }
}

// Solution::traceback_as_edges (solver.h:2431-2480): per layer, cells row by
// row (border included), each cell's traceback entries in order:
// ((i-1, j-1), (u, v)).  traceback_1 cells always hold one entry.
static void tb_edges(const oc_handle *h, int32_t layer, int32_t *out, int64_t *count) {
    const auto &s = *h->solver;
    int64_t n = 0;
    auto emit = [&](int64_t i, int64_t j, int32_t u, int32_t v) {
        if (out) {
            out[4 * n] = int32_t(i - 1);
            out[4 * n + 1] = int32_t(j - 1);
            out[4 * n + 2] = u;
            out[4 * n + 3] = v;
        }
        n++;
    };
    for (int64_t i = 0; i <= h->len_s; i++)
        for (int64_t j = 0; j <= h->len_t; j++) {
            const int64_t c = s.idx(layer, i - 1, j - 1);
            if (s.tbu) emit(i, j, s.tbu[c], s.tbv[c]);
            else if (s.all) for (const Pt &p : s.tbn[c]) emit(i, j, p.u, p.v);
        }
    *count = n;
}
int64_t oc_tb_edges_len(const oc_handle *h, int32_t layer) {
    int64_t n = 0;
    tb_edges(h, layer, nullptr, &n);
    return n;
}
void oc_tb_edges(const oc_handle *h, int32_t layer, int32_t *out) {
    int64_t n = 0;
    tb_edges(h, layer, out, &n);
}

// ---------------------------------------------------------------------------
// CPU baseline: the reference's batched score path (B lanes, Morton-free
// greedy grouping done by the caller) over indexed sequences, on T threads.
// Each thread owns one LaneSolver<16>, as one reference Solver per thread.
// codes_s/codes_t are uint8 per position; off_* index into them.
// Returns elapsed seconds of the solve loop; scores are written per problem.
// ---------------------------------------------------------------------------
double oc_baseline_indexed_score(const oc_config *cfg, const uint8_t *codes_s, const int64_t *off_s,
                                 const int64_t *len_s, const uint8_t *codes_t, const int64_t *off_t,
                                 const int64_t *len_t, const float *sim, int64_t K,
                                 const int32_t *batch_members, const int64_t *batch_off,
                                 int64_t n_batches, int32_t threads, float *scores) {
    constexpr int BL = 16;
    std::atomic<int64_t> next{0};
    std::atomic<int> failed{0};
    auto t0 = std::chrono::steady_clock::now();
    auto worker = [&]() {
        try {
            std::unique_ptr<LaneSolver<BL>> solver;
            while (true) {
                const int64_t bi = next.fetch_add(1);
                if (bi >= n_batches) break;
                const int32_t *mem = batch_members + batch_off[bi];
                const int64_t cnt = batch_off[bi + 1] - batch_off[bi];
                int64_t ls[BL], lt[BL], os[BL], ot[BL];
                int64_t mls = 0, mlt = 0;
                for (int b = 0; b < BL; b++) {
                    const int32_t p = mem[b < cnt ? b : 0];
                    ls[b] = len_s[p]; lt[b] = len_t[p]; os[b] = off_s[p]; ot[b] = off_t[p];
                    mls = std::max(mls, ls[b]);
                    mlt = std::max(mlt, lt[b]);
                }
                if (!solver) solver = std::make_unique<LaneSolver<BL>>(*cfg);
                // indexed_matrix_form::operator() (factory.h:624-630); positions
                // past a lane's own length read code 0 (result-neutral).
                solver->solve(
                    [&](int64_t i, int64_t j, float *out) {
                        for (int b = 0; b < BL; b++) {
                            const int a = i < ls[b] ? codes_s[os[b] + i] : 0;
                            const int c = j < lt[b] ? codes_t[ot[b] + j] : 0;
                            out[b] = sim[a * K + c];
                        }
                    },
                    mls, mlt, ls, lt);
                for (int b = 0; b < cnt; b++) {
                    LaneResult r = solver->one(b);
                    scores[mem[b]] = r.has_seed ? r.score : worst_val(cfg->direction);
                }
            }
        } catch (const std::exception &e) {
            g_err = e.what();
            failed = 1;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) pool.emplace_back(worker);
    for (auto &th : pool) th.join();
    auto t1 = std::chrono::steady_clock::now();
    if (failed) return -1.0;
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
