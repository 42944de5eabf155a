// pa_internal.h -- data layouts shared by the HIP kernels and the host library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pa {

enum { LOC_LOCAL = 0, LOC_GLOBAL = 1, LOC_SEMI = 2 };
enum { DIR_MIN = 0, DIR_MAX = 1 };
enum { GAP_LIN = 0, GAP_AFF = 1, GAP_DTW = 2, GAP_GEN = 3 };
// Similarity sources.  F_TABLE reads a (K+1)x(K+1) table (row = s symbol,
// last symbol = pad whose entries are the worst value) staged in LDS with
// uint8 codes; F_TABLE_G reads it from global memory (L2-resident) with
// uint16 codes, for tables too large for LDS (more than 128 symbols).
// F_DROWS is F_DENSE with every problem's rows contiguous (column stride 1)
// and at least DROWS_MIN_LT columns: each lane stages windows of its rows
// through LDS with 16-byte loads (a launch-level choice; the payload and
// everything outside the DP kernel treat it as F_DENSE).
enum { F_DENSE = 0, F_TABLE = 1, F_BINARY = 2, F_SERIES = 3, F_TABLE_G = 4, F_DROWS = 5 };
constexpr int DROWS_WIN = 8;       // columns per staged window (== the strip height R)
constexpr int DROWS_MIN_LT = 8;
constexpr int DROWS_LANE_FLOATS = 2 * DROWS_WIN * 8 + 16;  // two windows of R = 8 rows, 8 floats of slack each side
constexpr size_t DROWS_LDS = size_t(256) * DROWS_LANE_FLOATS * 4;
// What the forward pass keeps: score only, one traceback direction per cell
// (2 bits linear/DTW, 4 bits affine) or the full tie set (8 bits per cell).
enum { TR_SCORE = 0, TR_ONE = 1, TR_ALL = 2 };

constexpr int WAVE = 64;
constexpr int MAX_LDS_TABLE = 129;  // K + 1 pad symbol
constexpr int32_t NO_SEED = INT32_MIN;
constexpr int64_t NO_TB_OFF = INT64_MIN;  // Prob.tb_off of a problem whose traceback is not kept

// One problem.  Offsets are element offsets into the form's payload.
struct Prob {
    int64_t off_s;   // codes / series rows of s, or dense base
    int64_t off_t;   // codes / series rows of t
    int64_t tb_off;  // first 32-bit word of this problem's traceback
    int64_t rb_off;  // float2 offset of this problem's strip-boundary row buffer
    int64_t dv_off;  // offset of this problem's cell values: layer-major, then row-major
    int32_t len_s, len_t;
    int32_t si, sj;  // dense strides
    int32_t out;     // output slot (caller's problem index)
    int32_t L;       // lanes cooperating on this problem
    int32_t R;       // rows per lane
    int32_t band;    // band mode: first strip slot (progress, partials); -1 otherwise
    int32_t pack;    // long mode: traceback codes packed along the steps, long_pack(bits) per word; 0 otherwise
};

// One wavefront's work: 64/L problems (segment k <- probs[first + k]).
struct Group {
    int32_t first, count;
    int32_t L, lgL;      // lanes per problem and its log2
    int32_t max_lt;      // over members: step loop bound is max_lt + L - 1
    int32_t max_nst;     // over members: strips of L*R rows
    int32_t R;           // rows per lane (kernels built for several R dispatch on it)
    int32_t strip;       // band mode: the one strip this group computes; -1: all strips
    int64_t cost;        // scheduling cost
};

// Band mode (long problems, few of them): each strip of L*R rows is its own
// group, so the strips of one problem run concurrently on different
// wavefronts.  Strip st reads the bottom row of strip st-1 from its row
// buffer as soon as it is published: progress[band + st - 1] counts the
// published columns (release/acquire at agent scope).  Groups are queued in
// strip order, so a strip only ever waits for a strip taken earlier by a
// running wavefront.  Each strip leaves its share of the seed search in
// partials[band + st]; pa_band_combine_kernel folds them in strip order.
struct BandPartial {
    float v1;        // local: best; global: D(ls-1, lt-1); semiglobal: last-column best
    float v2;        // semiglobal: last-row best
    int32_t i1, j1;  // local: best cell; semiglobal: last-column row
    int32_t j2;      // semiglobal: last-row column
    int32_t pad;
};
#ifndef PA_BAND_PUB
#define PA_BAND_PUB 32
#endif
constexpr int BAND_PUB = PA_BAND_PUB;        // columns per progress publication
constexpr uint32_t BAND_SPIN_LIMIT = 1u << 20;  // polls before a wait gives up (about a second)

// Long-problem mode (pa_long_kernel): a workgroup of LONG_WAVES compute
// waves and one loader wave computes a band of LONG_WAVES strips of
// 64 x LONG_R rows; the waves run in phases of LONG_T steps between
// workgroup barriers, long_dw phases apart, handing bottom rows down through
// LDS rings of LONG_RING columns; the loader reads the band above LONG_K
// phases ahead (one: each band starts a phase's columns sooner; the 5045 x
// 10000 pair's kernels measured 2 % faster than with two).
#ifndef PA_LONG_WAVES
#define PA_LONG_WAVES 4
#endif
#ifndef PA_LONG_T
#define PA_LONG_T 16
#endif
#ifndef PA_LONG_K
#define PA_LONG_K 1
#endif
constexpr int LONG_WAVES = PA_LONG_WAVES;
constexpr int LONG_R = 1;
constexpr int LONG_T = PA_LONG_T;
constexpr int LONG_K = PA_LONG_K;
constexpr int LONG_RING = 64;
// Wave w runs b_w = -w mod LONG_T steps behind the phase grid, so that the
// LONG_T columns its lane 63 finishes in a phase and the LONG_T columns the
// next wave's lane 0 takes in a phase each fill one run of ring slots.  A
// wave reads its local phase k's entries at the start of phase k (the table
// form also the first of phase k + 1); the wave above wrote them in its
// local phase k + 64 / LONG_T (k + 64 / LONG_T - 1 above wave 1), which must
// end before that read: long_dw >= 64 / LONG_T + 1 (+ 1 for the table
// form).  It rewrites their slots 64 / LONG_T phases later, which must come
// after the read: long_dw < 128 / LONG_T - 1.
constexpr int long_dw(int form) { return 64 / LONG_T + (form == F_TABLE ? 2 : 1); }
static_assert((LONG_T == 8 || LONG_T == 16) && LONG_K * LONG_T <= 64, "long-mode phase shape");
static_assert(long_dw(F_TABLE) < 128 / LONG_T - 1 && LONG_WAVES <= LONG_T, "long-mode ring timing");
// A long problem's traceback codes (bits per cell: 4 affine, 2 linear),
// strip by strip, packed along the steps: lane l of the strip of wave w
// keeps the code of its step s in word (s + b_w) / long_pack(bits) at bit
// bits * ((s + b_w) % long_pack(bits)); a word index's 64 lanes are
// consecutive words, a strip's word indexes consecutive (a phase's codes
// are whole words, stored once per word).
constexpr int long_pack(int bits) { return 32 / bits < LONG_T ? 32 / bits : LONG_T; }
// words per lane and strip of steps_per_strip (len_t + 63) steps: the phases
// of the wave with the largest b_w
constexpr int64_t long_strip_words(int64_t steps_per_strip, int bits) {
    return (steps_per_strip + 2 * LONG_T - 2) / LONG_T * (LONG_T / long_pack(bits));
}
constexpr uint32_t LONG_SPIN_LIMIT = 1u << 20;  // re-reads before a wait gives up (about a second)

struct LongArgs {
    const int32_t *idx;    // the long problems' indices into KArgs::probs
    int32_t n;             // long problems
    int32_t total_bands;   // bands over all of them (tickets)
    uint32_t *ticket;      // band ticket counter (zeroed before launch)
};

struct KArgs {
    const Prob *probs;
    const Group *groups;
    int32_t n_groups;
    uint32_t *queue;  // work-queue ticket counter (zeroed before launch)
    const void *codes_s;
    const void *codes_t;
    const float *table;
    int32_t K1;       // table stride = K + 1
    int32_t dim;
    float eq, ne;
    const float *dense;
    float2 *rowbuf;
    const float *border_s;  // border_s[k] = D(k-1, -1), k = 0..lim_s
    const float *border_t;  // border_t[k] = D(-1, k-1), k = 0..lim_t
    float go_s, ge_s, go_t, ge_t;  // signed penalties: w1*sgn, extend*sgn (linear: g*sgn in go)
    float *scores;
    int32_t *seeds;    // [2*out]: seed (i, j), i = NO_SEED when there is none
    uint32_t *tb;      // traceback words
    float *dvals;      // cell values (null: none kept): D, then P and Q when val_layers == 3
    int32_t val_layers;
    uint32_t *progress;     // band mode: published columns per strip
    BandPartial *partials;  // band mode: seed search per strip
    uint32_t *band_err;     // band mode: set when a wait gave up
    int32_t band_sync;      // band mode: 1 = strips of a problem run concurrently (progress
                            // counters hand rows down); 0 = every row above is complete
    // SAFE_FLOATS readable floats: [0] = -inf, [1] = +inf, the rest 0; the
    // address a lane or row with nothing to read loads from
    const float *safe;
    // SINK_WORDS writable words: where a lane with nothing to store stores
    uint32_t *sink;
    // series form (DTW): a word that holds 1 when every series value is
    // finite, so that the cells' costs and values are never NaN or -0 and
    // min / max folds return the strict folds' bits (null: strict folds)
    const uint32_t *finite;
};
constexpr int SAFE_FLOATS = 16;
constexpr int SINK_WORDS = 16;

// R = 0 selects the kernel that serves every strip height in FAST_RS from
// one queue (the group carries its R).
#ifndef PA_FAST_RS_LIST
#define PA_FAST_RS_LIST 8, 10, 12, 14, 16, 20, 24, 28, 32
#endif
constexpr int FAST_RS[] = {PA_FAST_RS_LIST};

struct LaunchSpec {
    int loc, dir, gap, form, tr, R;
    bool fastmax;   // v_max_f32 is exact (local maximize score, NaN-free table)
    int grid;       // workgroups of 256 threads (persistent)
    size_t lds_bytes;
    int pair;       // fast kernel: LDS table of row-symbol pairs (K1^2 * (K1-1) float2)
    int block;      // threads per workgroup
};
// largest pair table the fast kernel stages (two workgroups per CU still fit)
constexpr size_t PAIR_LDS_MAX = 80 * 1024;
// fast kernel workgroup: its twelve waves share one LDS table, and at up to
// 168 VGPRs three waves per SIMD stay resident (768 measured 12% faster than
// 256-thread groups, which the 70 KB pair table held to two waves per SIMD)
#ifndef PA_FAST_BLOCK
#define PA_FAST_BLOCK 768
#endif
constexpr int FAST_BLOCK = PA_FAST_BLOCK;

hipError_t launch_solve(const LaunchSpec &spec, const KArgs &args, hipStream_t stream);
// Long-problem mode (forms F_TABLE / F_BINARY, score or one traceback,
// linear or affine gaps); `grid` workgroups take the bands in order.
// fold_max: max / min folds instead of strict compare-and-select, traceback
// codes from equality tests (the caller has checked that they return the
// same bits).
hipError_t launch_long(int loc, int dir, int gap, int form, int tr, bool fold_max, const KArgs &args,
                       const LongArgs &la, int grid, hipStream_t stream);
size_t long_lds_bytes(int form, int K1);
bool long_supported(int gap, int form, int tr);
hipError_t launch_walk(int gap, const Prob *probs, int64_t n, const uint32_t *tb, const int32_t *seeds,
                       int32_t *path, const int64_t *path_off, int32_t *path_len, hipStream_t stream);
hipError_t launch_zero_u32(uint32_t *p, int64_t n, hipStream_t stream);
// Folds band-mode strips' partials into scores and seeds (one thread per problem).
hipError_t launch_band_combine(int loc, int dir, int tr, const Prob *probs, const int32_t *band_probs, int n,
                               const BandPartial *partials, float *scores, int32_t *seeds, hipStream_t stream);

// Backward pass over one-optimal tracebacks (one thread per problem).
struct BwArgs {
    const Prob *probs;
    int64_t n;
    const uint32_t *tb;
    const int32_t *seeds;
    int global;          // count the border gap run that ends a global path
    float gsign;         // d score / d gap cost: -1 maximising, +1 minimising
    const float *grad_scores;
    float *grad_dense;   // dense form, payload layout
    const float *xs, *xt;
    float *grad_xs, *grad_xt;
    float *grad_gap;     // [2 * out]
};
hipError_t launch_backward(int gap, int form, int dim, const BwArgs &args, hipStream_t stream);

// Checkpointed backward: one problem's walk continued strip by strip, last
// strip first, each strip's traceback recomputed just before its walk.
struct StripWalk {
    int32_t i, j;      // current cell
    int32_t mode;      // 0: at a cell; 1: scanning a vertical gap run up column j
    int32_t pu, pv;    // the cell the vertical run started from
    int32_t cur_i, cur_j;  // DTW series: rows / columns of the running sums
    int32_t done;
    float opens, extends;
    float gx[4], gy[4];
};
// strip_of[q]: the strip problem q walks in this pass (-1: none)
hipError_t launch_backward_strips(int gap, int form, int dim, const BwArgs &args, StripWalk *state,
                                  const int32_t *strip_of, hipStream_t stream);
hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t value, hipStream_t stream);
// *flag = 1 when the nx floats at x and the ny floats at y are all finite, else 0
hipError_t launch_all_finite(const float *x, int64_t nx, const float *y, int64_t ny, uint32_t *flag,
                             hipStream_t stream);
// Raw symbols (width 1 or 4 bytes) to codes: uint8 or uint16 (out_width 1,
// 2) through lut, or uint32 values (out_width 4) when lut is null; *bad gets
// the smallest position without a code.
hipError_t launch_translate(const void *raw, int width, int64_t count, const uint8_t *lut, int32_t lut_len,
                            void *codes, int out_width, unsigned long long *bad, int cus, hipStream_t stream);
// General gap costs: one workgroup per problem.  At dvals + dv_off it writes
// four ls x lt planes: D, D transposed, S and the traceback code of every
// cell (WSB_* below); then the seed and score, and with `path` set the
// one-optimal path (as launch_walk writes it).  ws / wt hold w(k) * sgn for
// k = 0 .. lim.  threads: workgroup size, a multiple of 64 up to 1024.
// Traceback code of cell (u, v): the first candidate in push order that
// holds the cell's value: 0 the local zero, 1 the diagonal, 2 + (k + 1) the
// cell (k, v) above for k = -1 .. u-1, u + 3 + (k + 1) the cell (u, k) to
// the left for k = -1 .. v-1.
constexpr int WSB_PLANES = 4;
hipError_t launch_wsb(int loc, int dir, int form, const KArgs &args, int64_t n, const float *ws, const float *wt,
                      int32_t *path, const int64_t *path_off, int32_t *path_len, int threads,
                      hipStream_t stream);
bool supported_R(int form, int tr, bool fastmax, int R);
// The allocation-free entry point's device memory: a file-scope device array
// of NOALLOC_SCRATCH_BYTES per GPU.  noalloc_scratch gives the current
// device's copy (looked up once per device) and its CU count.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
hipError_t noalloc_scratch(void **ptr, int *cus);
// strip height of band-mode strips (long problems): 4 where the general
// kernel is built for it (more, shorter strips on concurrent wavefronts), else 8
int band_R(int form, int tr);
// resident waves per CU for this instantiation (dim: series dimensions, selects the kernel)
int occupancy_waves(const LaunchSpec &spec, int dim);

// Traceback bits per cell and 32-bit words per lane per step.
inline int tb_bits(int gap, int tr) { return tr == TR_ALL ? 8 : (tr == TR_ONE ? (gap == GAP_AFF ? 4 : 2) : 0); }
inline int tb_words(int gap, int tr, int R) { return (tb_bits(gap, tr) * R + 31) / 32; }

}  // namespace pa
