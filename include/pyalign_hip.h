/*
 * pyalign_hip.h -- C ABI of libpyalign_hip.so, the MI355X (gfx950) DP
 * alignment solver.
 *
 * This boundary replaces the reference's pybind11 extension module
 * pyalign.algorithm.<variant>.algorithm (pyalign/algorithm/module.cpp:1-18,
 * pyalign/algorithm/pyalign.h:12-102).  Every entry point below names the
 * reference interface it stands in for.  Plain pointers and sizes only; all
 * pointers are host pointers unless the name says _device.
 *
 * Every function returns a status code (PA_OK on success) and leaves a
 * thread-local message for pa_last_error().  PA_EINVAL maps to Python's
 * ValueError (std::invalid_argument in the reference), PA_ERUNTIME,
 * PA_ELENGTH and PA_EHIP to RuntimeError.
 *
 * Threading (as the reference, factory.h:705-792): one solver is not
 * re-entrant; distinct solvers are independent.  Each solver owns one HIP
 * stream and its device buffers.
 */
#ifndef PYALIGN_HIP_H
#define PYALIGN_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PA_OK = 0,
    PA_EINVAL = 1,   /* std::invalid_argument (factory.h:387-394, 526-533, 606-611) */
    PA_ERUNTIME = 2, /* std::runtime_error */
    PA_ELENGTH = 3,  /* exceeded_configured/implementation_length (solver.h:94-152) */
    PA_EHIP = 4      /* a HIP runtime call failed */
};

/* enums::Type, Direction, Detail, Count, Locality (pyalign/algorithm/enum.h:6-31) */
enum { PA_TYPE_ALIGNMENT = 0, PA_TYPE_DTW = 1 };
enum { PA_MINIMIZE = 0, PA_MAXIMIZE = 1 };
enum { PA_DETAIL_SCORE = 0, PA_DETAIL_ALIGNMENT = 1, PA_DETAIL_SOLUTION = 2 };
enum { PA_COUNT_ONE = 0, PA_COUNT_ALL = 1 };
enum { PA_LOCAL = 0, PA_GLOBAL = 1, PA_SEMIGLOBAL = 2 };

/* Gap cost kinds after GapCost.to_special_case() (pyalign/gaps.py:13-14,
 * 102-108, 140-143, 181-184; factory.h:33-64). */
enum { PA_GAP_NONE = 0, PA_GAP_LINEAR = 1, PA_GAP_AFFINE = 2, PA_GAP_GENERAL = 3 };

/* Problem forms (pyalign/problems/instance.py:7-10). */
enum { PA_FORM_DENSE = 0, PA_FORM_INDEXED = 1, PA_FORM_BINARY = 2, PA_FORM_SERIES = 3 };

/* What a solve produces: the goal of factory.h:1195-1239. */
enum { PA_GOAL_SCORE = 0, PA_GOAL_ONE = 1, PA_GOAL_ALL = 2 };

/*
 * The options dict of create_options() (options.h:16-44) with its gap costs
 * already mapped by GapCostOptions (factory.h:33-64).  gap_s/gap_t hold the
 * to_special_case() numbers in float64: linear {u, -}; affine {u, v} with
 * pyalign.gaps' w(k) = u + v*k, i.e. u = open - extend, v = extend.
 */
typedef struct pa_options_desc {
    int32_t solver_type; /* PA_TYPE_* */
    int32_t batch;       /* options['batch'] */
    int32_t direction;   /* PA_MINIMIZE / PA_MAXIMIZE */
    int32_t return_dup;  /* options['return_dup'] */
    int32_t detail;      /* PA_DETAIL_* (codomain.detail) */
    int32_t count;       /* PA_COUNT_* (codomain.count) */
    int32_t locality;    /* PA_LOCAL / PA_GLOBAL / PA_SEMIGLOBAL */
    int32_t gap_kind_s;  /* PA_GAP_* for s (rows) */
    int32_t gap_kind_t;  /* PA_GAP_* for t (columns) */
    int32_t index_bits;  /* 16 for m16 (Options<float,int16_t>), 32 for m32 */
    double gap_s[2];
    double gap_t[2];
} pa_options_desc;

typedef struct pa_options pa_options;
typedef struct pa_solver pa_solver;
typedef struct pa_result pa_result;

/* Thread-local text of the last error. */
const char *pa_last_error(void);

/* Number of HIP devices visible (0 when none); never fails. */
int pa_device_count(void);

/* create_options(dict) -> Options (pyalign.h:12-15, options.h:51-60). */
int pa_create_options(const pa_options_desc *desc, pa_options **out);
void pa_options_free(pa_options *opts);

/* create_solver(max_len_s, max_len_t, options) -> Solver
 * (factory.h:1268-1278).  The limits fix the solver's border vectors exactly
 * as the reference's MatrixFactory/solver constructors do. */
int pa_create_solver(int64_t max_len_s, int64_t max_len_t, const pa_options *opts,
                     pa_solver **out);
void pa_solver_free(pa_solver *solver);
/* General gap costs (both sides PA_GAP_GENERAL): the cost vectors w(k),
 * k = 0 .. max_len, that GeneralGapCostSolver's constructor takes from the
 * gap objects' costs(max_len + 1) (solver.h:3093-3125, factory.h:20-31);
 * lengths are checked as check_gap_tensor_shape does (solver.h:3047-3054).
 * Must be called before the first solve of such a solver. */
int pa_solver_set_gap_costs(pa_solver *solver, const float *w_s, int64_t n_s, const float *w_t,
                            int64_t n_t);

/* Linear or affine gap costs changed in place (no reference counterpart:
 * a reference solver's costs are fixed at creation, factory.h:33-64).
 * gap_s / gap_t hold the to_special_case() numbers as in pa_options_desc
 * (linear {g, -}, affine {u = open - extend, v = extend}); the solver keeps
 * its limits and device buffers and recomputes its border vectors exactly as
 * pa_create_solver does.  For differentiable training loops whose gap costs
 * are parameters: one solver serves every step.  Work already queued on the
 * solver's stream sees the old costs, later solves the new ones; a tape keeps
 * the costs of its forward. */
int pa_solver_set_gap_params(pa_solver *solver, const double *gap_s, const double *gap_t);

/* Solver.batch_size (pyalign.h:24, factory.h:422): lanes per call. */
int pa_solver_batch_size(const pa_solver *solver);

/* Solver.solve_for_{score,alignment,solution,...}: the dense MatrixForm
 * (factory.h:540-580).  sim is float32[Ls][Lt][B] (lane innermost), len is
 * int32[2][B] (row 0 = len_s, row 1 = len_t). */
int pa_solve_dense(pa_solver *solver, const float *sim, int64_t Ls, int64_t Lt, int64_t B,
                   const int32_t *len, pa_result **out);

/* Solver.solve_indexed_for_*: IndexedMatrixForm (factory.h:582-631).
 * a is uint32[B][La], b uint32[B][Lb], sim float32[K0][K1]. */
int pa_solve_indexed(pa_solver *solver, const uint32_t *a, int64_t La, const uint32_t *b,
                     int64_t Lb, int64_t B, const float *sim, int64_t K0, int64_t K1,
                     const int32_t *len, pa_result **out);

/* Solver.solve_binary_for_*: BinaryMatrixForm (factory.h:633-680). */
int pa_solve_binary(pa_solver *solver, const uint32_t *a, int64_t La, const uint32_t *b,
                    int64_t Lb, int64_t B, float eq, float ne, const int32_t *len,
                    pa_result **out);

/*
 * Ragged batch: N problems of individual lengths in one call, the whole bag
 * at once instead of the reference's per-B Python loop (solve.py:554-560).
 * Batching is result-neutral (each lane's cells depend only on its own
 * region, solver.h:2593-2705), so results equal the per-batch calls.
 *
 * Indexed form: codes_s/codes_t are symbol codes, uint8 (code_bytes 0 or 1)
 *   or uint16 (code_bytes 2); sequence k occupies
 *   codes_s[off_s[k] .. off_s[k]+len_s[k]).  The table is float32[K][K_t]
 *   (K_t = 0: square), s codes index its rows (< K), t codes its columns
 *   (< K_t), as indexed_matrix_form checks them (factory.h:602-614).  Any
 *   size up to 32,767 symbols per side: tables of up to 128 symbols are
 *   staged in LDS, larger ones are read from global memory (L2-resident)
 *   with uint16 codes (host uint8 codes are widened on the device).
 * Binary form: codes are uint32 values compared for equality.
 * Dense form: dense + dense_off[k] is problem k's row-major [len_s][len_t].
 * Series form (DTW): codes_s/codes_t are float32 series of `dim` columns,
 *   offsets in rows; cost = sum over dims of squared difference.
 * Lengths and offsets are host arrays; the payload pointers are host
 * pointers, or device pointers when on_device = 1 (inputs resident in HBM).
 */
typedef struct pa_ragged {
    int64_t n;
    int32_t form;       /* PA_FORM_* */
    int32_t on_device;  /* payload pointers are device pointers */
    const void *codes_s;
    const void *codes_t;
    const int64_t *off_s;
    const int64_t *off_t;
    const int32_t *len_s;
    const int32_t *len_t;
    const float *table; /* indexed: float32[K][K_t] (host) */
    int32_t K;          /* indexed: table rows (s symbols) */
    int32_t dim;        /* series dimension */
    float eq, ne;       /* binary */
    const float *dense; /* dense payload */
    const int64_t *dense_off;
    int64_t dense_ld;   /* dense row stride (0: rows of len_t, packed) */
    /* Problem encoding on the device (no reference counterpart: the
     * reference encodes each problem in Python, problems/factory.py:103-153,
     * 205-226).  raw_width 1 or 4: codes_s/codes_t hold raw symbols, 1-byte
     * (latin-1) or 4-byte (code points), offsets and lengths counted in
     * symbols.  Indexed form: lut (host, uint8[lut_len]) maps a symbol to its
     * code, 255 = not in the alphabet (the solve fails with PA_EINVAL "item
     * not in alphabet" before any DP work; see pa_solver_last_bad_symbol).
     * Binary form: the symbols are the values (lut = NULL).  0: codes given. */
    int32_t raw_width;
    int32_t lut_len;
    const uint8_t *lut;
    int32_t K_t;        /* indexed: table columns (t symbols); 0 = K */
    int32_t code_bytes; /* indexed codes given (raw_width 0): 1 = uint8 (0 too), 2 = uint16 */
} pa_ragged;

/* Scheduling of a ragged batch. */
enum {
    PA_SCHED_AUTO = 0,  /* cost-sorted work queue, per-bucket strip height (autotuned) */
    PA_SCHED_PLAIN = 1  /* the reference's Morton batches (instance.py:230-275) of 64,
                           one wavefront each, fixed strip height */
};

int pa_solve_ragged(pa_solver *solver, const pa_ragged *batch, int32_t schedule,
                    pa_result **out);

/* Score-only ragged solve writing float32 scores to a device buffer on the
 * solver's stream, without host synchronisation; for callers whose inputs
 * already sit in HBM.  Pair with pa_solver_sync. */
int pa_solve_ragged_scores_device(pa_solver *solver, const pa_ragged *batch, int32_t schedule,
                                  float *scores_device);
int pa_solver_sync(pa_solver *solver);

/* Allocation-free score solve (no reference counterpart: it is for callers
 * that cannot allocate device memory -- code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The same results as
 * pa_solve_ragged_scores_device, but every buffer its kernels use besides the
 * caller's inputs and outputs (problem and work-group descriptors, queue
 * tickets, borders, the similarity table, the strip-boundary row buffer)
 * lives in one static device array of 64 MiB per GPU, held until the process
 * ends; the call allocates nothing (no hipMalloc, hipFree, pinned memory,
 * stream or event of its own) and fails with PA_EINVAL when a batch needs
 * more scratch than that.  The payload must be device-resident (on_device =
 * 1, codes given, no raw symbols); general gap costs are not served.  Work
 * goes to `stream` (NULL: the null stream) without a host wait.  Every call
 * on one GPU shares that GPU's array: calls that overlap on one GPU (other
 * streams or threads) race, and a call's scratch holds its data only until
 * the next call. */
int pa_solve_ragged_scores_noalloc(pa_solver *solver, const pa_ragged *batch, int32_t schedule,
                                   float *scores_device, void *stream);

/* Timing of the last solve's DP kernel launches (HIP events on the solver's
 * stream): total milliseconds and launch count. */
int pa_solver_last_kernel_ms(const pa_solver *solver, double *ms, int32_t *launches,
                             double *cells);
/* Schedule of the last solve: work groups queued, problems run in band
 * mode (a long problem's strips computed concurrently by several
 * wavefronts) and, of those, problems whose forward was checkpointed (see
 * pa_solver_set_checkpoint).  No reference counterpart (diagnostics). */
int pa_solver_last_schedule(const pa_solver *solver, int32_t *groups, int32_t *band_problems,
                            int32_t *checkpointed);
/* Repeated launch sequences (same shapes and buffers) are captured as a
 * hipGraph and replayed; enable = 0 launches every kernel eagerly (also the
 * default when PYALIGN_AMD_NO_GRAPHS is set).  No reference counterpart. */
int pa_solver_set_graphs(pa_solver *solver, int32_t enable);
/* The solver's device work goes to `stream` (a hipStream_t; NULL is the
 * device's null stream) when external is 1, and to the solver's own stream
 * (the default) when it is 0.  On a caller's stream, pa_forward_ragged and
 * pa_backward return without waiting for the device: their results are
 * ordered on that stream like any other kernel's (a forward that ran problems
 * in band mode still needs pa_solver_sync to report a stalled strip
 * pipeline), and launch sequences are not captured as graphs.  A tape's
 * forward and backward must use the same stream.  No reference counterpart. */
int pa_solver_set_stream(pa_solver *solver, void *stream, int32_t external);

/* The stream the solver's device work goes to (its own, or the caller's set
 * by pa_solver_set_stream), so that a caller can order events or work of its
 * own with the solver's.  No reference counterpart. */
int pa_solver_stream(const pa_solver *solver, void **stream);
/* Checkpointed differentiation: in pa_forward_ragged, a problem (any
 * locality) on a device-resident dense or series payload whose longer side
 * exceeds min_len (default 2048) keeps only the bottom row of each strip of
 * 512 rows and the path's end (its seed), not its traceback; pa_backward
 * recomputes its traceback from those rows into a scratch buffer of the
 * solver (whole paths where a quarter of the free device memory holds them,
 * else one strip at a time, from the seed's strip up) and walks it.
 * Negative: never.  No reference counterpart. */
int pa_solver_set_checkpoint(pa_solver *solver, int64_t min_len);
/* Where the last solve from raw symbols met symbols outside the alphabet:
 * the first such position in s's and in t's symbol array, -1 for none.  No
 * reference counterpart (its encoders raise per problem, s before t). */
int pa_solver_last_bad_symbol(const pa_solver *solver, int64_t *pos_s, int64_t *pos_t);

/*
 * Differentiation (no reference counterpart: the reference has no
 * gradients).  The optimal score is piecewise linear in the similarity
 * values and the gap costs; on the optimal path
 *   dScore/dS(i, j)       = 1 on the path's matched cells (every path cell
 *                           for DTW, whose cost enters every cell),
 *   dScore/d(gap open)    = -(number of gap runs)        (maximize; + when minimizing),
 *   dScore/d(gap extend)  = -(sum over runs of length-1)  with w(k) = open + extend*(k-1),
 *   DTW with series:        dD/dx_i = sum over path cells (i, j) of 2 (x_i - y_j).
 * pa_forward_ragged runs the one-optimal forward pass on device-resident
 * inputs, writes the scores to scores_device and keeps the traceback on the
 * device in a tape; pa_backward walks every path once.  Gradient buffers
 * are accumulated into (+=) and have the forward payload's layout; grad_gap
 * is float32[n][2] (open, extend; linear gap costs use the open slot for
 * d/d(gap per symbol)).  Pass NULL for gradients not wanted.  The tape
 * borrows the forward's payload, which must outlive it, and the solver's
 * device buffers, which pa_tape_free hands back; a tape whose solver was
 * freed first can only be freed (pa_backward fails).
 */
typedef struct pa_tape pa_tape;
int pa_forward_ragged(pa_solver *solver, const pa_ragged *batch, float *scores_device, pa_tape **tape);
int pa_backward(pa_tape *tape, const float *grad_scores_device, float *grad_dense_device,
                float *grad_series_s_device, float *grad_series_t_device, float *grad_gap_device);
void pa_tape_free(pa_tape *tape);

/* ---- results (the tuple of per-lane objects, factory.h:699-808) ------ */

int64_t pa_result_count(const pa_result *r);
/* Scores of all problems; the worst value for a problem with no seed
 * (Solver::score, solver.h:2593-2608). */
int pa_result_scores(const pa_result *r, float *out);
/* 1 if problem k has a traceback seed (an optimal alignment exists). */
int pa_result_has_path(const pa_result *r, int64_t k);
/* Path of problem k, seed first (build_path, solver.h:858-935): n points. */
int64_t pa_result_path_len(const pa_result *r, int64_t k);
int pa_result_path(const pa_result *r, int64_t k, int32_t *uv);
/* Alignment of problem k (factory.h:130-193): s_to_t int32[len_s],
 * t_to_s int32[len_t], -1 where unaligned. */
int pa_result_alignment(const pa_result *r, int64_t k, int32_t *s_to_t, int32_t *t_to_s);
/* All problems of a one-optimal result at once (a whole bag's Solution
 * objects are built from two calls instead of two per problem; no reference
 * counterpart).  pa_result_paths: lens[k] = problem k's path length (-1: no
 * path); with uv non-NULL, the paths' (u, v) pairs concatenated in problem
 * order.  pa_result_alignments: every problem's s_to_t (len_s[k] entries)
 * and t_to_s (len_t[k] entries), concatenated in problem order, -1 for no
 * path. */
int pa_result_paths(const pa_result *r, int64_t *lens, int32_t *uv);
int pa_result_alignments(const pa_result *r, int32_t *s_to_t, int32_t *t_to_s);
/* All-optimal iteration (TracebackIterators<true>, solver.h:1568-1743):
 * returns 1 and fills score/path for the next alignment of problem k, 0 at
 * the end.  The path is read with pa_result_iter_path afterwards. */
int pa_result_iter_next(pa_result *r, int64_t k, float *score);
int64_t pa_result_iter_path_len(const pa_result *r, int64_t k);
int pa_result_iter_path(const pa_result *r, int64_t k, int32_t *uv);
void pa_result_free(pa_result *r);

/* Solution matrices (Solution::values / traceback_as_matrix /
 * traceback_as_edges, factory.h:251-262 and solver.h:2355-2471; read by
 * pyalign/solve.py:44-57).  pa_solve_ragged_matrices solves the batch again
 * keeping every layer's cell values and the traceback codes; the solver's
 * count (one or all optimal) fixes the traceback kind as in the reference
 * (traceback_1 / traceback_n).  Layers: 3 (D, P, Q) for affine gaps, else 1. */
int pa_solve_ragged_matrices(pa_solver *solver, const pa_ragged *batch, pa_result **out);
int32_t pa_result_layers(const pa_result *r); /* 0: no matrices kept */
/* float32[layers][len_s+1][len_t+1], border row and column included */
int pa_result_values(const pa_result *r, int64_t k, float *out);
/* int32[layers][len_s+1][len_t+1][2]: each cell's first predecessor (u, v),
 * the index type's minimum where there is none */
int pa_result_traceback(const pa_result *r, int64_t k, int32_t *out);
/* edges of one layer: int32[n][2][2] ((i-1, j-1), (u, v)) per predecessor,
 * cells row by row; pa_result_edges_len returns n (-1 on error) */
int64_t pa_result_edges_len(const pa_result *r, int64_t k, int32_t layer);
int pa_result_edges(const pa_result *r, int64_t k, int32_t layer, int32_t *out);

#ifdef __cplusplus
}
#endif

#endif /* PYALIGN_HIP_H */
