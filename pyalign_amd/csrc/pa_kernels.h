// pa_kernels.h -- the gfx950 DP kernel template (included by the pa_k_*.hip
// translation units, which instantiate disjoint subsets in parallel builds).
//
// Work decomposition.  A wavefront takes one group from a cost-sorted work
// queue.  A group holds 64/L problems; L consecutive lanes (a segment)
// cooperate on one problem.  Lane l of a segment owns R consecutive rows of
// the current strip (L*R rows); at step s it processes column j = s - l of
// its R rows, so the segment sweeps the matrix along anti-diagonals.  The
// bottom row's D and P of lane l reach lane l+1 one step later through a
// DPP wave shift (wave_shr:1); lane 0 of a segment reads the row above the
// strip from the border or, for long problems, from the row buffer written
// by the segment's last lane in the previous strip.  L = 1 is the
// inter-task mode (one problem per lane), L = 64 the intra-task mode (one
// problem per wavefront); the host autotuner picks (L, R) per problem.
//
// Every cell performs the reference's candidate fold in the reference's
// order with strict-improvement updates (pyalign/algorithm/solver.h:1311-1427;
// linear 2837-2854, Gotoh 2989-3035, DTW 3233-3250), so values are
// bit-identical to the reference's float32 arithmetic.  The `FAST` variant
// (local maximisation, score only, NaN-free table) uses v_max_f32: there
// every value is >= +0 and no -0 can arise (the zero candidate is first and
// a -0 never beats it, and +0 + -0 = +0), so max returns exactly the value
// the strict select keeps.
#pragma once

#include "pa_internal.h"

namespace pa {

template <int DIR>
__device__ __forceinline__ bool better(float a, float b) {
    if constexpr (DIR == DIR_MAX) return a > b;
    else return a < b;
}

template <int DIR>
__device__ __forceinline__ constexpr float worst() {
    return DIR == DIR_MAX ? -__builtin_inff() : __builtin_inff();
}

// lane k <- lane k-1 (wave_shr:1); lane 0 keeps its own value
__device__ __forceinline__ float shr1(float x) {
    return __int_as_float(
        __builtin_amdgcn_update_dpp(__float_as_int(x), __float_as_int(x), 0x138, 0xf, 0xf, false));
}

// Problems with an empty side: the seed comes from the border
// (solver.h:2042-2057, 2179-2187, 2273-2293).
template <int LOC, int DIR>
__device__ void write_empty(const KArgs &a, const Prob &pr) {
    const int ls = pr.len_s, lt = pr.len_t;
    float score = worst<DIR>();
    int si = NO_SEED, sj = 0;
    if (LOC == LOC_GLOBAL) {
        si = ls - 1;
        sj = lt - 1;
        score = ls == 0 ? a.border_t[lt] : a.border_s[ls];
    } else if (LOC == LOC_SEMI) {
        if (ls == 0 && lt > 0) { si = -1; sj = 0; score = a.border_t[1]; }
        else if (lt == 0 && ls > 0) { si = 0; sj = -1; score = a.border_s[1]; }
    }
    a.scores[pr.out] = score;
    if (a.seeds) {
        a.seeds[2 * pr.out] = si;
        a.seeds[2 * pr.out + 1] = sj;
    }
}

// Band mode: wait until strip `idx` has published `need` columns (lane 0
// only).  Bounded: after BAND_SPIN_LIMIT polls the wait gives up and flags
// the failure instead of spinning on.
__device__ __forceinline__ void band_wait(const KArgs &a, int idx, uint32_t need, uint32_t &seen) {
    uint32_t spins = 0;
    while (seen < need) {
        seen = __hip_atomic_load(a.progress + idx, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        if (seen >= need) break;
        ++spins;
        // give up after the limit, or at once when another wait already has
        if (spins > BAND_SPIN_LIMIT ||
            ((spins & 255u) == 0 && __hip_atomic_load(a.band_err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            __hip_atomic_store(a.band_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            seen = 0xffffffffu;
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
}

extern __shared__ float pa_dyn_lds[];  // the launch's dynamic LDS (table or staging windows)

template <int LOC, int DIR, int GAP, int FORM, int TR, int R, bool FAST, int SD, bool BAND = false>
__device__ __forceinline__ void solve_group(const KArgs &a, const Group &g, const float *__restrict__ tab,
                                            int lane) {
    constexpr float W = worst<DIR>();
    constexpr bool DENSE = FORM == F_DENSE || FORM == F_DROWS;
    constexpr bool TABLE = FORM == F_TABLE || FORM == F_TABLE_G;  // tab: LDS or global memory
    constexpr bool STAGED = FORM == F_DROWS;
    static_assert(!STAGED || R == DROWS_WIN, "staged dense rows need R == 8");
    constexpr int BITS = TR == TR_ALL ? 8 : (TR == TR_ONE ? (GAP == GAP_AFF ? 4 : 2) : 0);
    constexpr int WPC = TR == TR_SCORE ? 1 : (BITS * R + 31) / 32;
    constexpr int SDD = SD > 0 ? SD : 1;

    const int L = g.L;
    const int seg = lane >> g.lgL;
    const int l = lane & (L - 1);
    const bool has = seg < g.count;
    Prob pr;
    if (has) {
        pr = a.probs[g.first + seg];
    } else {
        pr.off_s = pr.off_t = pr.rb_off = pr.dv_off = 0;
        pr.tb_off = NO_TB_OFF;  // nothing stored, also by steps that compute unguarded
        pr.len_s = pr.len_t = 0;
        pr.si = pr.sj = pr.out = 0;
        pr.band = -1;
    }
    const int ls = pr.len_s, lt = pr.len_t;
    const int rps = L * R;
    const int nst = (ls > 0 && lt > 0) ? (ls + rps - 1) / rps : 0;
    const int steps = g.max_lt + L - 1;
    const int nsteps_p = lt + L - 1;  // this problem's steps per strip (traceback layout)
    const int PAD = a.K1 - 1;
    const float go_s = a.go_s, ge_s = a.ge_s, go_t = a.go_t, ge_t = a.ge_t;

    // ---- seed / score tracking (per lane, reduced over the segment) -------
    float best = W;
    int bi = -1, bj = -1;
    float rbest = W, cbest = W;
    int rj = -1, ci = -1;
    float gscore = W;
    // steady steps: the members' shortest length and fewest strips
    int min_lt = has ? lt : 0x7fffffff, min_nst = has ? nst : 0x7fffffff;
    for (int m = 1; m < WAVE; m <<= 1) {
        min_lt = min(min_lt, __shfl_xor(min_lt, m, WAVE));
        min_nst = min(min_nst, __shfl_xor(min_nst, m, WAVE));
    }
    min_lt = __builtin_amdgcn_readfirstlane(min_lt);  // the same in every lane
    min_nst = __builtin_amdgcn_readfirstlane(min_nst);
    const int last = ls - 1;
    const int st_last = last >= 0 ? last / rps : -1;
    const int l_last = last >= 0 ? (last % rps) / R : -1;
    const int r_last = last >= 0 ? last % R : -1;

    // band mode computes one strip (its rows' part of the seed search
    // goes to partials); otherwise the segment sweeps every strip
    const int st_lo = BAND ? g.strip : 0;
    const int st_hi = BAND ? g.strip + 1 : g.max_nst;
    for (int st = st_lo; st < st_hi; ++st) {
        const bool act = st < nst;
        const int row0 = st * rps + l * R;
        // row buffers: one reused row, or one row per strip boundary in band mode
        // a band slot (band mode, or a checkpointed forward) keeps every
        // strip's bottom row; otherwise one row is handed down and reused
        const bool keep_rows = BAND || pr.band >= 0;
        const int64_t rb_in = keep_rows ? pr.rb_off + int64_t(st - 1) * lt : pr.rb_off;
        const int64_t rb_out = keep_rows ? pr.rb_off + int64_t(st) * lt : pr.rb_off;
        [[maybe_unused]] uint32_t seen = 0;  // band mode: columns of strip st-1 known published

        int rowsym[R];
        bool rv[R];
        [[maybe_unused]] int64_t rowbase[R];
        float sa[R][SDD];
        float DlA[R], DlB[R], Ql[R];
        // Every load of the step loop is unconditional: a lane or row with
        // nothing to read loads a.safe (valid memory; its worst value for a
        // padded row) and a column index past the problem is clamped, so no
        // load sits behind a branch.
        const float *safeW = a.safe + (DIR == DIR_MIN ? 1 : 0);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row0 + r;
            rv[r] = act && i < ls;
            if constexpr (FORM == F_TABLE) {
                const uint8_t *cs = (const uint8_t *)a.codes_s;
                rowsym[r] = (rv[r] ? int(cs[pr.off_s + i]) : PAD) * a.K1;
            } else if constexpr (FORM == F_TABLE_G) {
                // a code beyond the table reads the pad row (no read outside it)
                const uint16_t *cs = (const uint16_t *)a.codes_s;
                rowsym[r] = (rv[r] ? min(int(cs[pr.off_s + i]), PAD) : PAD) * a.K1;
            } else if constexpr (FORM == F_BINARY) {
                const uint32_t *cs = (const uint32_t *)a.codes_s;
                rowsym[r] = rv[r] ? int(cs[pr.off_s + i]) : 0;
            } else if constexpr (DENSE) {
                rowbase[r] = rv[r] ? int64_t(i) * pr.si : 0;  // from the problem's first element
            } else if constexpr (FORM == F_SERIES) {
                const float *xs = (const float *)a.codes_s;
#pragma unroll
                for (int d = 0; d < SDD; ++d) sa[r][d] = rv[r] ? xs[(pr.off_s + i) * SD + d] : 0.0f;
            }
            DlA[r] = rv[r] ? a.border_s[i + 1] : W;  // D(i, -1)
            DlB[r] = DlA[r];
            Ql[r] = W;                               // Q(i, -1) (solver.h:2926-2933)
        }
        float botD = DlA[R - 1], botP = W;  // this lane's last row, handed down
        // diagonal predecessor of the strip's first row at j = 0: D(row0-1, -1)
        float prevD = W;
        if (l == 0) prevD = (row0 == 0) ? a.border_t[0] : (act && row0 <= ls ? a.border_s[row0] : W);

        // The column's inputs (symbol, series row or similarity column) and,
        // for lane 0, the row above the strip (border or row buffer) are
        // loaded one step ahead, read from these bases at the clamped column.
        [[maybe_unused]] const uint8_t *tb8 =
            act ? (const uint8_t *)a.codes_t + pr.off_t : (const uint8_t *)a.safe;
        [[maybe_unused]] const uint16_t *tb16 =
            act ? (const uint16_t *)a.codes_t + pr.off_t : (const uint16_t *)a.safe;
        [[maybe_unused]] const uint32_t *tb32 =
            act ? (const uint32_t *)a.codes_t + pr.off_t : (const uint32_t *)a.safe;
        [[maybe_unused]] const float *tbf = act ? (const float *)a.codes_t + pr.off_t * SDD : a.safe;
        [[maybe_unused]] const float *dbase = act ? a.dense + pr.off_s : safeW;
        [[maybe_unused]] const int64_t dsj = act ? int64_t(pr.sj) : 0;
        const float *upD = safeW, *upP = safeW;  // D(-1, j) / P(-1, j) or the strip above's bottom row
        int upDs = 0, upPs = 0;                  // their strides in floats
        if (l == 0 && act) {
            if (st == 0) {
                upD = a.border_t + 1;
                upDs = 1;
            } else {
                upD = reinterpret_cast<const float *>(a.rowbuf + rb_in);
                upP = upD + 1;
                upDs = upPs = 2;
            }
        }
        // a column's inputs, loaded two steps ahead into one of two slots
        struct Col {
            int sym;
            float v[SDD];
            float S[DENSE ? R : 1];
            float D, P;
        };
        Col cA, cB;
        // steady fetches: the column needs no clamp (the steady bounds keep it
        // inside every member), lanes without a member read index 0, and the
        // row-above strides (0, 1 or 2 floats) are a mask and a shift
        const int fmask = act ? -1 : 0;
        const int upMD = upDs != 0 ? -1 : 0, upSD = upDs >> 1;
        const int upMP = upPs != 0 ? -1 : 0, upSP = upPs >> 1;
        auto fetch = [&](auto steady, int s, Col &c) {
            constexpr bool STEADY = decltype(steady)::value;
            const int j = s - l;
            const int jc = STEADY ? (j & fmask) : (act ? (j < 0 ? 0 : (j < lt ? j : lt - 1)) : 0);
            if constexpr (FORM == F_TABLE) {
                c.sym = int(tb8[jc]);
            } else if constexpr (FORM == F_TABLE_G) {
                c.sym = min(int(tb16[jc]), PAD);
            } else if constexpr (FORM == F_BINARY) {
                c.sym = int(tb32[jc]);
            } else if constexpr (FORM == F_DENSE) {
#pragma unroll
                for (int r = 0; r < R; ++r) c.S[r] = dbase[rowbase[r] + int64_t(jc) * dsj];
            } else if constexpr (FORM == F_SERIES) {
#pragma unroll
                for (int d = 0; d < SDD; ++d) c.v[d] = tbf[jc * SD + d];
            }
            if constexpr (BAND) {
                if (a.band_sync && l == 0 && act && st > 0 && j < lt)
                    band_wait(a, pr.band + st - 1, uint32_t(j + 1), seen);
            }
            if constexpr (STEADY) {
                c.D = upD[(jc & upMD) << upSD];
                c.P = upP[(jc & upMP) << upSP];
            } else {
                c.D = upD[jc * upDs];
                c.P = upP[jc * upPs];
            }
        };

        // Stores of a step (the traceback words and the last lane's bottom
        // row; in band mode also its published column count) go out at the
        // top of the next step, after that step's loads.  Like the loads they
        // are unconditional (a lane with nothing to store stores to a.sink),
        // so the compiler counts every memory operation exactly and waits
        // for a load only where its value is used, two steps after it left.
        uint32_t *const sink_tb = a.sink;
        float2 *const sink_rb = reinterpret_cast<float2 *>(a.sink + 8);
        [[maybe_unused]] uint32_t *pend_dst = sink_tb;
        [[maybe_unused]] uint32_t pend_w[WPC];
#pragma unroll
        for (int w = 0; w < WPC; ++w) pend_w[w] = 0;
        float2 *pend_rb = sink_rb;
        float2 pend_v = make_float2(0.0f, 0.0f);
        [[maybe_unused]] bool pend_pub = false;
        [[maybe_unused]] uint32_t pend_cols = 0;
        auto flush = [&]() {
            if constexpr (TR != TR_SCORE) {
#pragma unroll
                for (int w = 0; w < WPC; ++w) pend_dst[w] = pend_w[w];
                pend_dst = sink_tb;
            }
            *pend_rb = pend_v;
            pend_rb = sink_rb;
            if constexpr (BAND) {
                if (pend_pub)
                    __hip_atomic_store(a.progress + pr.band + st, pend_cols, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
                pend_pub = false;
            }
        };
        fetch(std::integral_constant<bool, false>{}, 0, cA);
        fetch(std::integral_constant<bool, false>{}, 1, cB);

        // F_DROWS: window w holds this lane's R rows at columns [8w - l,
        // 8w - l + 8), read from global memory with two 16-byte loads per row
        // at a start clamped into the problem, kept in LDS (two windows per
        // lane, used alternately).  Row q of window w+1 is loaded four steps
        // before it is written to LDS at step 8w + q, during window w, so the
        // loads have four steps to arrive; step s reads its column from
        // window s / 8 with the window's offset (start - clamped start).
        [[maybe_unused]] float *lw = STAGED ? pa_dyn_lds + threadIdx.x * DROWS_LANE_FLOATS + 8 : nullptr;
        [[maybe_unused]] float stg[4][DROWS_WIN];
        auto wclamp = [&](int w) { return act ? min(max(DROWS_WIN * w - l, 0), lt - DROWS_WIN) : 0; };
        auto wload = [&](int w, int q, float (&dst)[DROWS_WIN]) {
            const float *src = dbase + rowbase[q] + wclamp(w);
            float4 v0, v1;
            __builtin_memcpy(&v0, src, 16);
            __builtin_memcpy(&v1, src + 4, 16);
            dst[0] = v0.x; dst[1] = v0.y; dst[2] = v0.z; dst[3] = v0.w;
            dst[4] = v1.x; dst[5] = v1.y; dst[6] = v1.z; dst[7] = v1.w;
        };
        auto wstore = [&](int w, int q, const float (&src)[DROWS_WIN]) {
            float *d = lw + (w & 1) * (R * DROWS_WIN) + q * DROWS_WIN;
            *reinterpret_cast<float4 *>(d) = make_float4(src[0], src[1], src[2], src[3]);
            *reinterpret_cast<float4 *>(d + 4) = make_float4(src[4], src[5], src[6], src[7]);
        };
        [[maybe_unused]] const float *wread = lw;  // window of the current step, shifted by its offset
        if constexpr (STAGED) {
#pragma unroll
            for (int q = 0; q < R; ++q) {  // window 0, then rows 0..3 of window 1 in flight
                float tmp[DROWS_WIN];
                wload(0, q, tmp);
                wstore(0, q, tmp);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) wload(1, q, stg[q]);
        }

        // one step: column j = s - l of this lane's R rows; the left column
        // D(i, j-1) is read from Din and this column written to Dout (two
        // banks used alternately: row r+1 still needs row r's old value as
        // its diagonal, so one bank would cost a register copy per cell)
        // A steady step (every lane of every member inside its problem's strip
        // and columns) computes its cells unguarded; lanes without a member
        // compute from safe inputs and store nothing (no traceback offset).
        auto step = [&](auto steady, auto phase, auto keepv, int s, float (&Din)[R], float (&Dout)[R], Col &cur) {
            constexpr bool STEADY = decltype(steady)::value;
            constexpr int C = decltype(phase)::value;  // s % 8 (staged dense rows only)
            constexpr int KM = decltype(keepv)::value;
            constexpr bool KEEPV = KM == 1;  // cell values stored (a.dvals set)
            // no traceback stored (a checkpointed forward's group): the codes are not computed
            constexpr bool NOTB = KM == 2 || KM == 4;
            // DTW over finite series: min / max folds, codes from equality tests
            constexpr bool MXF = KM >= 3;
            const int j = s - l;
            const int tsym = cur.sym;
            [[maybe_unused]] float tv[SDD];
            [[maybe_unused]] float Scol[DENSE ? R : 1];
            if constexpr (FORM == F_SERIES) {
#pragma unroll
                for (int d = 0; d < SDD; ++d) tv[d] = cur.v[d];
            }
            if constexpr (FORM == F_DENSE) {
#pragma unroll
                for (int r = 0; r < R; ++r) Scol[r] = cur.S[r];
            }
            if constexpr (STAGED) {
                const int w = s >> 3;
                wstore(w + 1, C, stg[C & 3]);      // row C of the next window, loaded four steps ago
                wload(w + ((C + 4) >> 3) + 1, (C + 4) & 7, stg[C & 3]);
                if constexpr (C == 0) {
                    // the window's offset; within +-7 wherever a cell of this
                    // lane is computed, clamped elsewhere so the address stays
                    // inside the lane's slack
                    const int off = act ? min(max(DROWS_WIN * w - l - wclamp(w), -8), 8) : 0;
                    wread = lw + (w & 1) * (R * DROWS_WIN) + off;
                }
#pragma unroll
                for (int r = 0; r < R; ++r) Scol[r] = wread[r * DROWS_WIN + C];
            }
            float inD = shr1(botD), inP = shr1(botP);
            if (l == 0) {
                inD = cur.D;
                inP = cur.P;
            }
            fetch(steady, s + 2, cur);
            flush();
            const float Ddiag = prevD;
            prevD = inD;
            if (STEADY || (act && j >= 0 && j < lt)) {
                float Dd = Ddiag, Dup = inD, Pup = inP;
                uint32_t tbw[WPC];
#pragma unroll
                for (int w = 0; w < WPC; ++w) tbw[w] = 0;

#pragma unroll
                for (int r = 0; r < R; ++r) {
                    // similarity S(i, j) (factory.h:573-576, 624-630, 673-679).
                    // A padded row (i >= len_s, the last strip's end) feeds only
                    // padded rows below it, stores nothing and enters no seed
                    // search (the searches test rv), so its value is free except
                    // for the score-only local maximum, which folds every row:
                    // there it reads the worst similarity.
                    constexpr bool PADW = LOC == LOC_LOCAL && TR == TR_SCORE;
                    float S;
                    if constexpr (TABLE) {
                        S = tab[rowsym[r] + tsym];
                    } else if constexpr (FORM == F_BINARY) {
                        S = rowsym[r] == tsym ? a.eq : a.ne;
                        if constexpr (PADW) S = rv[r] ? S : W;
                    } else if constexpr (DENSE) {
                        S = PADW ? (rv[r] ? Scol[r] : W) : Scol[r];
                    } else {
                        // squared Euclidean cost, dimensions summed in order
                        float c = 0.0f;
#pragma unroll
                        for (int d = 0; d < SDD; ++d) {
                            const float df = sa[r][d] - tv[d];
                            c = (d == 0) ? df * df : c + df * df;
                        }
                        S = PADW ? (rv[r] ? c : W) : c;
                    }

                    float D, P = 0.0f, Q = 0.0f;
                    uint32_t bits = 0;
                    if constexpr (GAP == GAP_DTW) {
                        // D = opt(diag, up, left) + cost (solver.h:3236-3248)
                        if constexpr (MXF) {
                            // finite series: every candidate is +0 or above
                            // (or the worst value), never -0 or NaN, so the
                            // fold's value is the strict fold's and its
                            // choice the first candidate equal to it
                            const float m = DIR == DIR_MAX ? fmaxf(fmaxf(Dd, Dup), Din[r])
                                                           : fminf(fminf(Dd, Dup), Din[r]);
                            D = m + S;
                            bits = Dd == m ? 1u : (Dup == m ? 2u : 3u);
                        } else {
                            float m = Dd;
                            int src = 1;
                            if (better<DIR>(Dup, m)) { m = Dup; src = 2; }
                            if (better<DIR>(Din[r], m)) { m = Din[r]; src = 3; }
                            D = m + S;
                            bits = uint32_t(src);
                        }
                    } else if constexpr (GAP == GAP_LIN) {
                        // [0], diag, up, left (solver.h:2841-2852)
                        const float dg = Dd + S;
                        const float up = Dup + go_s;
                        const float lf = Din[r] + go_t;
                        if constexpr (FAST) {
                            D = fmaxf(fmaxf(fmaxf(dg, 0.0f), up), lf);
                        } else {
                            int src;
                            if constexpr (LOC == LOC_LOCAL) {
                                D = 0.0f;
                                src = 0;
                                if (better<DIR>(dg, D)) { D = dg; src = 1; }
                            } else {
                                D = dg;
                                src = 1;
                            }
                            if (better<DIR>(up, D)) { D = up; src = 2; }
                            if (better<DIR>(lf, D)) { D = lf; src = 3; }
                            if constexpr (TR == TR_ONE) bits = uint32_t(src);
                            if constexpr (TR == TR_ALL)
                                bits = (LOC == LOC_LOCAL && D == 0.0f ? 1u : 0u) | (dg == D ? 2u : 0u) |
                                       (up == D ? 4u : 0u) | (lf == D ? 8u : 0u);
                        }
                    } else {
                        // Gotoh (solver.h:2993-3033)
                        const float po = Dup + go_s, pe = Pup + ge_s;
                        const float qo = Din[r] + go_t, qe = Ql[r] + ge_t;
                        const float dg = Dd + S;
                        if constexpr (FAST) {
                            P = fmaxf(po, pe);
                            Q = fmaxf(qo, qe);
                            D = fmaxf(fmaxf(fmaxf(dg, 0.0f), P), Q);
                        } else {
                            const bool pext = better<DIR>(pe, po);
                            P = pext ? pe : po;
                            const bool qext = better<DIR>(qe, qo);
                            Q = qext ? qe : qo;
                            int src;
                            if constexpr (LOC == LOC_LOCAL) {
                                D = 0.0f;
                                src = 0;
                                if (better<DIR>(dg, D)) { D = dg; src = 1; }
                            } else {
                                D = dg;
                                src = 1;
                            }
                            if (better<DIR>(P, D)) { D = P; src = 2; }
                            if (better<DIR>(Q, D)) { D = Q; src = 3; }
                            if constexpr (TR == TR_ONE)
                                bits = uint32_t(src) | (pext ? 4u : 0u) | (qext ? 8u : 0u);
                            if constexpr (TR == TR_ALL)
                                bits = (LOC == LOC_LOCAL && D == 0.0f ? 1u : 0u) | (dg == D ? 2u : 0u) |
                                       (P == D ? 4u : 0u) | (Q == D ? 8u : 0u) | (po == P ? 16u : 0u) |
                                       (pe == P ? 32u : 0u) | (qo == Q ? 64u : 0u) | (qe == Q ? 128u : 0u);
                        }
                    }

                    // ---- seeds ------------------------------------------
                    const int i = row0 + r;
                    if constexpr (LOC == LOC_LOCAL) {
                        if constexpr (TR == TR_SCORE) {
                            // pad rows never exceed the optimum: their S is the
                            // worst value and gap costs are non-negative
                            if constexpr (FAST) best = fmaxf(best, D);
                            else best = better<DIR>(D, best) ? D : best;
                        } else {
                            // Optima over i desc, j desc, first strict optimum
                            // wins (solver.h:2050-2056): largest i, then largest j
                            if (rv[r] && (better<DIR>(D, best) || (D == best && i >= bi))) {
                                best = D;
                                bi = i;
                                bj = j;
                            }
                        }
                    }
                    // cell values: local all-optimal seeds read D; the Solution
                    // matrices read every layer (solver.h:2355-2375)
                    if constexpr (TR != TR_SCORE && KEEPV) {
                        if (rv[r]) {
                            float *dv = a.dvals + pr.dv_off + int64_t(i) * lt + j;
                            dv[0] = D;
                            if constexpr (GAP == GAP_AFF) {
                                if (a.val_layers == 3) {
                                    const int64_t lay = int64_t(ls) * lt;
                                    dv[lay] = P;
                                    dv[2 * lay] = Q;
                                }
                            }
                        }
                    }
                    if constexpr (TR != TR_SCORE && !NOTB) tbw[(r * BITS) / 32] |= bits << ((r * BITS) % 32);

                    Dd = Din[r];
                    Dout[r] = D;
                    Ql[r] = Q;
                    Dup = D;
                    Pup = P;
                }
                botD = Dup;
                botP = Pup;
                if (l == L - 1 && st + 1 < nst) {
                    pend_rb = a.rowbuf + rb_out + j;
                    pend_v = make_float2(botD, botP);
                    if constexpr (BAND) {
                        pend_pub = a.band_sync && (((j + 1) % BAND_PUB) == 0 || j == lt - 1);
                        pend_cols = uint32_t(j + 1);
                    }
                }

                if constexpr (TR != TR_SCORE && !NOTB) {
                    // problems of a checkpointed forward keep no traceback
                    if (pr.tb_off != NO_TB_OFF) {
                        pend_dst = a.tb + pr.tb_off + (int64_t(st) * nsteps_p + s) * int64_t(L) * WPC + l * WPC;
#pragma unroll
                        for (int w = 0; w < WPC; ++w) pend_w[w] = tbw[w];
                    }
                }

                // ---- seeds from the finished column (global / semiglobal) --
                // (global: D(ls-1, lt-1) is read once the strip is done)
                if constexpr (LOC == LOC_SEMI) {
                    if (st == st_last && l == l_last) {
                        float v = Dout[0];
#pragma unroll
                        for (int r = 1; r < R; ++r) v = (r == r_last) ? Dout[r] : v;
                        // last row, j ascending, strict (solver.h:2285-2287)
                        if (better<DIR>(v, rbest)) { rbest = v; rj = j; }
                    }
                    if (j == lt - 1) {
                        // last column, i ascending, strict (solver.h:2289-2291)
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            if (rv[r] && better<DIR>(Dout[r], cbest)) { cbest = Dout[r]; ci = row0 + r; }
                    }
                }
            }
        };
        using std::integral_constant;
        // the step loop, built once per kind of group so that the usual loop
        // carries no per-cell test: 0 codes stored, 1 codes and cell values
        // (a.dvals), 2 neither (the groups of a checkpointed forward keep
        // strip rows only; their folds keep their order, their codes go);
        // 3 and 4 as 0 and 2 for DTW over finite series, with min / max folds
        using Gen = integral_constant<bool, false>;
        auto sweep = [&](auto kv) {
            int s = 0;
            if constexpr (STAGED) {
                // eight steps a round: the phase s % 8 is a compile-time constant
                for (; s < steps; s += 8) {
                    step(Gen{}, integral_constant<int, 0>{}, kv, s, DlA, DlB, cA);
                    if (s + 1 < steps) step(Gen{}, integral_constant<int, 1>{}, kv, s + 1, DlB, DlA, cB);
                    if (s + 2 < steps) step(Gen{}, integral_constant<int, 2>{}, kv, s + 2, DlA, DlB, cA);
                    if (s + 3 < steps) step(Gen{}, integral_constant<int, 3>{}, kv, s + 3, DlB, DlA, cB);
                    if (s + 4 < steps) step(Gen{}, integral_constant<int, 4>{}, kv, s + 4, DlA, DlB, cA);
                    if (s + 5 < steps) step(Gen{}, integral_constant<int, 5>{}, kv, s + 5, DlB, DlA, cB);
                    if (s + 6 < steps) step(Gen{}, integral_constant<int, 6>{}, kv, s + 6, DlA, DlB, cA);
                    if (s + 7 < steps) step(Gen{}, integral_constant<int, 7>{}, kv, s + 7, DlB, DlA, cB);
                }
            } else {
                if constexpr (!BAND) {
                    // steps [sa, sb) are steady: s >= L - 1 puts every lane at a
                    // column >= 0, s + 2 < the members' shortest length keeps
                    // every lane's column and the one it fetches inside its
                    // member (only when every member still has this strip); both
                    // ends even, so the banks keep their order
                    const int sa = L & ~1;
                    const int sb = st < min_nst ? ((min_lt - 2) & ~1) : 0;
                    if (sa < sb) {
                        for (; s < sa; s += 2) {
                            step(Gen{}, integral_constant<int, 0>{}, kv, s, DlA, DlB, cA);
                            step(Gen{}, integral_constant<int, 0>{}, kv, s + 1, DlB, DlA, cB);
                        }
                        for (int k = (sb - sa) >> 1; k > 0; --k, s += 2) {  // a uniform trip count
                            step(integral_constant<bool, true>{}, integral_constant<int, 0>{}, kv, s, DlA, DlB, cA);
                            step(integral_constant<bool, true>{}, integral_constant<int, 0>{}, kv, s + 1, DlB, DlA,
                                 cB);
                        }
                    }
                }
                for (; s + 1 < steps; s += 2) {
                    step(Gen{}, integral_constant<int, 0>{}, kv, s, DlA, DlB, cA);
                    step(Gen{}, integral_constant<int, 0>{}, kv, s + 1, DlB, DlA, cB);
                }
                if (s < steps) step(Gen{}, integral_constant<int, 0>{}, kv, s, DlA, DlB, cA);
            }
        };
        if constexpr (TR == TR_ONE) {
            const bool notb = a.probs[g.first].tb_off == NO_TB_OFF;  // (one problem a group)
            bool mx = false;
            if constexpr (GAP == GAP_DTW && FORM == F_SERIES) mx = a.finite != nullptr && *a.finite != 0u;
            if (a.dvals != nullptr) {
                sweep(integral_constant<int, 1>{});
            } else if (mx) {
                if constexpr (GAP == GAP_DTW && FORM == F_SERIES) {
                    if (notb) sweep(integral_constant<int, 4>{});
                    else sweep(integral_constant<int, 3>{});
                }
            } else if (notb) {
                sweep(integral_constant<int, 2>{});
            } else {
                sweep(integral_constant<int, 0>{});
            }
        } else if constexpr (TR != TR_SCORE) {
            if (a.dvals != nullptr) sweep(integral_constant<int, 1>{});
            else sweep(integral_constant<int, 0>{});
        } else if constexpr (GAP == GAP_DTW && FORM == F_SERIES) {
            if (a.finite != nullptr && *a.finite != 0u) sweep(integral_constant<int, 3>{});  // (no codes to compute)
            else sweep(integral_constant<int, 0>{});
        } else {
            sweep(integral_constant<int, 0>{});
        }
        if constexpr (LOC == LOC_GLOBAL) {
            // D(ls-1, lt-1): column lt-1 was this lane's step lt-1+l, which
            // wrote DlB when even and DlA when odd; later steps write nothing
            if (st == st_last && l == l_last) {
                const bool odd = ((lt - 1 + l) & 1) != 0;
                float v = odd ? DlA[0] : DlB[0];
#pragma unroll
                for (int r = 1; r < R; ++r) v = (r == r_last) ? (odd ? DlA[r] : DlB[r]) : v;
                gscore = v;
            }
        }
        flush();
        // the row buffer written by this strip's last lane is read next strip
        __threadfence_block();
    }

    // ---- reduce over the segment's lanes, lane 0 writes -----------------
    float score = W;
    int si = NO_SEED, sj = 0;
    if constexpr (LOC == LOC_LOCAL) {
        for (int m = 1; m < L; m <<= 1) {
            const float ob = __shfl_xor(best, m, WAVE);
            if constexpr (TR == TR_SCORE) {
                if constexpr (FAST) best = fmaxf(best, ob);
                else best = better<DIR>(ob, best) ? ob : best;
            } else {
                const int oi = __shfl_xor(bi, m, WAVE);
                const int oj = __shfl_xor(bj, m, WAVE);
                if (better<DIR>(ob, best) || (ob == best && oi > bi)) {
                    best = ob;
                    bi = oi;
                    bj = oj;
                }
            }
        }
        score = best;
        si = bi;
        sj = bj;
    } else if constexpr (LOC == LOC_GLOBAL) {
        const int src = (seg << g.lgL) + (l_last > 0 ? l_last : 0);
        score = __shfl(gscore, src, WAVE);
        si = ls - 1;
        sj = lt - 1;
    } else {
        const int src = (seg << g.lgL) + (l_last > 0 ? l_last : 0);
        rbest = __shfl(rbest, src, WAVE);
        rj = __shfl(rj, src, WAVE);
        for (int m = 1; m < L; m <<= 1) {
            const float oc = __shfl_xor(cbest, m, WAVE);
            const int oi = __shfl_xor(ci, m, WAVE);
            if (better<DIR>(oc, cbest) || (oc == cbest && oi >= 0 && (ci < 0 || oi < ci))) {
                cbest = oc;
                ci = oi;
            }
        }
        // Optima: the row scan first; the column scan replaces only on strict
        // improvement
        if (better<DIR>(cbest, rbest)) { score = cbest; si = ci; sj = lt - 1; }
        else { score = rbest; si = ls - 1; sj = rj; }
    }
    if constexpr (BAND) {
        if (has && l == 0 && g.strip < nst) {
            BandPartial bp;
            bp.v1 = LOC == LOC_LOCAL ? best : (LOC == LOC_GLOBAL ? score : cbest);
            bp.i1 = LOC == LOC_LOCAL ? bi : ci;
            bp.j1 = bj;
            bp.v2 = rbest;
            bp.j2 = rj;
            bp.pad = 0;
            a.partials[pr.band + g.strip] = bp;
        }
        return;
    }
    if (has && l == 0) {
        if (ls == 0 || lt == 0) {
            write_empty<LOC, DIR>(a, pr);
        } else {
            if (LOC == LOC_LOCAL && TR == TR_SCORE) { si = 0; sj = 0; }
            a.scores[pr.out] = score;
            if (a.seeds) {
                a.seeds[2 * pr.out] = si;
                a.seeds[2 * pr.out + 1] = sj;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// The score-only local maximisation path over a LDS table (the batched local
// Gotoh / Smith-Waterman workload).  Same decomposition as solve_group with
// the traceback and seed bookkeeping removed: every value is >= +0, so
// v_max_f32 reproduces the strict select exactly (see the file comment),
// the local borders are +0, and the optimum is a plain running maximum
// (padded rows read the pad symbol, whose worst-valued entries keep them
// below the optimum).  The column's symbol is loaded one step ahead.
//
// Gotoh here carries P' = max(P, 0) and Q' = max(Q, 0) instead of P and Q:
// with a non-positive extension term e, max(0, max(P, 0) + e) =
// max(0, P + e), so P'(i) = max(0, D(i-1) + o, P'(i-1) + e) = max(0, P(i))
// by induction (the same float additions whenever P >= 0, both sides
// clamped to 0 otherwise), and D = max(0, diag, P, Q) = max3(diag, P', Q'):
// one v_max3_f32 per layer instead of a separate zero clamp.
//
// The left column D(i, j-1) lives in two register banks used alternately
// by consecutive steps (steps go in pairs): row r+1 still needs the old
// D(i, j-1) of row r as its diagonal after row r's new value exists, so a
// single bank would cost a register copy per cell.  Both banks start at the
// border value and a lane leaves them untouched while it waits for its
// first column, so whichever bank a step reads holds the border.
//
// PAIR: the LDS holds the table for row-symbol pairs, entry ((x1 * K1 + x2)
// * Kc + y) = (S(x1, y), S(x2, y)) with Kc = K1 - 1 real column symbols, so
// rows 2q and 2q+1 of a lane get their similarities from one ds_read_b64 at
// one address add (K1^2 * Kc * 8 bytes: 70,560 for 20 symbols and the pad).
// ---------------------------------------------------------------------------
template <int GAP, bool PAIR, int R>
__device__ __forceinline__ void solve_group_fast(const KArgs &a, const Group &g, const float *__restrict__ tab,
                                                 int lane) {
    constexpr float NEG = -__builtin_inff();
    const int L = g.L;
    const int seg = lane >> g.lgL;
    const int l = lane & (L - 1);
    const bool has = seg < g.count;
    int64_t off_s = 0, off_t = 0, rb_off = 0;
    int ls = 0, lt = 0, out = 0;
    if (has) {
        const Prob &pr = a.probs[g.first + seg];
        off_s = pr.off_s;
        off_t = pr.off_t;
        rb_off = pr.rb_off;
        ls = pr.len_s;
        lt = pr.len_t;
        out = pr.out;
    }
    const int rps = L * R;
    const int nst = (ls > 0 && lt > 0) ? (ls + rps - 1) / rps : 0;
    const int steps = g.max_lt + L - 1;
    const int K1 = a.K1;
    const int PAD = K1 - 1;
    const float go_s = a.go_s, ge_s = a.ge_s, go_t = a.go_t, ge_t = a.ge_t;
    const uint8_t *__restrict__ cs = (const uint8_t *)a.codes_s;
    const uint8_t *__restrict__ ct = (const uint8_t *)a.codes_t;
    float best = NEG;
    // steady steps: the members' shortest length and fewest strips, and safe
    // symbol addresses (the first member's) for lanes without a member
    int min_lt = has ? lt : 0x7fffffff, min_nst = has ? nst : 0x7fffffff;
    for (int m = 1; m < WAVE; m <<= 1) {
        min_lt = min(min_lt, __shfl_xor(min_lt, m, WAVE));
        min_nst = min(min_nst, __shfl_xor(min_nst, m, WAVE));
    }
    min_lt = __builtin_amdgcn_readfirstlane(min_lt);  // the same in every lane
    min_nst = __builtin_amdgcn_readfirstlane(min_nst);
    // index of the symbol of column s - l + 1 is t_ix + s
    const int64_t t_ix = (has ? off_t : a.probs[g.first].off_t) + 1 - l;

    for (int st = 0; st < g.max_nst; ++st) {
        const bool act = st < nst;
        const bool rd_top = has && l == 0 && st > 0;         // lane 0 reads the strip above's row
        const bool wr_bot = l == L - 1 && st + 1 < nst;      // the last lane writes it for the next
        const int row0 = st * rps + l * R;
        int rowsym[R];
        [[maybe_unused]] int rowpair[R / 2 > 0 ? R / 2 : 1];
        float DlA[R], DlB[R], Ql[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row0 + r;
            rowsym[r] = ((act && i < ls) ? int(cs[off_s + i]) : PAD) * K1;
            DlA[r] = 0.0f;  // local border D(i, -1) = +0 (solver.h:2021-2029)
            DlB[r] = 0.0f;
            Ql[r] = NEG;    // Q(i, -1)
        }
        if constexpr (PAIR) {
            // byte offset of the (row 2q, row 2q+1) symbol pair's block
#pragma unroll
            for (int q = 0; q < R / 2; ++q) rowpair[q] = (rowsym[2 * q] + rowsym[2 * q + 1] / K1) * (K1 - 1) * 8;
        }
        float botD = 0.0f, botP = NEG;
        float prevD = 0.0f;  // D(row0-1, -1) = +0
        // global memory is touched only at the top of a step, one step
        // ahead: the column's symbol and (lane 0) the row above the strip are
        // loaded for the next step, and the last lane's bottom row of the
        // previous step is stored, so every load and store has a whole step
        // of cells to complete before anything waits on it
        int tnext = 0;
        if (act && l == 0 && lt > 0) tnext = int(ct[off_t]);
        // lane 0's row above, loaded a step ahead into alternating registers
        // (a step reads one and loads the other: no copies between steps)
        float2 nxA = make_float2(0.0f, NEG), nxB = nxA;  // D(-1, j) = +0, P(-1, j) = worst on the first strip
        if (l == 0 && st > 0 && act && lt > 0) nxA = a.rowbuf[rb_off];
        bool pend = false;  // a bottom row value waiting to be stored
        float2 pendv = make_float2(0.0f, 0.0f);
        int pendj = 0;

        // one step: column j = s - l of this lane's R rows, left column read
        // from Din, this column written to Dout.  A steady step (every lane of
        // every member inside its problem's strip and columns, the next column
        // too) drops the range tests; lanes without a member read the first
        // member's symbols and discard what they compute.
        auto step = [&](auto steady, int s, float (&Din)[R], float (&Dout)[R], const float2 &nxin, float2 &nxout) {
            constexpr bool STEADY = decltype(steady)::value;
            const int j = s - l;
            const int tsym = tnext;
            const float rD = nxin.x, rP = nxin.y;
            if constexpr (STEADY) {
                tnext = int(ct[t_ix + s]);
                if (rd_top) nxout = a.rowbuf[rb_off + 1 + s];  // lane 0: column s + 1
            } else {
                if (act && j + 1 >= 0 && j + 1 < lt) tnext = int(ct[off_t + j + 1]);
                if (l == 0 && st > 0 && act && j + 1 < lt) nxout = a.rowbuf[rb_off + j + 1];
            }
            if (pend) {  // after the loads: the next wait on them need not wait for it
                if constexpr (STEADY) a.rowbuf[rb_off - 1 - l + s] = pendv;  // column s - 1 - l
                else a.rowbuf[rb_off + pendj] = pendv;
                pend = false;
            }
            float inD = shr1(botD), inP = shr1(botP);
            if (l == 0) {
                inD = rD;
                inP = rP;
            }
            const float Ddiag = prevD;
            prevD = inD;
            if (STEADY || (act && j >= 0 && j < lt)) {
                float Dd = Ddiag, Pup = inP;
                float Gup = inD + go_s;
                [[maybe_unused]] float Sv[R];
                if constexpr (PAIR) {
                    const int t8 = tsym * 8;
#pragma unroll
                    for (int q = 0; q < R / 2; ++q) {
                        const float2 v = *reinterpret_cast<const float2 *>(
                            reinterpret_cast<const char *>(tab) + rowpair[q] + t8);
                        Sv[2 * q] = v.x;
                        Sv[2 * q + 1] = v.y;
                    }
                }
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    float S;
                    if constexpr (PAIR) S = Sv[r];
                    else S = tab[rowsym[r] + tsym];
                    const float dg = Dd + S;
                    float D;
                    if constexpr (GAP == GAP_AFF) {
                        const float Pc = fmaxf(fmaxf(Gup, Pup + ge_s), 0.0f);
                        const float Qc = fmaxf(fmaxf(Din[r] + go_t, Ql[r] + ge_t), 0.0f);
                        D = fmaxf(fmaxf(dg, Pc), Qc);
                        Ql[r] = Qc;
                        Pup = Pc;
                    } else {
                        D = fmaxf(fmaxf(fmaxf(dg, 0.0f), Gup), Din[r] + go_t);
                    }
                    best = fmaxf(best, D);
                    Dd = Din[r];
                    Dout[r] = D;
                    Gup = D + go_s;
                }
                botD = Dout[R - 1];
                botP = Pup;
                if constexpr (STEADY) {  // pendv is read only where pend is set
                    pend = wr_bot;
                    pendv = make_float2(botD, botP);
                } else if (wr_bot) {
                    pend = true;
                    pendv = make_float2(botD, botP);
                    pendj = j;
                }
            }
        };
        using General = std::integral_constant<bool, false>;
        using Steady = std::integral_constant<bool, true>;
        // steps [sa, sb) are steady: s >= L - 1 puts every lane at a column
        // >= 0, s + 1 < the members' shortest length keeps every lane's next
        // column inside its member (only when every member still has this
        // strip); both ends are even so the register banks keep their order
        const int sa = L & ~1;
        const int sb = st < min_nst ? ((min_lt - 1) & ~1) : 0;
        int s = 0;
        if (sa < sb) {
            for (; s < sa; s += 2) {
                step(General{}, s, DlA, DlB, nxA, nxB);
                step(General{}, s + 1, DlB, DlA, nxB, nxA);
            }
            for (int k = (sb - sa) >> 1; k > 0; --k, s += 2) {  // a uniform trip count
                step(Steady{}, s, DlA, DlB, nxA, nxB);
                step(Steady{}, s + 1, DlB, DlA, nxB, nxA);
            }
            pendj = s - 1 - l;  // the column of a store still pending
        }
        for (; s + 1 < steps; s += 2) {
            step(General{}, s, DlA, DlB, nxA, nxB);
            step(General{}, s + 1, DlB, DlA, nxB, nxA);
        }
        if (s < steps) step(General{}, s, DlA, DlB, nxA, nxB);
        if (pend) a.rowbuf[rb_off + pendj] = pendv;
        __threadfence_block();
    }
    for (int m = 1; m < L; m <<= 1) best = fmaxf(best, __shfl_xor(best, m, WAVE));
    if (has && l == 0) a.scores[out] = (ls == 0 || lt == 0) ? NEG : best;
}

template <int GAP, bool PAIR, int... RS>
__global__ __launch_bounds__(FAST_BLOCK) void pa_fast_kernel(KArgs a) {
    extern __shared__ float lds_table[];
    const int K1 = a.K1;
    if constexpr (PAIR) {
        const int Kc = K1 - 1;
        const int n = K1 * K1 * Kc;
        float2 *pt = reinterpret_cast<float2 *>(lds_table);
        for (int k = threadIdx.x; k < n; k += blockDim.x) {
            const int y = k % Kc, p = k / Kc;
            pt[k] = make_float2(a.table[(p / K1) * K1 + y], a.table[(p % K1) * K1 + y]);
        }
    } else {
        const int n = K1 * K1;
        for (int k = threadIdx.x; k < n; k += blockDim.x) lds_table[k] = a.table[k];
    }
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1);
    while (true) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(a.queue, 1u);
        t = uint32_t(__builtin_amdgcn_readlane(int(t), 0));  // uniform: the group's fields load as scalars
        if (t >= uint32_t(a.n_groups)) break;
        const Group &g = a.groups[t];
        const int R = g.R;
        (void)((R == RS ? (solve_group_fast<GAP, PAIR, RS>(a, g, lds_table, lane), true) : false) || ...);
    }
}

template <int GAP, bool PAIR>
static hipError_t go_fast_t(const LaunchSpec &s, const KArgs &a, hipStream_t st, int *occ_out) {
    const void *fn = reinterpret_cast<const void *>(&pa_fast_kernel<GAP, PAIR, PA_FAST_RS_LIST>);
    if (occ_out) {
        int blocks = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, fn, FAST_BLOCK, s.lds_bytes) != hipSuccess)
            blocks = 1;
        *occ_out = blocks * (FAST_BLOCK / WAVE);
        return hipSuccess;
    }
    const int block = s.block > 0 && s.block <= FAST_BLOCK && s.block % WAVE == 0 ? s.block : 256;
    hipLaunchKernelGGL((pa_fast_kernel<GAP, PAIR, PA_FAST_RS_LIST>), dim3(s.grid), dim3(block), s.lds_bytes, st, a);
    return hipGetLastError();
}

template <int GAP>
static hipError_t go_fast(const LaunchSpec &s, const KArgs &a, hipStream_t st, int *occ_out) {
    return s.pair ? go_fast_t<GAP, true>(s, a, st, occ_out) : go_fast_t<GAP, false>(s, a, st, occ_out);
}

template <int LOC, int DIR, int GAP, int FORM, int TR, int R, bool FAST, int SD>
__global__ __launch_bounds__(256) void pa_solve_kernel(KArgs a) {
    extern __shared__ float lds_table[];
    if constexpr (FORM == F_TABLE) {
        const int n = a.K1 * a.K1;
        for (int k = threadIdx.x; k < n; k += blockDim.x) lds_table[k] = a.table[k];
        __syncthreads();
    }
    // a table too large for LDS is read where it is (global memory, L2-resident)
    const float *tab = FORM == F_TABLE_G ? a.table : lds_table;
    const int lane = threadIdx.x & (WAVE - 1);
    // Persistent wavefronts drain the cost-sorted group queue; every wave
    // leaves once its ticket passes the group count.
    while (true) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(a.queue, 1u);
        t = uint32_t(__builtin_amdgcn_readlane(int(t), 0));  // uniform: the group's fields load as scalars
        if (t >= uint32_t(a.n_groups)) break;
        const Group &g = a.groups[t];
        if constexpr (!FAST && TR != TR_ALL) {
            if (g.strip >= 0) {
                solve_group<LOC, DIR, GAP, FORM, TR, R, FAST, SD, true>(a, g, tab, lane);
                continue;
            }
        }
        solve_group<LOC, DIR, GAP, FORM, TR, R, FAST, SD>(a, g, tab, lane);
    }
}

template <int LOC, int DIR, int GAP, int FORM, int TR, int R, bool FAST, int SD>
static hipError_t go(const LaunchSpec &s, const KArgs &a, hipStream_t st) {
    hipLaunchKernelGGL((pa_solve_kernel<LOC, DIR, GAP, FORM, TR, R, FAST, SD>), dim3(s.grid), dim3(256),
                       s.lds_bytes, st, a);
    return hipGetLastError();
}

template <int LOC, int DIR, int GAP, int FORM, int TR, int R, bool FAST, int SD>
static int occ(const LaunchSpec &s) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, reinterpret_cast<const void *>(&pa_solve_kernel<LOC, DIR, GAP, FORM, TR, R, FAST, SD>), 256,
            s.lds_bytes) != hipSuccess)
        return 4;
    return blocks * 4;
}

// Generic dispatch for one (LOC, DIR, GAP): forms x traceback modes at the
// general strip height R = 8; the hot score-only table path gets more.
template <int LOC, int DIR, int GAP, bool OCC>
static hipError_t dispatch_lg(const LaunchSpec &s, const KArgs &a, hipStream_t st, int *occ_out) {
#define PA_GO(FORM, TR, R, FAST, SD)                                                        \
    do {                                                                                    \
        if (OCC) {                                                                          \
            *occ_out = occ<LOC, DIR, GAP, FORM, TR, R, FAST, SD>(s);                        \
            return hipSuccess;                                                              \
        }                                                                                   \
        return go<LOC, DIR, GAP, FORM, TR, R, FAST, SD>(s, a, st);                          \
    } while (0)
    if constexpr (GAP == GAP_DTW) {
        if (s.tr == TR_ALL) return hipErrorInvalidValue;
        if (s.R != 8) return hipErrorInvalidValue;
        if (s.form == F_DENSE) {
            if (s.tr == TR_SCORE) PA_GO(F_DENSE, TR_SCORE, 8, false, 0);
            PA_GO(F_DENSE, TR_ONE, 8, false, 0);
        }
        if (s.form == F_DROWS) {
            if (s.tr == TR_SCORE) PA_GO(F_DROWS, TR_SCORE, 8, false, 0);
            PA_GO(F_DROWS, TR_ONE, 8, false, 0);
        }
        if (s.form == F_SERIES) {
            switch (a.dim) {
                case 1: if (s.tr == TR_SCORE) PA_GO(F_SERIES, TR_SCORE, 8, false, 1); PA_GO(F_SERIES, TR_ONE, 8, false, 1);
                case 2: if (s.tr == TR_SCORE) PA_GO(F_SERIES, TR_SCORE, 8, false, 2); PA_GO(F_SERIES, TR_ONE, 8, false, 2);
                case 3: if (s.tr == TR_SCORE) PA_GO(F_SERIES, TR_SCORE, 8, false, 3); PA_GO(F_SERIES, TR_ONE, 8, false, 3);
                case 4: if (s.tr == TR_SCORE) PA_GO(F_SERIES, TR_SCORE, 8, false, 4); PA_GO(F_SERIES, TR_ONE, 8, false, 4);
                default: return hipErrorInvalidValue;
            }
        }
        return hipErrorInvalidValue;
    } else {
        if (s.form == F_TABLE && s.tr == TR_SCORE) {
            if (s.fastmax) {
                if constexpr (LOC == LOC_LOCAL && DIR == DIR_MAX) {
                    if (s.R == 0) return go_fast<GAP>(s, a, st, occ_out);
                    switch (s.R) {
                        case 1: PA_GO(F_TABLE, TR_SCORE, 1, true, 0);
                        case 2: PA_GO(F_TABLE, TR_SCORE, 2, true, 0);
                        case 3: PA_GO(F_TABLE, TR_SCORE, 3, true, 0);
                        case 4: PA_GO(F_TABLE, TR_SCORE, 4, true, 0);
                        case 5: PA_GO(F_TABLE, TR_SCORE, 5, true, 0);
                        case 6: PA_GO(F_TABLE, TR_SCORE, 6, true, 0);
                        case 7: PA_GO(F_TABLE, TR_SCORE, 7, true, 0);
                        case 8: PA_GO(F_TABLE, TR_SCORE, 8, true, 0);
                        case 10: PA_GO(F_TABLE, TR_SCORE, 10, true, 0);
                        case 12: PA_GO(F_TABLE, TR_SCORE, 12, true, 0);
                        case 14: PA_GO(F_TABLE, TR_SCORE, 14, true, 0);
                        case 16: PA_GO(F_TABLE, TR_SCORE, 16, true, 0);
                        case 20: PA_GO(F_TABLE, TR_SCORE, 20, true, 0);
                        case 24: PA_GO(F_TABLE, TR_SCORE, 24, true, 0);
                        case 28: PA_GO(F_TABLE, TR_SCORE, 28, true, 0);
                        case 32: PA_GO(F_TABLE, TR_SCORE, 32, true, 0);
                        default: return hipErrorInvalidValue;
                    }
                }
                return hipErrorInvalidValue;
            }
            switch (s.R) {
                case 4: PA_GO(F_TABLE, TR_SCORE, 4, false, 0);
                case 8: PA_GO(F_TABLE, TR_SCORE, 8, false, 0);
                case 16: PA_GO(F_TABLE, TR_SCORE, 16, false, 0);
                case 32: PA_GO(F_TABLE, TR_SCORE, 32, false, 0);
                default: return hipErrorInvalidValue;
            }
        }
        if (s.R == 4) {  // band-mode strips of a long problem (more, shorter strips)
            if (s.form == F_TABLE && s.tr == TR_ONE) PA_GO(F_TABLE, TR_ONE, 4, false, 0);
            if (s.form == F_BINARY && s.tr == TR_SCORE) PA_GO(F_BINARY, TR_SCORE, 4, false, 0);
            if (s.form == F_BINARY && s.tr == TR_ONE) PA_GO(F_BINARY, TR_ONE, 4, false, 0);
            return hipErrorInvalidValue;
        }
        if (s.R != 8) return hipErrorInvalidValue;
        switch (s.form) {
            case F_TABLE:
                if (s.tr == TR_ONE) PA_GO(F_TABLE, TR_ONE, 8, false, 0);
                PA_GO(F_TABLE, TR_ALL, 8, false, 0);
            case F_TABLE_G:
                if (s.tr == TR_SCORE) PA_GO(F_TABLE_G, TR_SCORE, 8, false, 0);
                if (s.tr == TR_ONE) PA_GO(F_TABLE_G, TR_ONE, 8, false, 0);
                PA_GO(F_TABLE_G, TR_ALL, 8, false, 0);
            case F_BINARY:
                if (s.tr == TR_SCORE) PA_GO(F_BINARY, TR_SCORE, 8, false, 0);
                if (s.tr == TR_ONE) PA_GO(F_BINARY, TR_ONE, 8, false, 0);
                PA_GO(F_BINARY, TR_ALL, 8, false, 0);
            case F_DENSE:
                if (s.tr == TR_SCORE) PA_GO(F_DENSE, TR_SCORE, 8, false, 0);
                if (s.tr == TR_ONE) PA_GO(F_DENSE, TR_ONE, 8, false, 0);
                PA_GO(F_DENSE, TR_ALL, 8, false, 0);
            case F_DROWS:
                if (s.tr == TR_SCORE) PA_GO(F_DROWS, TR_SCORE, 8, false, 0);
                if (s.tr == TR_ONE) PA_GO(F_DROWS, TR_ONE, 8, false, 0);
                PA_GO(F_DROWS, TR_ALL, 8, false, 0);
            default: return hipErrorInvalidValue;
        }
    }
#undef PA_GO
}

}  // namespace pa

#include "pa_long.h"
