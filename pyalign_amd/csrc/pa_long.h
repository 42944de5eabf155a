// pa_long.h -- intra-task mode for long problems: one problem's strips swept
// by the waves of a workgroup, bottom rows handed from wave to wave through
// LDS (included at the end of pa_kernels.h).
//
// Decomposition.  A problem of ls x lt is cut into strips of 64 x R rows
// (lane l of a strip owns rows row0 + l*R .. +R-1 and at its step s computes
// column j = s - l, an anti-diagonal sweep).  LONG_WAVES consecutive strips
// form a band; a workgroup of LONG_WAVES waves computes one band, wave w
// strip band * LONG_WAVES + w.  Bands are taken in order from a ticket
// counter, so a band only waits for the band above it, taken earlier by a
// running workgroup: no circular wait whatever the grid.
//
// Hand-offs.  Inside a strip, lane l's bottom row (D, P) and the column
// symbol reach lane l+1 one step later through a DPP wave shift
// (wave_shr:1); lane 0's inputs -- the row above the strip and the column's
// symbol -- arrive as the shift's fill value, read from an LDS ring filled by
// the wave above (its lane 63 writes (D, P, symbol) of each column it
// finishes).  The waves of a workgroup run in phases of LONG_T steps
// separated by workgroup barriers; wave w runs long_dw phases behind wave
// w-1, so every ring entry a wave reads was written in an earlier phase (the
// barrier orders the LDS write before the read).  Lane 0 alone reads a
// phase's ring entries, at the start of the phase (a read by every lane
// would move 64 copies of each through the CU's LDS, which the four waves
// share), so each of its inputs is one DPP move whose fill operand is
// already in a register; lane 63's (D, P) pass through a shift register
// (lane 0 <- lane 63 by a wave rotate, the rest shifted down a lane) and
// lanes 0 .. LONG_T - 1 write the phase's entries at its end.  The loader
// wave writes every ring entry's symbol a phase before it is read.  Wave
// 0's ring holds the strip border (the first band) or the band above's
// bottom row: the band's last wave writes its phase's entries to global memory as
// two 8-byte words per column tagged with the column (relaxed atomics at
// device scope: no cache flush), and the loader wave of the band below
// reads them LONG_K phases before it stages them, re-reading (bounded) any
// whose tag is not there yet.  A band starts once the band above has
// published far enough ahead that those reads find their words (a re-read
// costs the whole workgroup a round trip to L2, and the delays would add up
// band after band).  The tagged buffer is cleared before the launch.
//
// Cells.  The same fold as the general kernel in the reference's push order
// with strict improvement (pyalign/algorithm/solver.h:1311-1427; linear
// 2837-2854, Gotoh 2989-3035), so values, traceback bits (the general
// kernel's layout: the walk kernel reads them unchanged) and seeds are
// identical.  A phase in which every lane's column lies inside the problem
// runs without branches: rows below the last one score the worst
// similarity, so they never win a seed (scores) or are masked out of it
// (local paths), and every lane tracks its own last-row optimum, of which
// the last row's lane is kept.  Each strip leaves its part of the seed
// search in a band-mode partial that pa_band_combine_kernel folds in strip
// order.
#pragma once

#include "pa_internal.h"

namespace pa {

// lane k <- lane k-1 (wave_shr:1); lane 0 takes `in` (the strip's input)
__device__ __forceinline__ float shr1_in(float x, float in) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(in), __float_as_int(x), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ int shr1_in_i(int x, int in) {
    return __builtin_amdgcn_update_dpp(in, x, 0x138, 0xf, 0xf, false);
}
// shift register step: lane k <- reg lane k-1, lane 0 <- x lane 63 (wave_ror:1)
__device__ __forceinline__ float push63(float reg, float x) {
    return shr1_in(reg, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x13C, 0xf, 0xf, false)));
}


// the step offset of compute wave w behind the phase grid (-w mod LONG_T)
__device__ __forceinline__ constexpr int long_boff(int w) { return (LONG_T - w % LONG_T) % LONG_T; }
// the first ring slot of local phase k: lane 0 of wave w reads column
// k T - b_w + t from slot (k T + t) mod RING; lane 63 of wave w writes column
// k T - b_w + t - 63 to slot (k T + long_ooff(w) + t) mod RING of ring w + 1
__device__ __forceinline__ constexpr int long_ooff(int w) {
    return ((long_boff(w + 1) - long_boff(w) - 63) % LONG_RING + LONG_RING) % LONG_RING;
}

// MX: the strict compare-and-select of every candidate fold done as max
// (maximising) or min; the host chooses it where it returns the same bits:
// similarities finite and never -0, and (maximising global problems) gap
// costs above zero, so that no candidate is -0 or NaN (x + y is -0 only when
// both are -0) and max / min keep the strict fold's value.  The strict fold
// then chooses the first candidate in push order equal to that value, so
// traceback codes come from equality tests beside the value chain instead
// of lengthening it.
template <int LOC, int DIR, int GAP, int FORM, int TR, bool MX>
__global__ __launch_bounds__((LONG_WAVES + 1) * WAVE) void pa_long_kernel(KArgs a, LongArgs la) {
    static_assert(LONG_R == 1, "one row per lane");
    constexpr int T = LONG_T;
    constexpr int K = LONG_K;
    constexpr int RING = LONG_RING;
    constexpr int LW = LONG_WAVES;
    constexpr int DW = long_dw(FORM);
    constexpr float W = worst<DIR>();
    static_assert(FORM == F_TABLE || FORM == F_BINARY, "long mode reads table codes or binary symbols");

    extern __shared__ float pa_dyn_lds[];
    // ring[w]: (D, P, symbol bits, -) of the row above compute wave w's
    // strip at each column (D, P from wave w - 1 or, for wave 0, the
    // loader; the symbol from the loader); then the similarity table
    float4 *ring = reinterpret_cast<float4 *>(pa_dyn_lds);
    float *tab = pa_dyn_lds + LW * RING * 4;
    __shared__ int s_ticket;
    if constexpr (FORM == F_TABLE) {
        const int n = a.K1 * a.K1;
        for (int k = threadIdx.x; k < n; k += blockDim.x) tab[k] = a.table[k];
    }
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int l = threadIdx.x & (WAVE - 1);
    const bool loader = wave == LW;

    while (true) {
        __syncthreads();  // the table is loaded; the last band is done with the rings
        if (threadIdx.x == 0) s_ticket = int(atomicAdd(la.ticket, 1u));
        __syncthreads();
        const int ticket = s_ticket;
        if (ticket >= la.total_bands) break;
        // ticket -> (problem, band)
        int q = 0, band = ticket;
        Prob pr;
        int nst = 0;
        for (;;) {
            pr = a.probs[la.idx[q]];
            nst = (pr.len_s + WAVE - 1) / WAVE;
            const int nbands = (nst + LW - 1) / LW;
            if (band < nbands || q + 1 >= la.n) break;
            band -= nbands;
            ++q;
        }
        const int lt = pr.len_t;
        const int nsp = lt + WAVE - 1;  // steps per strip
        // every wave runs the same number of phases (workgroup barriers): up
        // to the last wave's last step, a multiple of K (the loader's
        // unrolled read-ahead)
        const int phases = ((LW - 1) * DW + (nsp + T - 1) / T + 1 + K - 1) / K * K;
        // the band above hands its bottom row down through xb (two 8-byte
        // words per column, tagged with the column + 1)
        uint64_t *xb = reinterpret_cast<uint64_t *>(a.rowbuf + pr.rb_off);

        if (loader) {
            // ---- wave 0's ring: the strip border or the band above --------
            // Every load is unconditional (all lanes, columns clamped into
            // the problem) so that no wait is placed at a branch join: the
            // words of phase p are read at phase p - K and waited for only
            // when they are staged.
            auto run = [&](auto from_band) {
                constexpr bool XB = decltype(from_band)::value;
                const uint64_t *xb_in = xb + 2 * int64_t(XB ? band - 1 : 0) * lt;
                float4 *ring0 = ring;
                auto issue = [&](int c, uint64_t &d, uint64_t &p, uint32_t &sy) {
                    const int cc = min(max(c, 0), lt - 1);
                    if constexpr (FORM == F_TABLE) sy = ((const uint8_t *)a.codes_t)[pr.off_t + cc];
                    else sy = ((const uint32_t *)a.codes_t)[pr.off_t + cc];
                    if constexpr (XB) {
                        d = __hip_atomic_load(xb_in + 2 * cc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        p = __hip_atomic_load(xb_in + 2 * cc + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    } else {
                        d = __float_as_uint(a.border_t[cc + 1]);  // D(-1, c)
                        p = 0;
                    }
                };
                // column c (lanes with `mine`) into the ring; words whose tag
                // is not there yet are read again (bounded; a wait that gives
                // up flags the failure)
                auto stage = [&](int c, bool mine, uint64_t &d, uint64_t &p, uint32_t sy) {
                    if constexpr (XB) {
                        const uint32_t tag = uint32_t(c + 1);
                        bool good = !mine || c >= lt || (uint32_t(d >> 32) == tag && uint32_t(p >> 32) == tag);
                        uint32_t spins = 0;
                        while (!__all(good)) {
                            if (++spins > LONG_SPIN_LIMIT ||
                                ((spins & 63u) == 0 &&
                                 __hip_atomic_load(a.band_err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                                __hip_atomic_store(a.band_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                break;
                            }
                            __builtin_amdgcn_s_sleep(2);
                            if (!good) {
                                d = __hip_atomic_load(xb_in + 2 * c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                p = __hip_atomic_load(xb_in + 2 * c + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                                good = uint32_t(d >> 32) == tag && uint32_t(p >> 32) == tag;
                            }
                        }
                    }
                    if (mine) {
                        float4 e;
                        if (c >= lt) e = make_float4(W, W, 0.0f, 0.0f);
                        else e = make_float4(__uint_as_float(uint32_t(d)), XB ? __uint_as_float(uint32_t(p)) : W,
                                             __uint_as_float(sy), 0.0f);  // P(-1, c) is the worst value
                        ring0[c & (RING - 1)] = e;
                    }
                };
                // wave 0 reads phase k's columns [k T, (k + 1) T) at the start
                // of phase k (the table form also column (k + 1) T): lane
                // l < T stages column due(p) = (p + DUE) T + l at phase p
                constexpr int DUE = FORM == F_TABLE ? 2 : 1;
                const bool mine = l < T;
                auto due = [&](int p) { return (p + DUE) * T + (l < T ? l : T - 1); };
                if constexpr (XB) {
                    // start once the band above has published the column
                    // this band's read-ahead reaches a few phases in
                    const int c0 = min(lt - 1, (DUE + K + 1) * T - 1);
                    const uint32_t tag = uint32_t(c0 + 1);
                    uint32_t spins = 0;
                    while (uint32_t(__hip_atomic_load(xb_in + 2 * c0 + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >>
                                    32) != tag) {
                        if (++spins > LONG_SPIN_LIMIT ||
                            ((spins & 63u) == 0 &&
                             __hip_atomic_load(a.band_err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
                            __hip_atomic_store(a.band_err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                        __builtin_amdgcn_s_sleep(2);
                    }
                }
                {   // the first phases: columns [0, DUE T)
                    const bool m0 = l < DUE * T;
                    uint64_t d, p;
                    uint32_t sy;
                    issue(l, d, p, sy);
                    stage(l, m0, d, p, sy);
                }
                // the symbols of compute wave sw = 1 + l / T (lanes below
                // (LW - 1) T): its local phase k's at global phase
                // P_sw + k - DUE, column k T - b_sw + l mod T into slot
                // (k T + l mod T) mod RING
                const int sw = 1 + l / T, st_sw = band * LW + sw;
                const bool sym_lane = sw < LW && st_sw < nst;
                const int sk0 = DUE - sw * DW;  // local phase = global phase + sk0
                auto sym_col = [&](int p) { return (p + sk0) * T - long_boff(sw) + (l & (T - 1)); };
                auto sym_issue = [&](int p, uint32_t &sy) {
                    const int cc = min(max(sym_col(p), 0), lt - 1);
                    if constexpr (FORM == F_TABLE) sy = ((const uint8_t *)a.codes_t)[pr.off_t + cc];
                    else sy = ((const uint32_t *)a.codes_t)[pr.off_t + cc];
                };
                auto sym_stage = [&](int p, uint32_t sy) {
                    const int k = p + sk0;
                    if (sym_lane && k >= 0)
                        reinterpret_cast<float *>(ring + sw * RING + ((k * T + (l & (T - 1))) & (RING - 1)))[2] =
                            __uint_as_float(sy);
                };
                uint64_t hD[K], hP[K];
                uint32_t hs[K], hy[K];
#pragma unroll
                for (int u = 0; u < K; ++u) {
                    issue(due(u), hD[u], hP[u], hs[u]);
                    sym_issue(u, hy[u]);
                }
                __syncthreads();  // the compute waves' first ring entries
                for (int p0 = 0; p0 < phases; p0 += K) {
#pragma unroll
                    for (int u = 0; u < K; ++u) {
                        stage(due(p0 + u), mine, hD[u], hP[u], hs[u]);
                        sym_stage(p0 + u, hy[u]);
                        issue(due(p0 + u + K), hD[u], hP[u], hs[u]);
                        sym_issue(p0 + u + K, hy[u]);
                        __syncthreads();
                    }
                }
            };
            if (band > 0) run(std::true_type{});
            else run(std::false_type{});
            continue;
        }

        // ---- a compute wave: strip st, one row per lane ---------------------
        const int ls = pr.len_s;
        const int st = band * LW + wave;
        const bool has = st < nst;
        const int row = st * WAVE + l;
        const bool rv = has && row < ls;
        const int last = ls - 1;
        const int st_last = last / WAVE;
        const int l_last = last % WAVE;
        const int bw = long_boff(wave);
        const int first = wave * DW;  // this wave's local phase 0
        const int kmax = (nsp + bw + T - 1) / T;  // its phases: steps -b_w .. nsp - 1
        const float4 *in_ring = ring + wave * RING;
        float4 *out_ring = ring + (wave + 1) * RING;  // (wave + 1 < LW)
        // the band's last wave hands its rows to the band below
        const bool to_band = wave == LW - 1 && st + 1 < nst;
        uint64_t *xb_out = xb + 2 * int64_t(band) * lt;
        const int ooff = long_ooff(wave);
        const float go_s = a.go_s, ge_s = a.ge_s, go_t = a.go_t, ge_t = a.ge_t;
        [[maybe_unused]] const int PAD = a.K1 - 1;

        // the row's symbol; rows below the last one score the worst value
        int rowsym;
        float veq = W, vne = W;
        if constexpr (FORM == F_TABLE) {
            rowsym = (rv ? int(((const uint8_t *)a.codes_s)[pr.off_s + row]) : PAD) * a.K1;
        } else {
            rowsym = rv ? int(((const uint32_t *)a.codes_s)[pr.off_s + row]) : 0;
            if (rv) {
                veq = a.eq;
                vne = a.ne;
            }
        }
        float Dl = rv ? a.border_s[row + 1] : W;  // D(i, j - 1), at the start D(i, -1)
        float Ql = W;                              // Q(i, j - 1)
        // this lane's D and P as the next lane sees them one step later
        float botD = Dl, botP = W;
        // D(i - 1, j - 1); at the start D(i - 1, -1); rows below the last take W
        float prevD = (row == 0) ? a.border_t[0] : (has && row <= ls ? a.border_s[row] : W);
        int msym = 0;  // the column symbol this lane used last step
        [[maybe_unused]] float Snext = 0.0f;  // F_TABLE: next step's similarity
        float best = W;
        int bj = -1;
        float rbest = W, cbest = W;
        int rj = -1, ci = -1;
        float gscore = W;
        // traceback words: codes of PK steps each (step s at position
        // (s + bw) % PK of word (s + bw) / PK), a phase's T steps in T / PK
        // whole words
        constexpr int TB = GAP == GAP_AFF ? 4 : 2;
        constexpr int PK = long_pack(TB);
        [[maybe_unused]] uint32_t *tbp =
            TR == TR_ONE ? a.tb + pr.tb_off + int64_t(st) * long_strip_words(nsp, TB) * WAVE + l : nullptr;
        [[maybe_unused]] uint32_t tacc = 0;  // the codes of the current word so far

        // the fold of one direction: max / min where exact
        auto opt = [](float x, float y) {
            if constexpr (DIR == DIR_MAX) return fmaxf(x, y);
            else return fminf(x, y);
        };

        // one step: column j = s - l; `e` is lane 0's ring entry (column s),
        // `en` the next step's
        // lane 63's last T outputs (D, P), newest in lane 0
        float oD = W, oP = W;
        // checks a phase needs (bits): LO, some lanes before column 0 (the
        // strip's first phases); HI, some lanes at or past the last column
        // (its last phases); NEG, steps before the strip's first step (its
        // first phase)
        constexpr int LO = 1, HI = 2, NEG = 4;
        // tbq: this lane's first traceback word of the phase (step s0 + t's
        // code goes to tbq[(t / PK) * WAVE]: constant offsets within a phase)
        auto body = [&](int s, uint32_t *tbq, auto tc, const float4 &e, const float4 &en, auto checked) {
            [[maybe_unused]] constexpr int t = decltype(tc)::value;
            constexpr int CHK = decltype(checked)::value;
            const float inD = shr1_in(botD, e.x);
            const float inP = shr1_in(botP, e.y);
            const int tsym = shr1_in_i(msym, __float_as_int(e.z));
            msym = tsym;
            float S;
            if constexpr (FORM == F_TABLE) {
                if constexpr ((CHK & LO) != 0) {
                    if (s == 0) Snext = tab[rowsym + int(min(uint32_t(tsym), uint32_t(PAD)))];  // (no step before)
                }
                S = Snext;
                // the next step's symbol is the previous lane's current one
                const int nsym = shr1_in_i(tsym, __float_as_int(en.z));
                Snext = tab[rowsym + int(min(uint32_t(nsym), uint32_t(PAD)))];
            } else {
                S = rowsym == tsym ? veq : vne;
            }
            const float Dd = prevD;
            prevD = inD;
            const int j = s - l;
            // checked steps: every lane computes a cell, those whose column
            // lies outside the problem keep their state (selects, no branch)
            const bool act = ((CHK & LO) == 0 || l <= s) && ((CHK & HI) == 0 || l > s - lt);
            {
                float D, P = 0.0f, Q = 0.0f;
                [[maybe_unused]] uint32_t bits = 0;
                const float dg = Dd + S;
                if constexpr (MX) {
                    float up, lf;  // the vertical and horizontal candidates
                    if constexpr (GAP == GAP_LIN) {
                        up = inD + go_s;
                        lf = Dl + go_t;
                    } else {
                        P = opt(inD + go_s, inP + ge_s);
                        Q = opt(Dl + go_t, Ql + ge_t);
                        up = P;
                        lf = Q;
                    }
                    D = opt(opt(dg, up), lf);
                    if constexpr (LOC == LOC_LOCAL) D = opt(D, 0.0f);
                    if constexpr (TR == TR_ONE) {
                        // the first candidate equal to the fold, in push
                        // order (local: zero, diagonal, vertical, horizontal)
                        uint32_t src = up == D ? 2u : 3u;
                        src = dg == D ? 1u : src;
                        if constexpr (LOC == LOC_LOCAL) src = D == 0.0f ? 0u : src;
                        bits = src;
                        if constexpr (GAP == GAP_AFF) {
                            bits |= better<DIR>(inP + ge_s, inD + go_s) ? 4u : 0u;
                            bits |= better<DIR>(Ql + ge_t, Dl + go_t) ? 8u : 0u;
                        }
                    }
                } else if constexpr (GAP == GAP_LIN) {
                    const float up = inD + go_s;
                    const float lf = Dl + go_t;
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
                    bits = uint32_t(src);
                } else {
                    const float po = inD + go_s, pe = inP + ge_s;
                    const float qo = Dl + go_t, qe = Ql + ge_t;
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
                    bits = uint32_t(src) | (pext ? 4u : 0u) | (qext ? 8u : 0u);
                }
                if constexpr (LOC == LOC_LOCAL) {
                    if constexpr (TR == TR_SCORE) {
                        const float nb = MX ? opt(best, D) : (better<DIR>(D, best) ? D : best);  // padded rows stay below
                        best = act ? nb : best;
                    } else {
                        // Optima over i desc, j desc (solver.h:2050-2056): in
                        // one row a later column wins a tie
                        if (act && rv && (DIR == DIR_MAX ? D >= best : D <= best)) {
                            best = D;
                            bj = j;
                        }
                    }
                } else if constexpr (LOC == LOC_SEMI) {
                    // every lane tracks its row; the last row's lane is kept
                    // (last row, j ascending: solver.h:2285-2287)
                    if constexpr (TR == TR_SCORE && MX) {
                        rbest = act ? opt(rbest, D) : rbest;
                    } else if (act && better<DIR>(D, rbest)) {
                        rbest = D;
                        rj = j;
                    }
                    if constexpr ((CHK & HI) != 0) {
                        if (rv && j == lt - 1) {  // last column (solver.h:2289-2291)
                            cbest = D;
                            ci = row;
                        }
                    }
                } else {
                    if constexpr ((CHK & HI) != 0) {
                        if (j == lt - 1) gscore = D;
                    }
                }
                // (the words of lanes outside the problem, and the codes of
                // steps outside the strip, lie in the strip's area and are
                // never read)
                if constexpr (TR == TR_ONE) {
                    constexpr int r = t % PK;
                    tacc = r == 0 ? bits : tacc | (bits << (TB * r));
                    if constexpr (r == PK - 1) tbq[(t / PK) * WAVE] = tacc;
                }
                Dl = act ? D : Dl;
                Ql = act ? Q : Ql;
                botD = act ? D : botD;
                botP = act ? P : botP;
            }
        };

        auto step = [&](int s, uint32_t *tbq, auto tc, const float4 &e, const float4 &en, auto checked) {
            constexpr int CHK = decltype(checked)::value;
            if ((CHK & NEG) == 0 || s >= 0) body(s, tbq, tc, e, en, checked);  // (steps before the strip's first do nothing)
            // lane 63 hands its bottom row (column s - 63) down
            oD = push63(oD, botD);
            oP = push63(oP, botP);
        };

        // a phase's ring entries are read by lane 0 at its start (written in
        // an earlier phase); the first step waits for the first of them
        auto phase_steps = [&](int s0, int obase, auto checked) {
            using std::integral_constant;
            uint32_t *const tbq = TR == TR_ONE ? tbp + int64_t((s0 + bw) / T) * (T / PK) * WAVE : nullptr;
            if constexpr (TR == TR_ONE) tacc = 0;  // (steps before the strip's first leave their codes zero)
            float4 cur[T + 1];
            if (l == 0) {
                const float4 *src = in_ring + ((s0 + bw) & (RING - 1));
#pragma unroll
                for (int t = 0; t < T; ++t) cur[t] = src[t];
                if constexpr (FORM == F_TABLE) cur[T] = in_ring[(s0 + bw + T) & (RING - 1)];
            }
#define PA_LSTEP(t)                                                                                       \
    if constexpr ((t) < T) step(s0 + (t), tbq, integral_constant<int, (t) % T>{}, cur[(t) % T],              \
                                cur[(t) + 1 <= T ? (t) + 1 : T], checked)
            PA_LSTEP(0); PA_LSTEP(1); PA_LSTEP(2); PA_LSTEP(3); PA_LSTEP(4); PA_LSTEP(5); PA_LSTEP(6); PA_LSTEP(7);
            PA_LSTEP(8); PA_LSTEP(9); PA_LSTEP(10); PA_LSTEP(11); PA_LSTEP(12); PA_LSTEP(13); PA_LSTEP(14);
            PA_LSTEP(15);
#undef PA_LSTEP
            // lane k holds step T - 1 - k's output (column s0 + T - 1 - k - 63)
            if (l < T) {
                if (wave + 1 < LW) {
                    *reinterpret_cast<float2 *>(out_ring + obase + T - 1 - l) = make_float2(oD, oP);
                } else if (to_band) {
                    const int c = s0 + T - 1 - l - (WAVE - 1);
                    if (c >= 0 && c < lt) {
                        const uint64_t tag = uint64_t(c + 1) << 32;
                        __hip_atomic_store(xb_out + 2 * c, tag | __float_as_uint(oD), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(xb_out + 2 * c + 1, tag | __float_as_uint(oP), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
            }
        };
        __syncthreads();  // the loader's first ring entries
        for (int p = 0; p < phases; ++p) {
            const int k = p - first;  // this wave's local phase
            if (has && k >= 0 && k < kmax) {
                const int s0 = k * T - bw;
                const int obase = (k * T + ooff) & (RING - 1);
                // every lane on a column inside the problem and before the
                // last one (whose cells are seeds): no checks
                using std::integral_constant;
                const bool lo = s0 < WAVE - 1, hi = s0 + T >= lt;
                if (!lo && !hi) phase_steps(s0, obase, integral_constant<int, 0>{});
                else if (!hi) {
                    if (s0 >= 0) phase_steps(s0, obase, integral_constant<int, LO>{});
                    else phase_steps(s0, obase, integral_constant<int, LO | NEG>{});
                } else if (!lo) {
                    phase_steps(s0, obase, integral_constant<int, HI>{});
                } else {
                    phase_steps(s0, obase, integral_constant<int, LO | HI | NEG>{});
                }
            }
            __syncthreads();
        }

        // ---- this strip's part of the seed search ------------------------------
        if (has) {
            int bi = bj >= 0 ? row : -1;
            if constexpr (LOC == LOC_LOCAL) {
                for (int m = 1; m < WAVE; m <<= 1) {
                    const float ob = __shfl_xor(best, m, WAVE);
                    if constexpr (TR == TR_SCORE) {
                        best = better<DIR>(ob, best) ? ob : best;
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
            } else if constexpr (LOC == LOC_GLOBAL) {
                gscore = __shfl(gscore, l_last, WAVE);
            } else {
                rbest = __shfl(rbest, l_last, WAVE);
                rj = __shfl(rj, l_last, WAVE);
                for (int m = 1; m < WAVE; m <<= 1) {
                    const float oc = __shfl_xor(cbest, m, WAVE);
                    const int oi = __shfl_xor(ci, m, WAVE);
                    if (better<DIR>(oc, cbest) || (oc == cbest && oi >= 0 && (ci < 0 || oi < ci))) {
                        cbest = oc;
                        ci = oi;
                    }
                }
            }
            if (l == 0) {
                BandPartial bp;
                bp.v1 = LOC == LOC_LOCAL ? best : (LOC == LOC_GLOBAL ? gscore : cbest);
                bp.i1 = LOC == LOC_LOCAL ? bi : ci;
                bp.j1 = bj;
                bp.v2 = st == st_last ? rbest : W;
                bp.j2 = st == st_last ? rj : -1;
                bp.pad = 0;
                a.partials[pr.band + st] = bp;
            }
        }
    }
}

template <int LOC, int DIR, int GAP>
static hipError_t dispatch_long_lg(int form, int tr, bool mx, const KArgs &a, const LongArgs &la, int grid,
                                   size_t lds, hipStream_t st) {
#define PA_LONG(FM, TRM, MXM)                                                                                  \
    hipLaunchKernelGGL((pa_long_kernel<LOC, DIR, GAP, FM, TRM, MXM>), dim3(grid), dim3((LONG_WAVES + 1) * WAVE), \
                       lds, st, a, la)
    if (form == F_TABLE) {
        if (tr == TR_SCORE) {
            if (mx) PA_LONG(F_TABLE, TR_SCORE, true);
            else PA_LONG(F_TABLE, TR_SCORE, false);
        } else {
            if (mx) PA_LONG(F_TABLE, TR_ONE, true);
            else PA_LONG(F_TABLE, TR_ONE, false);
        }
    } else if (form == F_BINARY) {
        if (tr == TR_SCORE) {
            if (mx) PA_LONG(F_BINARY, TR_SCORE, true);
            else PA_LONG(F_BINARY, TR_SCORE, false);
        } else {
            if (mx) PA_LONG(F_BINARY, TR_ONE, true);
            else PA_LONG(F_BINARY, TR_ONE, false);
        }
    } else {
        return hipErrorInvalidValue;
    }
#undef PA_LONG
    return hipGetLastError();
}

}  // namespace pa
