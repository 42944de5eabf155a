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
// symbol -- come from an LDS ring filled by the wave above (its lane 63
// writes (D, P, symbol) of each column it finishes).  The waves of a
// workgroup run in phases of LONG_T steps separated by workgroup barriers;
// wave w runs LONG_DW phases behind wave w-1, so every ring entry a wave
// reads was written in an earlier phase (the barrier orders the LDS write
// before the read).  Wave 0 fills its own ring from the strip border (the
// first band) or from the band above: that band's last wave stores each
// column's bottom (D, P) to global memory as two 8-byte words tagged with
// the column (relaxed atomics at device scope: no cache flush), and wave 0
// reads them LONG_K phases before they are due, re-reading (bounded) any
// whose tag is not there yet.  The tagged buffer is cleared before the
// launch.
//
// Cells.  The same fold as the general kernel in the reference's push order
// with strict improvement (pyalign/algorithm/solver.h:1311-1427; linear
// 2837-2854, Gotoh 2989-3035), so values, traceback bits (the general
// kernel's layout: the walk kernel reads them unchanged) and seeds are
// identical.  Each strip leaves its part of the seed search in a band-mode
// partial that pa_band_combine_kernel folds in strip order.
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

// MX: the strict compare-and-select of every candidate fold done as max
// (maximising) or min; the host chooses it for score-only solves where it
// returns the same bits: similarities finite and never -0, and (maximising
// global problems) gap costs above zero, so that no candidate is -0 or NaN
// (x + y is -0 only when both are -0) and max / min keep the strict fold's
// value.
template <int LOC, int DIR, int GAP, int FORM, int TR, bool MX>
__global__ __launch_bounds__((LONG_WAVES + 1) * WAVE) void pa_long_kernel(KArgs a, LongArgs la) {
    constexpr int R = LONG_R;
    constexpr int T = LONG_T;
    constexpr int K = LONG_K;
    constexpr int RING = LONG_RING;
    constexpr float W = worst<DIR>();
    constexpr int BITS = TR == TR_ONE ? (GAP == GAP_AFF ? 4 : 2) : 0;
    constexpr int WPC = TR == TR_ONE ? (BITS * R + 31) / 32 : 1;
    static_assert(FORM == F_TABLE || FORM == F_BINARY, "long mode reads table codes or binary symbols");
    static_assert(!MX || TR == TR_SCORE, "max folds carry no traceback bits");

    extern __shared__ float pa_dyn_lds[];
    // ring[w][c % RING]: (D, P, symbol bits, -) of the row above compute wave
    // w's strip at column c; then the similarity table (F_TABLE)
    float4 *ring = reinterpret_cast<float4 *>(pa_dyn_lds);
    float *tab = pa_dyn_lds + LONG_WAVES * RING * 4;
    __shared__ int s_ticket;
    if constexpr (FORM == F_TABLE) {
        const int n = a.K1 * a.K1;
        for (int k = threadIdx.x; k < n; k += blockDim.x) tab[k] = a.table[k];
    }
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE);
    const int l = threadIdx.x & (WAVE - 1);
    const bool loader = wave == LONG_WAVES;

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
            nst = (pr.len_s + WAVE * R - 1) / (WAVE * R);
            const int nbands = (nst + LONG_WAVES - 1) / LONG_WAVES;
            if (band < nbands || q + 1 >= la.n) break;
            band -= nbands;
            ++q;
        }
        const int lt = pr.len_t;
        const int nsp = lt + WAVE - 1;  // steps per strip
        // every wave runs the same number of phases (workgroup barriers), a
        // multiple of K (the loader's unrolled read-ahead)
        const int phases = ((LONG_WAVES - 1) * LONG_DW + (nsp + T - 1) / T + K - 1) / K * K;
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
                // of phase k - 1 (phase 0's at its own start): lane l < T
                // stages column due(p) = (p + 2) T + l at phase p
                const bool mine = l < T;
                auto due = [&](int p) { return (p + 2) * T + (l < T ? l : T - 1); };
                {   // phases 0 and 1: columns [0, 2 T)
                    const bool m0 = l < 2 * T;
                    uint64_t d, p;
                    uint32_t sy;
                    issue(l, d, p, sy);
                    stage(l, m0, d, p, sy);
                }
                uint64_t hD[K], hP[K];
                uint32_t hs[K];
#pragma unroll
                for (int u = 0; u < K; ++u) issue(due(u), hD[u], hP[u], hs[u]);
                __syncthreads();  // the compute waves' first ring entries
                for (int p0 = 0; p0 < phases; p0 += K) {
#pragma unroll
                    for (int u = 0; u < K; ++u) {
                        stage(due(p0 + u), mine, hD[u], hP[u], hs[u]);
                        issue(due(p0 + u + K), hD[u], hP[u], hs[u]);
                        __syncthreads();
                    }
                }
            };
            if (band > 0) run(std::true_type{});
            else run(std::false_type{});
            continue;
        }

        // ---- a compute wave: strip st ---------------------------------------
        const int ls = pr.len_s;
        const int st = band * LONG_WAVES + wave;
        const bool has = st < nst;
        const int row0 = st * WAVE * R + l * R;
        const int last = ls - 1;
        const int st_last = last / (WAVE * R);
        const int l_last = (last % (WAVE * R)) / R;
        const int r_last = last % R;
        uint64_t *xb_out = xb + 2 * int64_t(band) * lt;
        const bool out_x = has && wave == LONG_WAVES - 1 && st + 1 < nst;
        const float4 *my_ring = ring + wave * RING;
        float4 *next_ring = ring + (wave + 1) * RING;
        const float go_s = a.go_s, ge_s = a.ge_s, go_t = a.go_t, ge_t = a.ge_t;
        [[maybe_unused]] const int PAD = a.K1 - 1;

        int rowsym[R];
        bool rv[R];
        float DlA[R], DlB[R], Ql[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row0 + r;
            rv[r] = has && i < ls;
            if constexpr (FORM == F_TABLE) {
                rowsym[r] = (rv[r] ? int(((const uint8_t *)a.codes_s)[pr.off_s + i]) : PAD) * a.K1;
            } else {
                rowsym[r] = rv[r] ? int(((const uint32_t *)a.codes_s)[pr.off_s + i]) : 0;
            }
            DlA[r] = rv[r] ? a.border_s[i + 1] : W;  // D(i, -1)
            DlB[r] = DlA[r];
            Ql[r] = W;  // Q(i, -1)
        }
        // this lane's bottom row as the next lane sees it one step later; at
        // the start D(row0 + R - 1, -1), the next lane's diagonal at j = 0
        float botD = DlA[R - 1], botP = W;
        // D(row0 - 1, -1); lanes below the last row (all rows padding) take W
        float prevD = (row0 == 0) ? a.border_t[0] : (has && row0 <= ls ? a.border_s[row0] : W);
        int msym = 0;                                 // the column symbol this lane used last step
        // ring entries of this phase's and the next phase's columns, lane t
        // holding column k T + t of local phase k (lane 0 of the strip takes
        // its inputs from them with readlane)
        float4 cur = make_float4(W, W, 0.0f, 0.0f), nxt = cur;
        [[maybe_unused]] float Snext[R];             // F_TABLE: next step's similarities
#pragma unroll
        for (int r = 0; r < R; ++r) Snext[r] = 0.0f;
        float best = W;
        int bi = -1, bj = -1;
        float rbest = W, cbest = W;
        int rj = -1, ci = -1;
        float gscore = W;
        [[maybe_unused]] uint32_t *tbp =
            TR == TR_ONE ? a.tb + pr.tb_off + int64_t(st) * nsp * WAVE * WPC + l * WPC : nullptr;

        // the fold of one direction: strict improvement keeps the first
        // candidate on ties (the old value wins), or max / min where exact
        auto opt = [](float x, float y) {
            if constexpr (DIR == DIR_MAX) return fmaxf(x, y);
            else return fminf(x, y);
        };
        // binary similarities held in registers (one select per cell)
        float veq = a.eq, vne = a.ne;
        asm volatile("" : "+v"(veq), "+v"(vne));
        const bool lane_out = l == WAVE - 1;  // the lane whose bottom row goes down
        const bool out_ring = has && wave + 1 < LONG_WAVES;

        // R cells of column j: Dd is D(row0 - 1, j - 1), (Dup, Pup) the row
        // above at j; Din holds D(i, j - 1), Dout gets D(i, j)
        auto cells = [&](int j, float Dd, float Dup, float Pup, const float (&S)[R], float (&Din)[R], float (&Dout)[R],
                         auto checked) -> uint32_t {
            constexpr bool CHK = decltype(checked)::value;
            uint32_t tbw = 0;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float D, P = 0.0f, Q = 0.0f;
                uint32_t bits = 0;
                const float dg = Dd + S[r];
                if constexpr (MX) {
                    if constexpr (GAP == GAP_LIN) {
                        D = opt(opt(dg, Dup + go_s), Din[r] + go_t);
                    } else {
                        P = opt(Dup + go_s, Pup + ge_s);
                        Q = opt(Din[r] + go_t, Ql[r] + ge_t);
                        D = opt(opt(dg, P), Q);
                    }
                    if constexpr (LOC == LOC_LOCAL) D = opt(D, 0.0f);
                } else if constexpr (GAP == GAP_LIN) {
                    const float up = Dup + go_s;
                    const float lf = Din[r] + go_t;
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
                    const float po = Dup + go_s, pe = Pup + ge_s;
                    const float qo = Din[r] + go_t, qe = Ql[r] + ge_t;
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
                    const int i = row0 + r;
                    if constexpr (TR == TR_SCORE) {
                        best = MX ? opt(best, D) : (better<DIR>(D, best) ? D : best);  // padded rows stay below
                    } else if (!CHK || rv[r]) {
                        // Optima over i desc, j desc (solver.h:2050-2056)
                        if (better<DIR>(D, best) || (D == best && i >= bi)) {
                            best = D;
                            bi = i;
                            bj = j;
                        }
                    }
                }
                if constexpr (TR == TR_ONE) tbw |= bits << (r * BITS);
                Dd = Din[r];
                Dout[r] = D;
                Ql[r] = Q;
                Dup = D;
                Pup = P;
            }
            botD = Dup;
            botP = Pup;
            return tbw;
        };

        auto lane_in = [](const float4 &e, int t, int k) {  // word k of lane t's entry
            const float x = k == 0 ? e.x : (k == 1 ? e.y : e.z);
            return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), t));
        };
        // one step: column j = s - l; t = s mod T (its ring entry is lane t's of cur)
        auto step = [&](int s, auto tc, float (&Din)[R], float (&Dout)[R], auto checked) {
            constexpr bool CHK = decltype(checked)::value;
            constexpr int t = decltype(tc)::value;
            const int j = s - l;
            const float inD = shr1_in(botD, lane_in(cur, t, 0));
            const float inP = shr1_in(botP, lane_in(cur, t, 1));
            const int tsym = shr1_in_i(msym, __float_as_int(lane_in(cur, t, 2)));
            msym = tsym;
            float S[R];
            if constexpr (FORM == F_TABLE) {
#pragma unroll
                for (int r = 0; r < R; ++r) S[r] = Snext[r];
                // the next step's symbol is the previous lane's current one
                const float ns = t + 1 < T ? lane_in(cur, (t + 1) % T, 2) : lane_in(nxt, 0, 2);
                const int nsym = shr1_in_i(tsym, __float_as_int(ns));
#pragma unroll
                for (int r = 0; r < R; ++r) Snext[r] = tab[rowsym[r] + min(nsym, PAD)];
            } else {
#pragma unroll
                for (int r = 0; r < R; ++r) S[r] = (rowsym[r] == tsym) ? veq : vne;
                if constexpr (CHK) {
#pragma unroll
                    for (int r = 0; r < R; ++r) S[r] = rv[r] ? S[r] : W;
                }
            }
            const float Ddiag = prevD;
            prevD = inD;
            if constexpr (!CHK) {
                // a steady step: every lane on a column before the last, every
                // row real and no seed row (the strip is not the last)
                [[maybe_unused]] const uint32_t tbw = cells(j, Ddiag, inD, inP, S, Din, Dout, checked);
                if constexpr (TR == TR_ONE) tbp[int64_t(s) * WAVE * WPC] = tbw;
                const int jo = s - (WAVE - 1);  // lane 63's column
                if (lane_out) {
                    if (out_ring) {
                        next_ring[jo & (RING - 1)] = make_float4(botD, botP, __uint_as_float(uint32_t(tsym)), 0.0f);
                    } else if (out_x) {
                        const uint64_t tag = uint64_t(jo + 1) << 32;
                        __hip_atomic_store(xb_out + 2 * jo, tag | __float_as_uint(botD), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                        __hip_atomic_store(xb_out + 2 * jo + 1, tag | __float_as_uint(botP), __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
                    }
                }
                return;
            }
            if (has && j >= 0 && j < lt) {
                [[maybe_unused]] const uint32_t tbw = cells(j, Ddiag, inD, inP, S, Din, Dout, checked);
                if constexpr (TR == TR_ONE) tbp[int64_t(s) * WAVE * WPC] = tbw;
                if constexpr (LOC == LOC_SEMI) {
                    if (st == st_last && l == l_last) {
                        float v = Dout[0];
#pragma unroll
                        for (int r = 1; r < R; ++r) v = (r == r_last) ? Dout[r] : v;
                        if (better<DIR>(v, rbest)) { rbest = v; rj = j; }  // last row, j ascending (solver.h:2285-2287)
                    }
                    if (j == lt - 1) {
#pragma unroll
                        for (int r = 0; r < R; ++r)  // last column, i ascending (solver.h:2289-2291)
                            if (rv[r] && better<DIR>(Dout[r], cbest)) { cbest = Dout[r]; ci = row0 + r; }
                    }
                }
                if constexpr (LOC == LOC_GLOBAL) {
                    if (st == st_last && l == l_last && j == lt - 1) {
                        float v = Dout[0];
#pragma unroll
                        for (int r = 1; r < R; ++r) v = (r == r_last) ? Dout[r] : v;
                        gscore = v;
                    }
                }
            }
            // lane 63 hands its bottom row (column s - 63) down
            if (lane_out && j >= 0 && j < lt && has) {
                if (out_ring) {
                    next_ring[j & (RING - 1)] = make_float4(botD, botP, __uint_as_float(uint32_t(tsym)), 0.0f);
                } else if (out_x) {
                    const uint64_t tag = uint64_t(j + 1) << 32;
                    __hip_atomic_store(xb_out + 2 * j, tag | __float_as_uint(botD), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                    __hip_atomic_store(xb_out + 2 * j + 1, tag | __float_as_uint(botP), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        };

        const int first = wave * LONG_DW;  // this wave's first phase
        // a strip whose rows are all real and that holds no seed row skips
        // the row and seed checks in its steady steps (every lane on a column
        // before the last one)
        const bool full = has && st < st_last;
        auto ring_entries = [&](int k) { return my_ring[(k * T + (l & (T - 1))) & (RING - 1)]; };
        __syncthreads();  // the loader's first ring entries
        for (int p = 0; p < phases; ++p) {
            const int k = p - first;  // this wave's local phase
            const int s0 = k * T;
            if (has && s0 + T > 0 && s0 < nsp) {
                if (k == 0) cur = ring_entries(0);
                else cur = nxt;
                nxt = ring_entries(k + 1);  // written before the last barrier
                if (k == 0) {
                    if constexpr (FORM == F_TABLE) {
                        const int nsym = shr1_in_i(0, __float_as_int(lane_in(cur, 0, 2)));
#pragma unroll
                        for (int r = 0; r < R; ++r) Snext[r] = tab[rowsym[r] + min(nsym, PAD)];
                    }
                }
                using std::integral_constant;
                auto phase_steps = [&](auto checked) {
                    step(s0 + 0, integral_constant<int, 0>{}, DlA, DlB, checked);
                    step(s0 + 1, integral_constant<int, 1>{}, DlB, DlA, checked);
                    step(s0 + 2, integral_constant<int, 2>{}, DlA, DlB, checked);
                    step(s0 + 3, integral_constant<int, 3>{}, DlB, DlA, checked);
                    if constexpr (T > 4) {
                        step(s0 + 4, integral_constant<int, 4 % T>{}, DlA, DlB, checked);
                        step(s0 + 5, integral_constant<int, 5 % T>{}, DlB, DlA, checked);
                        step(s0 + 6, integral_constant<int, 6 % T>{}, DlA, DlB, checked);
                        step(s0 + 7, integral_constant<int, 7 % T>{}, DlB, DlA, checked);
                    }
                    if constexpr (T > 8) {
                        step(s0 + 8, integral_constant<int, 8 % T>{}, DlA, DlB, checked);
                        step(s0 + 9, integral_constant<int, 9 % T>{}, DlB, DlA, checked);
                        step(s0 + 10, integral_constant<int, 10 % T>{}, DlA, DlB, checked);
                        step(s0 + 11, integral_constant<int, 11 % T>{}, DlB, DlA, checked);
                        step(s0 + 12, integral_constant<int, 12 % T>{}, DlA, DlB, checked);
                        step(s0 + 13, integral_constant<int, 13 % T>{}, DlB, DlA, checked);
                        step(s0 + 14, integral_constant<int, 14 % T>{}, DlA, DlB, checked);
                        step(s0 + 15, integral_constant<int, 15 % T>{}, DlB, DlA, checked);
                    }
                };
                if (full && s0 >= WAVE - 1 && s0 + T < lt) phase_steps(integral_constant<bool, false>{});
                else phase_steps(integral_constant<bool, true>{});
            }
            __syncthreads();
        }

        // ---- this strip's part of the seed search ------------------------------
        if (has) {
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
            PA_LONG(F_TABLE, TR_ONE, false);
        }
    } else if (form == F_BINARY) {
        if (tr == TR_SCORE) {
            if (mx) PA_LONG(F_BINARY, TR_SCORE, true);
            else PA_LONG(F_BINARY, TR_SCORE, false);
        } else {
            PA_LONG(F_BINARY, TR_ONE, false);
        }
    } else {
        return hipErrorInvalidValue;
    }
#undef PA_LONG
    return hipGetLastError();
}

}  // namespace pa
