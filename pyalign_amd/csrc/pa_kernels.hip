// pa_kernels.hip -- traceback walk, small helper kernels and the top-level
// dispatch into the per-locality instantiation units (pa_k_*.hip).
#include "pa_kernels.h"

#include <algorithm>
#include <mutex>

namespace pa {

hipError_t dispatch_local_max(const LaunchSpec &, const KArgs &, hipStream_t, int *);
hipError_t dispatch_local_min(const LaunchSpec &, const KArgs &, hipStream_t, int *);
hipError_t dispatch_global_max(const LaunchSpec &, const KArgs &, hipStream_t, int *);
hipError_t dispatch_global_min(const LaunchSpec &, const KArgs &, hipStream_t, int *);
hipError_t dispatch_semi_max(const LaunchSpec &, const KArgs &, hipStream_t, int *);
hipError_t dispatch_semi_min(const LaunchSpec &, const KArgs &, hipStream_t, int *);

hipError_t dispatch_long_local_max(int, int, int, bool, const KArgs &, const LongArgs &, int, size_t, hipStream_t);
hipError_t dispatch_long_local_min(int, int, int, bool, const KArgs &, const LongArgs &, int, size_t, hipStream_t);
hipError_t dispatch_long_global_max(int, int, int, bool, const KArgs &, const LongArgs &, int, size_t, hipStream_t);
hipError_t dispatch_long_global_min(int, int, int, bool, const KArgs &, const LongArgs &, int, size_t, hipStream_t);
hipError_t dispatch_long_semi_max(int, int, int, bool, const KArgs &, const LongArgs &, int, size_t, hipStream_t);
hipError_t dispatch_long_semi_min(int, int, int, bool, const KArgs &, const LongArgs &, int, size_t, hipStream_t);

bool long_supported(int gap, int form, int tr) {
    return (gap == GAP_LIN || gap == GAP_AFF) && (form == F_TABLE || form == F_BINARY) &&
           (tr == TR_SCORE || tr == TR_ONE);
}

size_t long_lds_bytes(int form, int K1) {
    return size_t(LONG_WAVES) * LONG_RING * 16 + (form == F_TABLE ? size_t(K1) * K1 * 4 : 0);
}

hipError_t launch_long(int loc, int dir, int gap, int form, int tr, bool fold_max, const KArgs &a,
                       const LongArgs &la, int grid, hipStream_t st) {
    if (!long_supported(gap, form, tr) || la.n <= 0 || grid <= 0) return hipErrorInvalidValue;
    const size_t lds = long_lds_bytes(form, a.K1);
    const bool mx = dir == DIR_MAX;
    switch (loc) {
        case LOC_LOCAL:
            return mx ? dispatch_long_local_max(gap, form, tr, fold_max, a, la, grid, lds, st)
                      : dispatch_long_local_min(gap, form, tr, fold_max, a, la, grid, lds, st);
        case LOC_GLOBAL:
            return mx ? dispatch_long_global_max(gap, form, tr, fold_max, a, la, grid, lds, st)
                      : dispatch_long_global_min(gap, form, tr, fold_max, a, la, grid, lds, st);
        case LOC_SEMI:
            return mx ? dispatch_long_semi_max(gap, form, tr, fold_max, a, la, grid, lds, st)
                      : dispatch_long_semi_min(gap, form, tr, fold_max, a, la, grid, lds, st);
    }
    return hipErrorInvalidValue;
}

static hipError_t dispatch(const LaunchSpec &s, const KArgs &a, hipStream_t st, int *occ) {
    const bool mx = s.dir == DIR_MAX;
    switch (s.loc) {
        case LOC_LOCAL: return mx ? dispatch_local_max(s, a, st, occ) : dispatch_local_min(s, a, st, occ);
        case LOC_GLOBAL: return mx ? dispatch_global_max(s, a, st, occ) : dispatch_global_min(s, a, st, occ);
        case LOC_SEMI: return mx ? dispatch_semi_max(s, a, st, occ) : dispatch_semi_min(s, a, st, occ);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_solve(const LaunchSpec &s, const KArgs &a, hipStream_t st) {
    return dispatch(s, a, st, nullptr);
}

int occupancy_waves(const LaunchSpec &s, int dim) {
    int waves = 0;
    KArgs dummy{};
    dummy.dim = dim;
    if (dispatch(s, dummy, nullptr, &waves) != hipSuccess || waves <= 0) return 4;
    return waves;
}

// F_TABLE_G (tables read from global memory) is built at R = 8 only
int band_R(int form, int tr) {
    const bool r4 = (form == F_TABLE && (tr == TR_SCORE || tr == TR_ONE)) ||
                    (form == F_BINARY && (tr == TR_SCORE || tr == TR_ONE));
    return r4 ? 4 : 8;
}

bool supported_R(int form, int tr, bool fastmax, int R) {
    if (form == F_TABLE && tr == TR_SCORE) {
        if (fastmax) {
            switch (R) {
                case 1: case 2: case 3: case 4: case 5: case 6: case 7: case 8:
                case 10: case 12: case 14: case 16: case 20: case 24: case 28: case 32:
                    return true;
                default: return false;
            }
        }
        return R == 4 || R == 8 || R == 16 || R == 32;
    }
    return R == 8;
}

// ---------------------------------------------------------------------------
// Traceback walks.  One wavefront follows one problem's path: the walk is
// sequential, so the wave's 64 lanes do not split it; instead they load the
// traceback words the walk is about to read, a tile at a time, into LDS
// with coalesced loads, and every lane then steps through the tile in
// lockstep (same state in every lane, made scalar with readfirstlane).  A
// path only moves up and left, so a tile that ends at the current cell
// (steps and lanes at or before it, same strip) holds the next stretch.
// Bit layout per problem: strip st, step s, lane k, word w at
// ((st * (len_t + L - 1) + s) * L + k) * words + w; cell (i, j) is lane
// k = (i mod L*R) / R, row r = i mod R of that lane, step s = j + k.
// ---------------------------------------------------------------------------
// LDS per wavefront: a traceback tile of `tile_words` words (host-chosen:
// small tiles and 4 wavefronts per block for many problems, one wavefront
// with a large tile for a few long ones).
#ifndef PA_WALK_KT_MAX
#define PA_WALK_KT_MAX WAVE
#endif
#ifndef PA_WALK_ST_MUL
#define PA_WALK_ST_MUL 1
#endif
#ifndef PA_WALK_LOADS
#define PA_WALK_LOADS 32
#endif
struct WalkShape {
    int waves;       // wavefronts (problems) per block
    int tile_words;  // traceback words per wavefront tile
};

inline int device_cus() {
    static int cus = 0;
    if (cus <= 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    return cus;
}

// extra: LDS words a wavefront needs besides its tile (series windows).
// Few problems: one wavefront per block with a large tile.  Otherwise four
// per block, the tile as large as the wavefronts resident per CU allow
// (larger tiles: fewer refills along a path).
#ifndef PA_WALK_TILE_MAX
#define PA_WALK_TILE_MAX 8192
#endif
inline WalkShape walk_shape(int64_t n, int extra) {
    if (n < 512) return WalkShape{1, 14336};
    const int64_t per_cu = (n + device_cus() - 1) / device_cus();
    int tile = per_cu <= 4 ? 8192 : (per_cu <= 8 ? 4096 : 2048);
    tile = tile < PA_WALK_TILE_MAX ? tile : PA_WALK_TILE_MAX;
    while (tile > 2048 && size_t(4) * (tile + extra) * 4 > size_t(160) * 1024) tile /= 2;
    return WalkShape{4, tile};
}

// A position (i, j) on one problem's traceback with its decomposition
// (strip st, lane k, bit position of row r in the lane's words, step
// s = j + k), moved up and left without divisions or multiplies, and the
// LDS tile the codes are read from (lanes [k0, k1), steps [s0, s1) of strip
// t_st, step-major).  li is the current cell's first word in the tile,
// kept up to date by the moves and recomputed when a tile is loaded.  ga is
// the cell's offset in a dense payload with strides (gsi, gsj).
struct TbWalker {
    uint32_t *lds;
    const uint32_t *tb;
    int64_t tb_off;
    int nsteps;
    int L, R, wpc, bits, KT, ST;
    int t_st, k0, k1;
    int s0, s1;
    int kww;  // words per tile step: (k1 - k0) * wpc
    int i, j, st, k, bit;
    int s, li;
    int lgR, lgrps;  // log2 R and log2 (L R): strip heights and lane counts are powers of two
    int64_t ga, gsi, gsj;

    __device__ void init(uint32_t *l, int tile_words, const uint32_t *t, const Prob &pr, int b) {
        lds = l;
        tb = t;
        tb_off = pr.tb_off;
        L = pr.L;
        R = pr.R;
        bits = b;
        wpc = (b * R + 31) / 32;
        lgR = __builtin_ctz(unsigned(R));
        lgrps = __builtin_ctz(unsigned(L * R));
        nsteps = pr.len_t + L - 1;
        // tile shape: a path going diagonally leaves a lane every R rows, one
        // step more than R, so tiles hold about R + 1 steps per lane
        KT = L < PA_WALK_KT_MAX ? L : PA_WALK_KT_MAX;
        while (KT > 1 && KT * KT * wpc * (R + 1) > tile_words) KT >>= 1;
        // steps: what a diagonal path uses crossing the tile's lanes (a
        // refill costs in proportion to the tile's size; measured against
        // twice that and against the whole tile)
        ST = tile_words / (KT * wpc);
        ST = ST < PA_WALK_ST_MUL * (R + 1) * KT ? ST : PA_WALK_ST_MUL * (R + 1) * KT;
        t_st = -1;
        k0 = k1 = 0;
        s0 = s1 = 0;
        kww = 0;
        li = 0;
        gsi = pr.si;
        gsj = pr.sj;
    }
    __device__ void seek(int i_, int j_) {
        i = i_;
        j = j_;
        const int rps = L * R;
        st = i >= 0 ? i / rps : -1;
        const int rem = i - st * rps;
        k = rem / R;
        bit = (rem - k * R) * bits;
        s = j + k;
        t_st = -1;  // the next code() loads a tile around this cell
        ga = int64_t(i) * gsi + int64_t(j) * gsj;
    }
    __device__ void up() {
        --i;
        ga -= gsi;
        bit -= bits;
        if (bit < 0) {  // row R-1 of the lane above, one step earlier
            bit = (R - 1) * bits;
            --s;
            li -= kww + wpc;
            if (--k < 0) {  // into the previous strip (the tile is reloaded)
                k = L - 1;
                s += L;
                --st;
            }
        }
    }
    __device__ void left() {
        --j;
        --s;
        li -= kww;
        ga -= gsj;
    }
    // moves to (ni, nj), up and left of the current cell, keeping the tile
    // (the next code() loads another when the cell lies outside it)
    __device__ void jump(int ni, int nj) {
        i = ni;
        j = nj;
        st = i >> lgrps;  // (-1 above the matrix)
        const int rem = i & ((1 << lgrps) - 1);
        k = rem >> lgR;
        bit = (rem & (R - 1)) * bits;
        s = j + k;
        ga = int64_t(i) * gsi + int64_t(j) * gsj;
        li = (s - s0) * kww + (k - k0) * wpc;
    }
    // this lane's cell of a run: lane m, the cell m moves of di rows up and
    // dj columns left from the current one; ok when it lies in the matrix
    // and the loaded tile (which holds the current cell)
    __device__ uint32_t gather(int di, int dj, int lane, bool &ok) const {
        const int ii = i - di * lane, jj = j - dj * lane;
        const int rem = ii & ((1 << lgrps) - 1);
        const int k2 = rem >> lgR;
        const int s2 = jj + k2;
        ok = ii >= 0 && jj >= 0 && (ii >> lgrps) == t_st && k2 >= k0 && s2 >= s0;
        const int b2 = (rem & (R - 1)) * bits;
        const uint32_t wd = ok ? lds[(s2 - s0) * kww + (k2 - k0) * wpc + (b2 >> 5)] : 0u;
        return (wd >> (b2 & 31)) & ((1u << bits) - 1u);
    }
    // traceback code of the current cell (i, j >= 0; the same in every lane)
    __device__ uint32_t code(int lane) {
        if (st != t_st || k < k0 || s < s0) {
            t_st = st;
            k1 = k + 1;
            k0 = k1 > KT ? k1 - KT : 0;
            s1 = s + 1;
            s0 = s1 > ST ? s1 - ST : 0;
            kww = (k1 - k0) * wpc;
            li = (s - s0) * kww + (k - k0) * wpc;
            const int kw = kww;
            const int cnt = (s1 - s0) * kw;
            const uint32_t *base = tb + tb_off + (int64_t(t_st) * nsteps + s0) * L * wpc + int64_t(k0) * wpc;
            const int sstride = L * wpc;
            // w / kw as w * (1 / kw) in float (w < 2^24: exact operands),
            // corrected by one step where the rounded product falls short
            const float rkw = 1.0f / float(kw);
            __builtin_amdgcn_wave_barrier();  // the old tile's reads are done
            // PA_WALK_LOADS loads in flight per lane before their LDS stores
            constexpr int U = PA_WALK_LOADS;
            for (int w0 = 0; w0 < cnt; w0 += U * WAVE) {
                uint32_t v[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int w = w0 + u * WAVE + lane;
                    int si = int(float(w) * rkw);
                    si += (si + 1) * kw <= w ? 1 : 0;
                    si -= si * kw > w ? 1 : 0;
                    v[u] = w < cnt ? base[int64_t(si) * sstride + (w - si * kw)] : 0u;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int w = w0 + u * WAVE + lane;
                    if (w < cnt) lds[w] = v[u];
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        const uint32_t w = lds[li + (bit >> 5)];
        return __builtin_amdgcn_readfirstlane((w >> (bit & 31)) & ((1u << bits) - 1u));
    }
};

// One move along a one-optimal path (TracebackIterators<false>::Iterator
// ::next, solver.h:1498-1536); returns the code's source (0: a local start,
// D(i, j) not above zero, no move).  Affine gaps are single jumps to the
// gap's origin found through the P/Q extend bits, because P and Q copy
// their predecessor's traceback (solver.h:2999-3016).
template <int GAP>
__device__ __forceinline__ int walk_step(TbWalker &w, int lane) {
    const int src = int(w.code(lane) & 3u);
    if (src == 1) {
        w.up();
        w.left();
    } else if (src == 2) {
        if constexpr (GAP == GAP_AFF) {
            while (w.i > 0 && (w.code(lane) & 4u)) w.up();
        }
        w.up();
    } else if (src == 3) {
        if constexpr (GAP == GAP_AFF) {
            while (w.j > 0 && (w.code(lane) & 8u)) w.left();
        }
        w.left();
    }
    return src;
}

// A walk's path cells, held one per lane (cell c in lane c % 64) and stored
// 64 at a time with one coalesced store, so that no move waits on a store of
// its own (the wait counter is shared with the tile loads).  Cell 0, the
// seed, sits in lane 0.
struct PathCells {
    int32_t *P;
    int pi, pj;  // this lane's cell
    int cnt;     // cells so far

    __device__ void init(int32_t *p, int u, int v) {
        P = p;
        pi = u;
        pj = v;
        cnt = 1;
    }
    __device__ void spill(int upto, int lane) {  // cells [upto & ~63, upto)
        const int b0 = (upto - 1) & ~(WAVE - 1);
        if (b0 + lane < upto) {
            P[2 * (b0 + lane)] = pi;
            P[2 * (b0 + lane) + 1] = pj;
        }
    }
    // cells cnt .. cnt + n - 1: (i0 - di (t + 1), j0 - dj (t + 1)), t < n
    __device__ void record(int n, int i0, int j0, int di, int dj, int lane) {
        int t0 = 0;
        while (t0 < n) {
            const int slot = cnt & (WAVE - 1);
            const int take = n - t0 < WAVE - slot ? n - t0 : WAVE - slot;
            const int t = t0 + (lane - slot) + 1;
            if (lane >= slot && lane < slot + take) {
                pi = i0 - di * t;
                pj = j0 - dj * t;
            }
            cnt += take;
            t0 += take;
            if ((cnt & (WAVE - 1)) == 0) spill(cnt, lane);
        }
    }
    __device__ int finish(int lane) {  // stores the last cells, returns the path's length
        if (cnt & (WAVE - 1)) spill(cnt, lane);
        return cnt;
    }
};

// The walk of a long-mode problem (strips of 64 rows, one row per lane,
// traceback codes packed along the steps: long_pack(bits) steps per word,
// see long_strip_words): the moves of walk_step, taken a run at a time.
// Lane m reads the code of the cell m moves ahead along the run the path
// is in (a diagonal run: m rows up and m columns left; an affine gap run:
// m cells up or left), and a ballot of "the run goes on there" gives the
// run's length: one LDS round trip for up to 64 moves instead of one per
// move.  A diagonal run continues while the cells' codes say diagonal; a
// gap run (one path move, solver.h:2999-3016) while their extend bits are
// set.  Runs stop at the strip's top and at the tile's bottom and resume
// from there.  Codes come from a tile of whole word rows (64 words: one
// word index of the strip's lanes) loaded straight into LDS; the next
// strip's tile is loaded ahead into a second buffer once the path is in the
// strip's top half.  Returns the path's length; cells are stored as
// pa_walk_kernel stores them.
template <int GAP>
__device__ __forceinline__ int walk_long64(const Prob &pr, const uint32_t *tb, uint32_t *lds, int tile_words, int u,
                                           int v, int32_t *P, int lane) {
    constexpr int TB = GAP == GAP_AFF ? 4 : 2;
    constexpr int PK = long_pack(TB);
    constexpr int LGPK = PK == 16 ? 4 : (PK == 8 ? 3 : 2);
    static_assert((1 << LGPK) == PK, "steps per word: a power of two");
    constexpr uint32_t MASK = (1u << TB) - 1u;
    const int nsp = pr.len_t + WAVE - 1;
    const int64_t sw = long_strip_words(nsp, TB);
    // two tiles (the current one and the next strip's, loaded ahead) of QT
    // word rows: more than a diagonal path's crossing of a strip
    const int QT = tile_words / (2 * WAVE) < 64 ? tile_words / (2 * WAVE) : 64;
    int i = u, j = v;
    int st = i / WAVE;
    int base = st * WAVE;  // the strip's first row
    static_assert((LONG_WAVES & (LONG_WAVES - 1)) == 0, "strip st's wave: st & (LONG_WAVES - 1)");
    int uq = j + (i - base) + long_boff(st & (LONG_WAVES - 1));  // the step plus the strip's offset
    int cur = 0, nxt = QT * WAVE;       // the LDS offsets of the tile read and the one loaded ahead
    int t_st = -1, q0 = 0;              // cur: word rows [q0, ...] of strip t_st
    int n_st = -1, nq0 = 0, nq1 = -1;  // nxt: rows [nq0, nq1] of strip n_st
    // loads rows [q1 - QT + 1, q1] (from 0) of strip sq to LDS offset dst, returns the first
    auto issue = [&](int dst, int sq, int q1) -> int {
        const int qa = q1 - QT + 1 > 0 ? q1 - QT + 1 : 0;
        const uint4 *b4 = reinterpret_cast<const uint4 *>(tb + pr.tb_off + (int64_t(sq) * sw + qa) * WAVE);
        const int n4 = (q1 - qa + 1) * (WAVE / 4);
        for (int c = 0; c * WAVE < n4; ++c) {
            if (c * WAVE + lane < n4)
                __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)(b4 + c * WAVE + lane),
                                                 (__attribute__((address_space(3))) void *)(lds + dst + c * WAVE * 4), 16,
                                                 0, 0);
        }
        return qa;
    };
    PathCells cells;
    cells.init(P, u, v);
    // the code of the cell dk * lane rows up and du * lane steps back (for
    // this lane), and whether that cell lies in the strip and the tile
    auto gather = [&](int dk, int du, bool &ok) -> uint32_t {
        const int r = (i - base) - dk * lane;
        const int uu = uq - du * lane;
        const int q = uu >> LGPK;
        ok = r >= 0 && q >= q0;
        const uint32_t w = ok ? lds[cur + (q - q0) * WAVE + r] : 0u;
        return (w >> (TB * (uu & (PK - 1)))) & MASK;  // (PK may be below 32 / TB: 8-step phases)
    };
    // a run of n moves of di rows and dj columns (the last one may leave the strip)
    auto advance = [&](int n, int di, int dj) {
        i -= di * n;
        j -= dj * n;
        uq -= (di + dj) * n;
        if (i < base) {  // into the strip above
            base -= WAVE;
            --st;
            uq = j + (i - base) + long_boff(st & (LONG_WAVES - 1));
        }
    };
    int mode = 0;  // 0: at a cell; 2 / 3: inside a vertical / horizontal gap run (affine)
    while ((i | j) >= 0) {
        const int qc = uq >> LGPK;
        if (st != t_st || qc < q0) {  // a tile holding this cell's word row
            if (st == n_st && qc >= nq0 && qc <= nq1) {  // the one loaded ahead
                const int t = cur;
                cur = nxt;
                nxt = t;
                q0 = nq0;
                n_st = -1;
            } else {
                __builtin_amdgcn_wave_barrier();  // the old tile's reads are done
                q0 = issue(cur, st, qc);
            }
            t_st = st;
            asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // in the strip's top half: the strip above, rows up to the path's
            // latest possible entry (column j, row 63), loaded ahead
        }
        if (i - base < WAVE / 2 && st > 0 && n_st != st - 1) {
            n_st = st - 1;
            nq1 = (j + WAVE - 1 + long_boff(n_st & (LONG_WAVES - 1))) >> LGPK;
            nq0 = issue(nxt, n_st, nq1);
        }
        if (mode == 0) {
            // a diagonal run: lane m, the cell m moves on (inside the matrix: j - m >= 0)
            bool ok;
            const uint32_t c = gather(1, 2, ok);
            ok = ok && j - lane >= 0;
            const uint64_t run = __ballot(ok && (c & 3u) == 1u);
            const int n = ~run == 0 ? WAVE : __builtin_ctzll(~run);
            const uint64_t valid = __ballot(ok);
            if (n > 0) {
                cells.record(n, i, j, 1, 1, lane);
                advance(n, 1, 1);
            }
            // the cell the run stopped at, when it lies in this strip and tile
            if (n < WAVE && ((valid >> n) & 1u)) {
                const uint32_t cn = uint32_t(__builtin_amdgcn_readlane(int(c), n));
                const uint32_t src = cn & 3u;
                if (src == 0) break;  // a local start
                if constexpr (GAP == GAP_AFF) {
                    mode = int(src);  // the gap run starts at this cell
                } else {
                    cells.record(1, i, j, src == 2 ? 1 : 0, src == 3 ? 1 : 0, lane);
                    advance(1, src == 2 ? 1 : 0, src == 3 ? 1 : 0);
                }
            }
        } else {
            // a gap run: lane m, the cell m moves on; the run goes on while
            // the extend bit is set away from the border (solver.h:2999-3016)
            const bool vert = mode == 2;
            bool ok;
            const uint32_t c = gather(vert ? 1 : 0, 1, ok);
            const int pos = (vert ? i : j) - lane;
            const uint64_t ext = __ballot(ok && pos > 0 && (c & (vert ? 4u : 8u)) != 0u);
            const int n = ~ext == 0 ? WAVE : __builtin_ctzll(~ext);
            const uint64_t valid = __ballot(ok);
            if (n < WAVE && ((valid >> n) & 1u)) {  // the run's last cell: one more move ends it
                const int di = vert ? 1 : 0, dj = vert ? 0 : 1;
                cells.record(1, i - di * n, j - dj * n, di, dj, lane);
                advance(n + 1, di, dj);
                mode = 0;
            } else {  // the run goes on past this strip or tile
                advance(n, vert ? 1 : 0, vert ? 0 : 1);
            }
        }
    }
    return cells.finish(lane);
}

template <int GAP>
__global__ __launch_bounds__(256) void pa_walk_kernel(const Prob *probs, int64_t n, const uint32_t *tb,
                                                      const int32_t *seeds, int32_t *path,
                                                      const int64_t *path_off, int32_t *path_len,
                                                      int tile_words) {
    extern __shared__ uint32_t smem[];
    // the wave index read as a scalar: the problem, its descriptor and the
    // whole walk state are then wave-uniform (scalar registers and scalar
    // branches; the vector lanes only load tiles)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE), lane = threadIdx.x % WAVE;
    const int64_t k = int64_t(blockIdx.x) * (blockDim.x / WAVE) + wave;
    if (k >= n) return;  // whole wavefronts leave; no block-wide barriers follow
    const Prob pr = probs[k];
    const int out = pr.out;
    const int u = seeds[2 * out], v = seeds[2 * out + 1];
    if (u == NO_SEED) {
        if (lane == 0) path_len[out] = -1;
        return;
    }
    int32_t *P = path + 2 * path_off[out];
    if (pr.pack) {
        const int cnt = walk_long64<GAP>(pr, tb, smem + wave * tile_words, tile_words, u, v, P, lane);
        if (lane == 0) path_len[out] = cnt;
        return;
    }
    TbWalker w;
    w.init(smem + wave * tile_words, tile_words, tb, pr, GAP == GAP_AFF ? 4 : 2);
    w.seek(u, v);
    PathCells cells;
    cells.init(P, u, v);
    // The moves of walk_step a run at a time (as the dense backward takes
    // them): lane m reads the cell m moves on along the run the path is in
    // (cells of one code; or an affine gap run, whose cells extend it while
    // their extend bit is set), and a ballot gives the run's length.
    int mode = 0;  // affine: 2 / 3 inside a vertical / horizontal gap run
    while (w.i >= 0 && w.j >= 0) {
        const uint32_t c0 = w.code(lane);  // (loads a tile holding the current cell)
        if (mode == 0) {
            const int src = int(c0 & 3u);
            if (src == 0) break;  // a local start
            if constexpr (GAP == GAP_AFF) {
                if (src >= 2) {  // a gap run starts at this cell
                    mode = src;
                    continue;
                }
            }
            const int di = src != 3 ? 1 : 0, dj = src != 2 ? 1 : 0;
            bool ok;
            const uint32_t cm = w.gather(di, dj, lane, ok);
            const uint64_t run = __ballot(ok && int(cm & 3u) == src);
            const int n = ~run == 0 ? WAVE : __builtin_ctzll(~run);  // (at least 1: the current cell)
            cells.record(n, w.i, w.j, di, dj, lane);
            w.jump(w.i - di * n, w.j - dj * n);
        } else {
            const bool vert = mode == 2;
            bool ok;
            const uint32_t cm = w.gather(vert ? 1 : 0, vert ? 0 : 1, lane, ok);
            const int pos = (vert ? w.i : w.j) - lane;
            const uint64_t ext = __ballot(ok && pos > 0 && (cm & (vert ? 4u : 8u)) != 0u);
            const int n = ~ext == 0 ? WAVE : __builtin_ctzll(~ext);
            const uint64_t valid = __ballot(ok);
            if (n < WAVE && ((valid >> n) & 1u)) {  // the run's last cell: one more move ends it
                w.jump(w.i - (vert ? n + 1 : 0), w.j - (vert ? 0 : n + 1));
                cells.record(1, w.i + (vert ? 1 : 0), w.j + (vert ? 0 : 1), vert ? 1 : 0, vert ? 0 : 1, lane);
                mode = 0;
            } else {  // the run goes on past the tile
                w.jump(w.i - (vert ? n : 0), w.j - (vert ? 0 : n));
            }
        }
    }
    const int cnt = cells.finish(lane);
    if (lane == 0) path_len[out] = cnt;
}

hipError_t launch_walk(int gap, const Prob *probs, int64_t n, const uint32_t *tb, const int32_t *seeds,
                       int32_t *path, const int64_t *path_off, int32_t *path_len, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const WalkShape ws = walk_shape(n, 0);
    const dim3 grid(unsigned((n + ws.waves - 1) / ws.waves)), block(ws.waves * WAVE);
    const size_t lds = size_t(ws.waves) * ws.tile_words * 4;
    if (gap == GAP_AFF)
        hipLaunchKernelGGL(pa_walk_kernel<GAP_AFF>, grid, block, lds, st, probs, n, tb, seeds, path, path_off,
                           path_len, ws.tile_words);
    else  // linear and DTW share the 2-bit moves
        hipLaunchKernelGGL(pa_walk_kernel<GAP_LIN>, grid, block, lds, st, probs, n, tb, seeds, path, path_off,
                           path_len, ws.tile_words);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Backward: one wavefront per problem walks its optimal path (the moves of
// pa_walk_kernel) and emits the gradients of the optimal score.  The score
// is a sum along the path: S(u, v) for every cell left by a diagonal move
// (the alignment's matched cells; path_compressor, solver.h:831-846), the
// cost of every path cell for DTW (solver.h:3236-3248), and -w(k) per gap
// run with w(k) = open + extend*(k - 1) (pyalign/gaps.py:164-176), including
// the border run that ends a global path at (u, -1) or (-1, v) (the border
// values w(u+1), solver.h:2943-2951).  Each gradient entry is touched once
// per call (path cells are distinct; a DTW path visits each row and each
// column in one contiguous run, summed in registers in path order), so the
// adds go out as no-return atomics and the walk never waits on them.  The
// series rows the DTW cost needs come through an LDS window that, like the
// traceback tile, ends at the current row and is refilled when left.
// ---------------------------------------------------------------------------
constexpr int SERIES_WIN = 64;  // rows per series window (x and y each; small: more resident walks)

// float add to global memory with no value returned (nothing waits on it)
__device__ __forceinline__ void add_noret(float *p, float v) { (void)unsafeAtomicAdd(p, v); }

// A walk's gradient adds, held one per lane and sent as one 64-lane
// no-return atomic instruction per 64 adds.  A global memory operation in
// every move would make the moves wait on it (the wait counter the walk's
// tile loads use also counts the atomics); queued, a move touches only LDS.
// Every lane of the wave calls push and flush (n is the same in all lanes).
struct AddQueue {
    float *p = nullptr;
    float v = 0.0f;
    int n = 0;
    __device__ void push(float *ptr, float val, int lane) {
        if (lane == n) {
            p = ptr;
            v = val;
        }
        if (++n == WAVE) flush(lane);
    }
    __device__ void flush(int lane) {
        if (lane < n) add_noret(p, v);
        n = 0;
    }
};

template <int SD>
struct RowWin {
    float *lds;        // SERIES_WIN * SD floats
    const float *src;  // series rows of this problem
    int lo, hi;        // rows [lo, hi) held

    __device__ void init(float *l, const float *s) {
        lds = l;
        src = s;
        lo = hi = 0;
    }
    __device__ const float *row(int i, int lane) {
        if (i < lo || i >= hi) {
            hi = i + 1;
            lo = hi > SERIES_WIN ? hi - SERIES_WIN : 0;
            __builtin_amdgcn_wave_barrier();
            for (int w = lane; w < (hi - lo) * SD; w += WAVE) lds[w] = src[int64_t(lo) * SD + w];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        return lds + (i - lo) * SD;
    }
};

template <int GAP, int FORM, int SD>
__global__ __launch_bounds__(256) void pa_backward_kernel(BwArgs a, int tile_words) {
    constexpr int SDD = SD > 0 ? SD : 1;
    constexpr int WIN = FORM == F_SERIES ? 2 * SERIES_WIN * SDD : 0;  // floats per wavefront
    extern __shared__ uint32_t smem[];
    // the wave index read as a scalar: the problem, its descriptor and the
    // whole walk state are then wave-uniform (scalar registers and scalar
    // branches; the vector lanes only load tiles)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE), lane = threadIdx.x % WAVE;
    const int64_t k = int64_t(blockIdx.x) * (blockDim.x / WAVE) + wave;
    if (k >= a.n) return;
    const Prob pr = a.probs[k];
    if (pr.tb_off == NO_TB_OFF) return;  // checkpointed: walked strip by strip
    const int out = pr.out;
    const float g = a.grad_scores[out];
    const int u0 = a.seeds[2 * out], v0 = a.seeds[2 * out + 1];
    float opens = 0.0f, extends = 0.0f;
    uint32_t *mine = smem + size_t(wave) * (tile_words + WIN);
    TbWalker w;
    w.init(mine, tile_words, a.tb, pr, GAP == GAP_AFF ? 4 : 2);
    AddQueue q;
    [[maybe_unused]] RowWin<SDD> wx, wy;
    // DTW series: the running sums of the current row / column
    [[maybe_unused]] float gx[SDD], gy[SDD];
    [[maybe_unused]] int cur_i = -1, cur_j = -1;
    if constexpr (FORM == F_SERIES) {
        float *win = reinterpret_cast<float *>(mine + tile_words);
        wx.init(win, a.xs + pr.off_s * SD);
        wy.init(win + SERIES_WIN * SDD, a.xt + pr.off_t * SD);
    }
    auto flush_row = [&]() {
        if constexpr (FORM == F_SERIES) {
            if (cur_i >= 0 && a.grad_xs)
                for (int d = 0; d < SD; ++d) q.push(a.grad_xs + (pr.off_s + cur_i) * SD + d, gx[d], lane);
        }
    };
    auto flush_col = [&]() {
        if constexpr (FORM == F_SERIES) {
            if (cur_j >= 0 && a.grad_xt)
                for (int d = 0; d < SD; ++d) q.push(a.grad_xt + (pr.off_t + cur_j) * SD + d, gy[d], lane);
        }
    };
    auto cost_grad = [&](int i, int j, int64_t ga) {  // ga: the cell's dense offset
        if constexpr (FORM == F_DENSE) {
            if (a.grad_dense) q.push(a.grad_dense + pr.off_s + ga, g, lane);
        } else if constexpr (FORM == F_SERIES) {
            if (i != cur_i) {
                flush_row();
                cur_i = i;
                for (int d = 0; d < SD; ++d) gx[d] = 0.0f;
            }
            if (j != cur_j) {
                flush_col();
                cur_j = j;
                for (int d = 0; d < SD; ++d) gy[d] = 0.0f;
            }
            const float *xr = wx.row(i, lane);
            const float *yr = wy.row(j, lane);
#pragma unroll
            for (int d = 0; d < SD; ++d) {
                const float gd = g * (2.0f * (xr[d] - yr[d]));
                gx[d] += gd;
                gy[d] -= gd;
            }
        }
    };
    int u = u0, v = v0;
    if (u0 != NO_SEED) {
        w.seek(u0, v0);
        if constexpr (FORM == F_DENSE) {
            // A run at a time: lane m reads the cell m moves on along the run
            // the path is in (cells of one code: diagonal moves, and linear gap
            // or DTW moves of one direction; or an affine gap run, whose cells
            // extend it while their extend bit is set), a ballot gives the run's
            // length, and the run's gradient cells (each touched once) go out as
            // one atomic instruction.  The moves are walk_step's.
            int mode = 0, pu = 0;  // affine: 2 / 3 inside a vertical / horizontal gap run from row / column pu
            while (w.i >= 0 && w.j >= 0) {
                const uint32_t c0 = w.code(lane);  // (loads a tile holding the current cell)
                if (mode == 0) {
                    const int src = int(c0 & 3u);
                    if (src == 0) {  // local: the path starts at a zero cell
                        if constexpr (GAP == GAP_DTW) cost_grad(w.i, w.j, w.ga);
                        break;
                    }
                    if constexpr (GAP == GAP_AFF) {
                        if (src >= 2) {  // a gap run starts at this cell
                            mode = src;
                            pu = src == 2 ? w.i : w.j;
                            continue;
                        }
                    }
                    const int di = src != 3 ? 1 : 0, dj = src != 2 ? 1 : 0;
                    bool ok;
                    const uint32_t cm = w.gather(di, dj, lane, ok);
                    const uint64_t run = __ballot(ok && int(cm & 3u) == src);
                    const int n = ~run == 0 ? WAVE : __builtin_ctzll(~run);  // (at least 1: the current cell)
                    if (GAP == GAP_DTW || src == 1) {
                        if (a.grad_dense && lane < n)
                            add_noret(a.grad_dense + pr.off_s + w.ga - int64_t(lane) * (di * w.gsi + dj * w.gsj), g);
                    } else {
                        opens += float(n);  // linear: one gap symbol per move
                    }
                    w.jump(w.i - di * n, w.j - dj * n);
                } else {
                    const bool vert = mode == 2;
                    bool ok;
                    const uint32_t cm = w.gather(vert ? 1 : 0, vert ? 0 : 1, lane, ok);
                    const int pos = (vert ? w.i : w.j) - lane;
                    const uint64_t ext = __ballot(ok && pos > 0 && (cm & (vert ? 4u : 8u)) != 0u);
                    const int n = ~ext == 0 ? WAVE : __builtin_ctzll(~ext);
                    const uint64_t valid = __ballot(ok);
                    if (n < WAVE && ((valid >> n) & 1u)) {  // the run's last cell: one more move ends it
                        w.jump(w.i - (vert ? n + 1 : 0), w.j - (vert ? 0 : n + 1));
                        opens += 1.0f;
                        extends += float(pu - (vert ? w.i : w.j) - 1);  // run length - 1
                        mode = 0;
                    } else {  // the run goes on past the tile
                        w.jump(w.i - (vert ? n : 0), w.j - (vert ? 0 : n));
                    }
                }
            }
        } else if constexpr (FORM == F_SERIES && GAP == GAP_DTW) {
            // DTW over series, a run at a time (moves of one direction):
            // lane t computes the cost gradient of the run's cell t.  A row
            // (column) the path crosses is one contiguous stretch of its
            // cells, summed in path order: across a run (diagonal, or up for
            // rows) each cell's row is complete within the run and its sum
            // (the pending one, for the run's first cell) goes out at once;
            // along a run (left for rows, up for columns) the cells add to
            // the running sum in order, lane by lane.  The sums are those of
            // cost_grad, bit for bit.
            while (w.i >= 0 && w.j >= 0) {
                const uint32_t c0 = w.code(lane);  // (loads a tile holding the current cell)
                wx.row(w.i, lane);                 // (series windows holding the current cell)
                wy.row(w.j, lane);
                const int src = int(c0 & 3u);
                if (src == 0) {
                    cost_grad(w.i, w.j, w.ga);
                    break;
                }
                const int di = src != 3 ? 1 : 0, dj = src != 2 ? 1 : 0;
                bool ok;
                const uint32_t cm = w.gather(di, dj, lane, ok);
                const int ii = w.i - di * lane, jj = w.j - dj * lane;
                ok = ok && ii >= wx.lo && jj >= wy.lo;
                const uint64_t run = __ballot(ok && int(cm & 3u) == src);
                const int n = ~run == 0 ? WAVE : __builtin_ctzll(~run);  // (at least 1: the current cell)
                float gd[SDD];
                const bool mine = lane < n;
#pragma unroll
                for (int d = 0; d < SD; ++d) {
                    const float xv = mine ? wx.lds[(ii - wx.lo) * SD + d] : 0.0f;
                    const float yv = mine ? wy.lds[(jj - wy.lo) * SD + d] : 0.0f;
                    gd[d] = g * (2.0f * (xv - yv));
                }
                // rows: complete per cell across the run, a running sum along it
                if (di == 1) {
                    if (mine && a.grad_xs) {
                        const bool cont = lane == 0 && cur_i == w.i;
#pragma unroll
                        for (int d = 0; d < SD; ++d)
                            add_noret(a.grad_xs + (pr.off_s + ii) * SD + d, (cont ? gx[d] : 0.0f) + gd[d]);
                    }
                    cur_i = -1;  // (nothing pending)
                } else {
                    if (cur_i != w.i) {
                        flush_row();
                        cur_i = w.i;
#pragma unroll
                        for (int d = 0; d < SD; ++d) gx[d] = 0.0f;
                    }
                    for (int t = 0; t < n; ++t)
#pragma unroll
                        for (int d = 0; d < SD; ++d)
                            gx[d] += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(gd[d]), t));
                }
                // columns likewise
                if (dj == 1) {
                    if (mine && a.grad_xt) {
                        const bool cont = lane == 0 && cur_j == w.j;
#pragma unroll
                        for (int d = 0; d < SD; ++d)
                            add_noret(a.grad_xt + (pr.off_t + jj) * SD + d, (cont ? gy[d] : 0.0f) - gd[d]);
                    }
                    cur_j = -1;
                } else {
                    if (cur_j != w.j) {
                        flush_col();
                        cur_j = w.j;
#pragma unroll
                        for (int d = 0; d < SD; ++d) gy[d] = 0.0f;
                    }
                    for (int t = 0; t < n; ++t)
#pragma unroll
                        for (int d = 0; d < SD; ++d)
                            gy[d] -= __int_as_float(__builtin_amdgcn_readlane(__float_as_int(gd[d]), t));
                }
                w.jump(w.i - di * n, w.j - dj * n);
            }
        } else {
            while (w.i >= 0 && w.j >= 0) {
                const int pu = w.i, pv = w.j;
                const int64_t pga = w.ga;
                const int src = walk_step<GAP>(w, lane);
                if constexpr (GAP == GAP_DTW) cost_grad(pu, pv, pga);
                if (src == 0) break;  // local: the path starts at a zero cell
                if (src == 1) {
                    if constexpr (GAP != GAP_DTW) cost_grad(pu, pv, pga);
                } else if constexpr (GAP == GAP_AFF) {
                    opens += 1.0f;
                    extends += float((pu - w.i) + (pv - w.j) - 1);  // run length - 1
                } else if constexpr (GAP == GAP_LIN) {
                    opens += 1.0f;  // linear: one gap symbol per move
                }
            }
        }
        u = w.i;
        v = w.j;
        flush_row();
        flush_col();
        if (GAP != GAP_DTW && a.global) {
            const int run = (u >= 0 && v < 0) ? u + 1 : ((v >= 0 && u < 0) ? v + 1 : 0);
            if (run > 0) {
                if constexpr (GAP == GAP_AFF) {
                    opens += 1.0f;
                    extends += float(run - 1);
                } else {
                    opens += float(run);
                }
            }
        }
    }
    if (a.grad_gap) {
        q.push(a.grad_gap + 2 * out, a.gsign * g * opens, lane);
        q.push(a.grad_gap + 2 * out + 1, a.gsign * g * extends, lane);
    }
    q.flush(lane);
}

hipError_t launch_backward(int gap, int form, int dim, const BwArgs &args, hipStream_t st) {
    if (args.n == 0) return hipSuccess;
    const int win = form == F_SERIES ? 2 * SERIES_WIN * dim : 0;
    const WalkShape ws = walk_shape(args.n, win);
    const dim3 grid(unsigned((args.n + ws.waves - 1) / ws.waves)), block(ws.waves * WAVE);
    const size_t lds = size_t(ws.waves) * (ws.tile_words + win) * 4;
#define PA_BW(G, F, D) \
    hipLaunchKernelGGL((pa_backward_kernel<G, F, D>), grid, block, lds, st, args, ws.tile_words)
    if (form == F_DENSE) {
        if (gap == GAP_AFF) PA_BW(GAP_AFF, F_DENSE, 0);
        else if (gap == GAP_LIN) PA_BW(GAP_LIN, F_DENSE, 0);
        else PA_BW(GAP_DTW, F_DENSE, 0);
    } else if (form == F_SERIES && gap == GAP_DTW) {
        switch (dim) {
            case 1: PA_BW(GAP_DTW, F_SERIES, 1); break;
            case 2: PA_BW(GAP_DTW, F_SERIES, 2); break;
            case 3: PA_BW(GAP_DTW, F_SERIES, 3); break;
            case 4: PA_BW(GAP_DTW, F_SERIES, 4); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        return hipErrorInvalidValue;
    }
#undef PA_BW
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Band mode: fold each problem's strip partials in strip order with the
// single-segment rules (local: larger i then larger j on ties; global: the
// last row's strip; semiglobal: last row first, then the last column in
// ascending i with strict improvement, solver.h:2042-2057, 2179-2187,
// 2273-2293).
// ---------------------------------------------------------------------------
template <int LOC, int DIR>
__global__ __launch_bounds__(256) void pa_band_combine_kernel(const Prob *probs, const int32_t *band_probs, int n,
                                                              const BandPartial *parts, float *scores,
                                                              int32_t *seeds, int score_only) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const Prob pr = probs[band_probs[t]];
    const int ls = pr.len_s, lt = pr.len_t;
    const int rps = pr.L * pr.R;
    const int nst = (ls + rps - 1) / rps;
    const BandPartial *p = parts + pr.band;
    constexpr float W = worst<DIR>();
    float score = W;
    int si = NO_SEED, sj = 0;
    if constexpr (LOC == LOC_LOCAL) {
        float best = W;
        int bi = -1, bj = -1;
        for (int st = 0; st < nst; ++st) {
            if (better<DIR>(p[st].v1, best) || (p[st].v1 == best && p[st].i1 > bi)) {
                best = p[st].v1;
                bi = p[st].i1;
                bj = p[st].j1;
            }
        }
        score = best;
        si = score_only ? 0 : bi;
        sj = score_only ? 0 : bj;
    } else if constexpr (LOC == LOC_GLOBAL) {
        score = p[nst - 1].v1;
        si = ls - 1;
        sj = lt - 1;
    } else {
        float cb = W;
        int ci = -1;
        for (int st = 0; st < nst; ++st) {
            if (better<DIR>(p[st].v1, cb) || (p[st].v1 == cb && p[st].i1 >= 0 && (ci < 0 || p[st].i1 < ci))) {
                cb = p[st].v1;
                ci = p[st].i1;
            }
        }
        const float rb = p[nst - 1].v2;
        if (better<DIR>(cb, rb)) {
            score = cb;
            si = ci;
            sj = lt - 1;
        } else {
            score = rb;
            si = ls - 1;
            sj = p[nst - 1].j2;
        }
    }
    scores[pr.out] = score;
    if (seeds) {
        seeds[2 * pr.out] = si;
        seeds[2 * pr.out + 1] = sj;
    }
}

hipError_t launch_band_combine(int loc, int dir, int tr, const Prob *probs, const int32_t *band_probs, int n,
                               const BandPartial *partials, float *scores, int32_t *seeds, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 grid(unsigned((n + 255) / 256)), block(256);
    const int so = tr == TR_SCORE ? 1 : 0;
#define PA_BC(LC, DR) \
    hipLaunchKernelGGL((pa_band_combine_kernel<LC, DR>), grid, block, 0, st, probs, band_probs, n, partials, scores, \
                       seeds, so)
    const bool mx = dir == DIR_MAX;
    switch (loc) {
        case LOC_LOCAL: if (mx) PA_BC(LOC_LOCAL, DIR_MAX); else PA_BC(LOC_LOCAL, DIR_MIN); break;
        case LOC_GLOBAL: if (mx) PA_BC(LOC_GLOBAL, DIR_MAX); else PA_BC(LOC_GLOBAL, DIR_MIN); break;
        default: if (mx) PA_BC(LOC_SEMI, DIR_MAX); else PA_BC(LOC_SEMI, DIR_MIN); break;
    }
#undef PA_BC
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Checkpointed backward: the walk of pa_backward_kernel continued one strip
// at a time (from the seed's strip up), on a traceback recomputed for that strip
// only.  The walk stops when it leaves the strip upwards and saves where it
// is, including a vertical gap run still searching for its origin (the
// only move that crosses strips), the gap counts so far and the DTW
// running sums; the next pass resumes in the strip above.
// ---------------------------------------------------------------------------
template <int GAP, int FORM, int SD>
__global__ __launch_bounds__(256) void pa_backward_strip_kernel(BwArgs a, StripWalk *state, const int32_t *strip_of,
                                                                int tile_words) {
    constexpr int SDD = SD > 0 ? SD : 1;
    constexpr int WIN = FORM == F_SERIES ? 2 * SERIES_WIN * SDD : 0;
    extern __shared__ uint32_t smem[];
    // the wave index read as a scalar: the problem, its descriptor and the
    // whole walk state are then wave-uniform (scalar registers and scalar
    // branches; the vector lanes only load tiles)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / WAVE), lane = threadIdx.x % WAVE;
    const int64_t k = int64_t(blockIdx.x) * (blockDim.x / WAVE) + wave;
    if (k >= a.n) return;
    const int st = strip_of[k];
    if (st < 0) return;
    StripWalk ws = state[k];
    if (ws.done) return;
    const Prob pr = a.probs[k];
    const int out = pr.out;
    const float g = a.grad_scores[out];
    const int lo = st * pr.L * pr.R;
    uint32_t *mine = smem + size_t(wave) * (tile_words + WIN);
    TbWalker w;
    w.init(mine, tile_words, a.tb, pr, GAP == GAP_AFF ? 4 : 2);
    AddQueue q;
    [[maybe_unused]] RowWin<SDD> wx, wy;
    if constexpr (FORM == F_SERIES) {
        float *win = reinterpret_cast<float *>(mine + tile_words);
        wx.init(win, a.xs + pr.off_s * SD);
        wy.init(win + SERIES_WIN * SDD, a.xt + pr.off_t * SD);
    }
    float opens = ws.opens, extends = ws.extends;
    [[maybe_unused]] float gx[SDD], gy[SDD];
#pragma unroll
    for (int d = 0; d < SDD; ++d) {
        gx[d] = ws.gx[d];
        gy[d] = ws.gy[d];
    }
    int cur_i = ws.cur_i, cur_j = ws.cur_j;
    auto flush_row = [&]() {
        if constexpr (FORM == F_SERIES) {
            if (cur_i >= 0 && a.grad_xs)
                for (int d = 0; d < SD; ++d) q.push(a.grad_xs + (pr.off_s + cur_i) * SD + d, gx[d], lane);
        }
    };
    auto flush_col = [&]() {
        if constexpr (FORM == F_SERIES) {
            if (cur_j >= 0 && a.grad_xt)
                for (int d = 0; d < SD; ++d) q.push(a.grad_xt + (pr.off_t + cur_j) * SD + d, gy[d], lane);
        }
    };
    auto cost_grad = [&](int i, int j, int64_t ga) {  // ga: the cell's dense offset
        if constexpr (FORM == F_DENSE) {
            if (a.grad_dense) q.push(a.grad_dense + pr.off_s + ga, g, lane);
        } else if constexpr (FORM == F_SERIES) {
            if (i != cur_i) {
                flush_row();
                cur_i = i;
                for (int d = 0; d < SD; ++d) gx[d] = 0.0f;
            }
            if (j != cur_j) {
                flush_col();
                cur_j = j;
                for (int d = 0; d < SD; ++d) gy[d] = 0.0f;
            }
            const float *xr = wx.row(i, lane);
            const float *yr = wy.row(j, lane);
#pragma unroll
            for (int d = 0; d < SD; ++d) {
                const float gd = g * (2.0f * (xr[d] - yr[d]));
                gx[d] += gd;
                gy[d] -= gd;
            }
        }
    };
    w.seek(ws.i, ws.j);
    int mode = ws.mode, pu = ws.pu;
    bool finished = false;
    while (true) {
        if (w.i < 0 || w.j < 0) {
            finished = true;
            break;
        }
        if (w.i < lo) break;  // continues in the strip above
        if (mode == 1) {      // vertical gap run: up column j to its origin
            if (w.i > 0 && (w.code(lane) & 4u)) {
                w.up();
                continue;
            }
            w.up();
            opens += 1.0f;
            extends += float(pu - w.i - 1);
            mode = 0;
            continue;
        }
        const int ci = w.i, cj = w.j;
        const int64_t cga = w.ga;
        const int src = int(w.code(lane) & 3u);
        if constexpr (GAP == GAP_DTW) cost_grad(ci, cj, cga);
        if (src == 0) {  // a local start
            finished = true;
            break;
        }
        if (src == 1) {
            if constexpr (GAP != GAP_DTW) cost_grad(ci, cj, cga);
            w.up();
            w.left();
        } else if (src == 2) {
            if constexpr (GAP == GAP_AFF) {
                mode = 1;  // the scan starts at this cell
                pu = ci;
            } else {
                opens += 1.0f;
                w.up();
            }
        } else {
            if constexpr (GAP == GAP_AFF) {
                while (w.j > 0 && (w.code(lane) & 8u)) w.left();
                w.left();
                opens += 1.0f;
                extends += float(cj - w.j - 1);
            } else {
                opens += 1.0f;
                w.left();
            }
        }
    }
    if (finished) {
        flush_row();
        flush_col();
        const int u = w.i, v = w.j;
        if (GAP != GAP_DTW && a.global) {
            const int run = (u >= 0 && v < 0) ? u + 1 : ((v >= 0 && u < 0) ? v + 1 : 0);
            if (run > 0) {
                if constexpr (GAP == GAP_AFF) {
                    opens += 1.0f;
                    extends += float(run - 1);
                } else {
                    opens += float(run);
                }
            }
        }
        if (a.grad_gap) {
            q.push(a.grad_gap + 2 * out, a.gsign * g * opens, lane);
            q.push(a.grad_gap + 2 * out + 1, a.gsign * g * extends, lane);
        }
    }
    q.flush(lane);
    if (lane == 0) {
        ws.i = w.i;
        ws.j = w.j;
        ws.mode = mode;
        ws.pu = pu;
        ws.cur_i = cur_i;
        ws.cur_j = cur_j;
        ws.done = finished ? 1 : 0;
        ws.opens = opens;
        ws.extends = extends;
#pragma unroll
        for (int d = 0; d < SDD; ++d) {
            ws.gx[d] = gx[d];
            ws.gy[d] = gy[d];
        }
        state[k] = ws;
    }
}

hipError_t launch_backward_strips(int gap, int form, int dim, const BwArgs &args, StripWalk *state,
                                  const int32_t *strip_of, hipStream_t st) {
    if (args.n == 0) return hipSuccess;
    const int win = form == F_SERIES ? 2 * SERIES_WIN * dim : 0;
    const WalkShape ws = walk_shape(args.n, win);
    const dim3 grid(unsigned((args.n + ws.waves - 1) / ws.waves)), block(ws.waves * WAVE);
    const size_t lds = size_t(ws.waves) * (ws.tile_words + win) * 4;
#define PA_BS(G, F, D)                                                                                         \
    hipLaunchKernelGGL((pa_backward_strip_kernel<G, F, D>), grid, block, lds, st, args, state, strip_of, \
                       ws.tile_words)
    if (form == F_DENSE) {
        if (gap == GAP_AFF) PA_BS(GAP_AFF, F_DENSE, 0);
        else if (gap == GAP_LIN) PA_BS(GAP_LIN, F_DENSE, 0);
        else PA_BS(GAP_DTW, F_DENSE, 0);
    } else if (form == F_SERIES && gap == GAP_DTW) {
        switch (dim) {
            case 1: PA_BS(GAP_DTW, F_SERIES, 1); break;
            case 2: PA_BS(GAP_DTW, F_SERIES, 2); break;
            case 3: PA_BS(GAP_DTW, F_SERIES, 3); break;
            case 4: PA_BS(GAP_DTW, F_SERIES, 4); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        return hipErrorInvalidValue;
    }
#undef PA_BS
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// General gap costs (GeneralGapCostSolver, solver.h:3047-3184: the
// Waterman-Smith-Beyer recurrence, O(n m (n + m))).  One workgroup per
// problem sweeps its anti-diagonals, one cell per thread; every cell folds
// the reference's candidates in push order with strict improvement:
// [local zero], the diagonal, the column above from row -1 down
// (D(k, j) + w_s(i - k)), the row to the left from column -1 on
// (D(i, k) + w_t(j - k)).  The threads of a wavefront hold consecutive rows
// i (and columns j = d - i), and scan in lockstep over k: the column scan
// reads row k of D and the row scan row k of the transposed copy, so both
// loads are adjacent across the wavefront.  After the sweep the workgroup
// reduces the seed (solver.h:2042-2057, 2179-2187, 2273-2293) and, for one
// optimal path, thread 0 follows the traceback codes (build_path,
// solver.h:858-935).  ws[k] / wt[k] hold w(k) * sgn.
// ---------------------------------------------------------------------------
constexpr int WSB_MAX_THREADS = 1024;
constexpr int WSB_UNROLL = 8;  // candidate loads issued before their compares

// Fold the candidates base[k * stride] + w[top - k], k = 0 .. count-1, into
// (acc, code) in order, code = code0 + k on a strict improvement.  The loads
// of WSB_UNROLL candidates go out together; the compares stay in order.
template <int DIR>
__device__ __forceinline__ void wsb_scan(const float *base, int64_t stride, const float *w, int top, int count,
                                         int code0, float &acc, int &code) {
    int k = 0;
    for (; k + WSB_UNROLL <= count; k += WSB_UNROLL) {
        float c[WSB_UNROLL];
#pragma unroll
        for (int q = 0; q < WSB_UNROLL; ++q) c[q] = base[int64_t(k + q) * stride] + w[top - k - q];
#pragma unroll
        for (int q = 0; q < WSB_UNROLL; ++q) {
            if (better<DIR>(c[q], acc)) {
                acc = c[q];
                code = code0 + k + q;
            }
        }
    }
    for (; k < count; ++k) {
        const float c = base[int64_t(k) * stride] + w[top - k];
        if (better<DIR>(c, acc)) {
            acc = c;
            code = code0 + k;
        }
    }
}

template <int LOC, int DIR, int FORM>
__global__ __launch_bounds__(WSB_MAX_THREADS) void pa_wsb_kernel(KArgs a, const float *ws, const float *wt,
                                                                 int32_t *path, const int64_t *path_off,
                                                                 int32_t *path_len) {
    __shared__ float red_v[WSB_MAX_THREADS];
    __shared__ int64_t red_p[WSB_MAX_THREADS];
    const Prob pr = a.probs[blockIdx.x];
    const int ls = pr.len_s, lt = pr.len_t;
    const int tid = int(threadIdx.x);
    if (ls <= 0 || lt <= 0) {
        if (tid == 0) {
            write_empty<LOC, DIR>(a, pr);
            if (path) {
                const int si = a.seeds[2 * pr.out], sj = a.seeds[2 * pr.out + 1];
                if (si == NO_SEED) {
                    path_len[pr.out] = -1;
                } else {
                    int32_t *P = path + 2 * path_off[pr.out];
                    P[0] = si;
                    P[1] = sj;
                    path_len[pr.out] = 1;
                }
            }
        }
        return;
    }
    const int64_t cells = int64_t(ls) * lt;
    float *D = a.dvals + pr.dv_off;
    float *DT = D + cells;
    float *SV = D + 2 * cells;
    int32_t *TB = reinterpret_cast<int32_t *>(D + 3 * cells);
    auto Dv = [&](int i, int j) -> float {  // D with its borders
        if (i < 0) return a.border_t[j + 1];
        if (j < 0) return a.border_s[i + 1];
        return D[int64_t(i) * lt + j];
    };
    for (int d = 0; d <= ls + lt - 2; ++d) {
        const int i0 = d - lt + 1 > 0 ? d - lt + 1 : 0;
        const int i1 = d < ls - 1 ? d : ls - 1;
        for (int i = i0 + tid; i <= i1; i += int(blockDim.x)) {
            const int j = d - i;
            float S;
            if constexpr (FORM == F_DENSE) {
                S = a.dense[pr.off_s + int64_t(i) * pr.si + int64_t(j) * pr.sj];
            } else if constexpr (FORM == F_TABLE) {
                const int x = ((const uint8_t *)a.codes_s)[pr.off_s + i];
                const int y = ((const uint8_t *)a.codes_t)[pr.off_t + j];
                S = a.table[x * a.K1 + y];
            } else if constexpr (FORM == F_TABLE_G) {
                // codes beyond the table read its pad row / column
                const int x = min(int(((const uint16_t *)a.codes_s)[pr.off_s + i]), a.K1 - 1);
                const int y = min(int(((const uint16_t *)a.codes_t)[pr.off_t + j]), a.K1 - 1);
                S = a.table[int64_t(x) * a.K1 + y];
            } else {
                const uint32_t x = ((const uint32_t *)a.codes_s)[pr.off_s + i];
                const uint32_t y = ((const uint32_t *)a.codes_t)[pr.off_t + j];
                S = x == y ? a.eq : a.ne;
            }
            const float dg = Dv(i - 1, j - 1) + S;
            float acc;
            int code;
            if constexpr (LOC == LOC_LOCAL) {
                acc = 0.0f;
                code = 0;
                if (better<DIR>(dg, acc)) {
                    acc = dg;
                    code = 1;
                }
            } else {
                acc = dg;
                code = 1;
            }
            {  // the column above, k = -1 .. i-1
                const float c = a.border_t[j + 1] + ws[i + 1];
                if (better<DIR>(c, acc)) {
                    acc = c;
                    code = 2;
                }
            }
            wsb_scan<DIR>(D + j, lt, ws, i, i, 3, acc, code);
            {  // the row to the left, k = -1 .. j-1
                const float c = a.border_s[i + 1] + wt[j + 1];
                if (better<DIR>(c, acc)) {
                    acc = c;
                    code = i + 3;
                }
            }
            wsb_scan<DIR>(DT + i, ls, wt, j, j, i + 4, acc, code);
            D[int64_t(i) * lt + j] = acc;
            DT[int64_t(j) * ls + i] = acc;
            SV[int64_t(i) * lt + j] = S;
            TB[int64_t(i) * lt + j] = code;
        }
        __syncthreads();
    }

    // ---- seed: the first optimum in the reference's visiting order ---------
    // local: rows and columns descending (the largest flat index first);
    // semiglobal: the last row left to right, then the last column top down
    // (position p); global: the corner.  Cells that do not beat the worst
    // value (and NaNs) never seed.
    constexpr float W = worst<DIR>();
    int si = ls - 1, sj = lt - 1;
    float score = D[cells - 1];
    if constexpr (LOC != LOC_GLOBAL) {
        float bv = W;
        int64_t bp = -1;
        auto fold = [&](float v, int64_t p) {
            if (!better<DIR>(v, W)) return;
            const bool earlier = LOC == LOC_LOCAL ? p > bp : p < bp;
            if (bp < 0 || better<DIR>(v, bv) || (v == bv && earlier)) {
                bv = v;
                bp = p;
            }
        };
        if constexpr (LOC == LOC_LOCAL) {
            for (int64_t q = tid; q < cells; q += blockDim.x) fold(D[q], q);
        } else {
            for (int64_t p = tid; p < int64_t(lt) + ls; p += blockDim.x)
                fold(p < lt ? D[int64_t(ls - 1) * lt + p] : D[(p - lt) * lt + (lt - 1)], p);
        }
        red_v[tid] = bv;
        red_p[tid] = bp;
        __syncthreads();
        if (tid == 0) {
            bv = W;
            bp = -1;
            for (int t = 0; t < int(blockDim.x); ++t)
                if (red_p[t] >= 0) fold(red_v[t], red_p[t]);
            if (bp < 0) {
                si = NO_SEED;
                sj = 0;
                score = W;
            } else if (LOC == LOC_LOCAL) {
                si = int(bp / lt);
                sj = int(bp % lt);
                score = bv;
            } else {
                si = bp < lt ? ls - 1 : int(bp - lt);
                sj = bp < lt ? int(bp) : lt - 1;
                score = bv;
            }
        }
    }
    if (tid != 0) return;
    a.scores[pr.out] = score;
    a.seeds[2 * pr.out] = si;
    a.seeds[2 * pr.out + 1] = sj;
    if (!path) return;
    if (si == NO_SEED) {
        path_len[pr.out] = -1;
        return;
    }
    // ---- one optimal path: follow the traceback codes ----------------------
    int32_t *P = path + 2 * path_off[pr.out];
    int u = si, v = sj, cnt = 1;
    P[0] = u;
    P[1] = v;
    while (u >= 0 && v >= 0) {
        const int64_t c = int64_t(u) * lt + v;
        // Local stops where D is not above zero (solver.h:2117-2127)
        if (LOC == LOC_LOCAL && !better<DIR>(D[c], 0.0f)) break;
        const int code = TB[c];
        if (code == 0) break;  // the local zero holds D = 0, stopped above
        if (code == 1) {
            u -= 1;
            v -= 1;
        } else if (code <= u + 2) {
            u = code - 3;
        } else {
            v = code - u - 4;
        }
        P[2 * cnt] = u;
        P[2 * cnt + 1] = v;
        ++cnt;
    }
    path_len[pr.out] = cnt;
}

hipError_t launch_wsb(int loc, int dir, int form, const KArgs &a, int64_t n, const float *ws, const float *wt,
                      int32_t *path, const int64_t *path_off, int32_t *path_len, int threads, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (threads < WAVE || threads > WSB_MAX_THREADS || threads % WAVE) return hipErrorInvalidValue;
    const dim3 grid(static_cast<unsigned>(n)), block(static_cast<unsigned>(threads));
#define PA_WSB(LC, DR, FM) \
    hipLaunchKernelGGL((pa_wsb_kernel<LC, DR, FM>), grid, block, 0, st, a, ws, wt, path, path_off, path_len)
#define PA_WSB_F(LC, DR)                                        \
    do {                                                        \
        if (form == F_DENSE) PA_WSB(LC, DR, F_DENSE);           \
        else if (form == F_TABLE) PA_WSB(LC, DR, F_TABLE);      \
        else if (form == F_TABLE_G) PA_WSB(LC, DR, F_TABLE_G);  \
        else if (form == F_BINARY) PA_WSB(LC, DR, F_BINARY);    \
        else return hipErrorInvalidValue;                       \
    } while (0)
    const bool mx = dir == DIR_MAX;
    switch (loc) {
        case LOC_LOCAL: if (mx) PA_WSB_F(LOC_LOCAL, DIR_MAX); else PA_WSB_F(LOC_LOCAL, DIR_MIN); break;
        case LOC_GLOBAL: if (mx) PA_WSB_F(LOC_GLOBAL, DIR_MAX); else PA_WSB_F(LOC_GLOBAL, DIR_MIN); break;
        default: if (mx) PA_WSB_F(LOC_SEMI, DIR_MAX); else PA_WSB_F(LOC_SEMI, DIR_MIN); break;
    }
#undef PA_WSB_F
#undef PA_WSB
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Problem encoding on the device: raw symbols of a whole bag (1-byte latin-1
// or 4-byte code points, the text the host joined) become the solver's codes,
// through `lut` into table codes for the indexed form (uint8, or uint16 for
// tables read from global memory), or widened to uint32 values for the
// binary form (lut == nullptr).  A symbol without a code (beyond lut_len, or
// 255 in lut) writes code 0 and leaves the smallest such position in *bad,
// which the host checks before any DP launch.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void pa_translate_kernel(const void *raw, int width, int64_t count,
                                                           const uint8_t *lut, int32_t lut_len, void *codes,
                                                           int out_width, unsigned long long *bad) {
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < count; q += stride) {
        const uint32_t x = width == 1 ? uint32_t(static_cast<const uint8_t *>(raw)[q])
                                      : static_cast<const uint32_t *>(raw)[q];
        if (lut) {
            uint32_t c = x < uint32_t(lut_len) ? uint32_t(lut[x]) : 255u;
            if (c == 255u) {
                atomicMin(bad, static_cast<unsigned long long>(q));
                c = 0;
            }
            if (out_width == 2) static_cast<uint16_t *>(codes)[q] = uint16_t(c);
            else static_cast<uint8_t *>(codes)[q] = uint8_t(c);
        } else {
            static_cast<uint32_t *>(codes)[q] = x;
        }
    }
}

hipError_t launch_translate(const void *raw, int width, int64_t count, const uint8_t *lut, int32_t lut_len,
                            void *codes, int out_width, unsigned long long *bad, int cus, hipStream_t st) {
    if (count == 0) return hipSuccess;
    if (width != 1 && width != 4) return hipErrorInvalidValue;
    if (lut ? (out_width != 1 && out_width != 2) : out_width != 4) return hipErrorInvalidValue;
    const int64_t blocks = std::min<int64_t>((count + 255) / 256, int64_t(std::max(cus, 1)) * 16);
    hipLaunchKernelGGL(pa_translate_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, st, raw, width,
                       count, lut, lut_len, codes, out_width, bad);
    return hipGetLastError();
}

__global__ void pa_fill_u32_kernel(uint32_t *p, int64_t n, uint32_t value) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = value;
}

hipError_t launch_fill_u32(uint32_t *p, int64_t n, uint32_t value, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pa_fill_u32_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, p, n, value);
    return hipGetLastError();
}

hipError_t launch_zero_u32(uint32_t *p, int64_t n, hipStream_t st) { return launch_fill_u32(p, n, 0u, st); }

// flag (set to 1 beforehand) cleared when a value is not finite
__global__ __launch_bounds__(256) void pa_all_finite_kernel(const float *x, int64_t nx, const float *y, int64_t ny,
                                                            uint32_t *flag) {
    bool bad = false;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < nx + ny; k += stride)
        bad = bad || !__builtin_isfinite(k < nx ? x[k] : y[k - nx]);
    if (__any(bad) && (threadIdx.x & (WAVE - 1)) == 0)
        __hip_atomic_store(flag, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

hipError_t launch_all_finite(const float *x, int64_t nx, const float *y, int64_t ny, uint32_t *flag,
                             hipStream_t st) {
    hipError_t e = launch_fill_u32(flag, 1, 1u, st);
    if (e != hipSuccess || nx + ny == 0) return e;
    const int64_t blocks = std::min<int64_t>((nx + ny + 255) / 256, int64_t(device_cus()) * 8);
    hipLaunchKernelGGL(pa_all_finite_kernel, dim3(unsigned(blocks)), dim3(256), 0, st, x, nx, y, ny, flag);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Device memory of the allocation-free entry point: every GPU the library
// runs on holds one copy of this array for the life of the process.  Host
// code never passes the array's name to a kernel or a copy (on the host it is
// only a stand-in for the symbol): the address of the current device's copy
// comes from hipGetSymbolAddress, looked up once per device and kept.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
int g_scratch_cus[NOALLOC_MAX_DEVICES] = {};
}  // namespace

hipError_t noalloc_scratch(void **ptr, int *cus) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        int n = 0;
        e = hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
        if (e != hipSuccess) return e;
        g_scratch_cus[dev] = n;
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    *cus = g_scratch_cus[dev];
    return hipSuccess;
}

}  // namespace pa
