// pa_kernels.hip -- traceback walk, small helper kernels and the top-level
// dispatch into the per-locality instantiation units (pa_k_*.hip).
#include "pa_internal.h"

namespace pa {

hipError_t dispatch_local_max(const LaunchSpec &, const KArgs &, hipStream_t, int *);
hipError_t dispatch_local_min(const LaunchSpec &, const KArgs &, hipStream_t, int *);
hipError_t dispatch_global_max(const LaunchSpec &, const KArgs &, hipStream_t, int *);
hipError_t dispatch_global_min(const LaunchSpec &, const KArgs &, hipStream_t, int *);
hipError_t dispatch_semi_max(const LaunchSpec &, const KArgs &, hipStream_t, int *);
hipError_t dispatch_semi_min(const LaunchSpec &, const KArgs &, hipStream_t, int *);

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

int occupancy_waves(const LaunchSpec &s) {
    int waves = 0;
    KArgs dummy{};
    if (dispatch(s, dummy, nullptr, &waves) != hipSuccess || waves <= 0) return 4;
    return waves;
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
// Traceback walk: one thread per problem follows the packed direction bits
// from the seed (TracebackIterators<false>::Iterator::next, solver.h:1498-1536).
// Affine gaps are single jumps to the gap's origin found through the P/Q
// extend bits, because P and Q copy their predecessor's traceback
// (solver.h:2999-3016).  Bit layout per problem: strip st, step s, lane k,
// word w at ((st * (len_t + L - 1) + s) * L + k) * words + w, cell (i, j)
// at lane k = (i mod L*R) / R, step s = j + k.
// ---------------------------------------------------------------------------
template <int GAP>
__global__ __launch_bounds__(256) void pa_walk_kernel(const Prob *probs, int64_t n, const uint32_t *tb,
                                                      const int32_t *seeds, int32_t *path,
                                                      const int64_t *path_off, int32_t *path_len) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Prob pr = probs[k];
    const int out = pr.out;
    int u = seeds[2 * out], v = seeds[2 * out + 1];
    if (u == NO_SEED) {
        path_len[out] = -1;
        return;
    }
    constexpr int BITS = GAP == GAP_AFF ? 4 : 2;
    const int L = pr.L, R = pr.R, lt = pr.len_t;
    const int rps = L * R;
    const int wpc = (BITS * R + 31) / 32;
    const int64_t nsteps = lt + L - 1;
    auto cell = [&](int i, int j) -> uint32_t {
        const int st = i / rps;
        const int rem = i - st * rps;
        const int kk = rem / R;
        const int r = rem - kk * R;
        const int64_t s = j + kk;
        const uint32_t w = tb[pr.tb_off + ((st * nsteps + s) * L + kk) * wpc + (r * BITS) / 32];
        return (w >> ((r * BITS) % 32)) & ((1u << BITS) - 1u);
    };
    int32_t *P = path + 2 * path_off[out];
    P[0] = u;
    P[1] = v;
    int cnt = 1;
    while (u >= 0 && v >= 0) {
        const uint32_t c = cell(u, v);
        const uint32_t src = c & 3u;
        if (src == 0) break;  // local: D(u, v) is not above zero
        if (src == 1) {
            --u;
            --v;
        } else if (src == 2) {
            if constexpr (GAP == GAP_AFF) {
                int kk = u;
                while (kk > 0 && (cell(kk, v) & 4u)) --kk;
                u = kk - 1;
            } else {
                --u;
            }
        } else {
            if constexpr (GAP == GAP_AFF) {
                int kk = v;
                while (kk > 0 && (cell(u, kk) & 8u)) --kk;
                v = kk - 1;
            } else {
                --v;
            }
        }
        P[2 * cnt] = u;
        P[2 * cnt + 1] = v;
        ++cnt;
    }
    path_len[out] = cnt;
}

hipError_t launch_walk(int gap, const Prob *probs, int64_t n, const uint32_t *tb, const int32_t *seeds,
                       int32_t *path, const int64_t *path_off, int32_t *path_len, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 grid(unsigned((n + 255) / 256)), block(256);
    if (gap == GAP_AFF)
        hipLaunchKernelGGL(pa_walk_kernel<GAP_AFF>, grid, block, 0, st, probs, n, tb, seeds, path, path_off,
                           path_len);
    else  // linear and DTW share the 2-bit moves
        hipLaunchKernelGGL(pa_walk_kernel<GAP_LIN>, grid, block, 0, st, probs, n, tb, seeds, path, path_off,
                           path_len);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Backward: one thread per problem walks its optimal path (same moves as
// pa_walk_kernel) and emits the gradients of the optimal score.  The score
// is a sum along the path: S(u, v) for every cell left by a diagonal move
// (the alignment's matched cells; path_compressor, solver.h:831-846), the
// cost of every path cell for DTW (solver.h:3236-3248), and -w(k) per gap
// run with w(k) = open + extend*(k - 1) (pyalign/gaps.py:164-176), including
// the border run that ends a global path at (u, -1) or (-1, v) (the border
// values w(u+1), solver.h:2943-2951).  Every problem writes only its own
// region, so no atomics are needed.
// ---------------------------------------------------------------------------
template <int GAP, int FORM, int SD>
__global__ __launch_bounds__(256) void pa_backward_kernel(BwArgs a) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= a.n) return;
    const Prob pr = a.probs[k];
    const int out = pr.out;
    const float g = a.grad_scores[out];
    int u = a.seeds[2 * out], v = a.seeds[2 * out + 1];
    float opens = 0.0f, extends = 0.0f;
    constexpr int BITS = GAP == GAP_AFF ? 4 : 2;
    const int L = pr.L, R = pr.R, lt = pr.len_t;
    const int rps = L * R;
    const int wpc = (BITS * R + 31) / 32;
    const int64_t nsteps = lt + L - 1;
    auto cell = [&](int i, int j) -> uint32_t {
        const int st = i / rps;
        const int rem = i - st * rps;
        const int kk = rem / R;
        const int r = rem - kk * R;
        const int64_t s = j + kk;
        const uint32_t w = a.tb[pr.tb_off + ((st * nsteps + s) * L + kk) * wpc + (r * BITS) / 32];
        return (w >> ((r * BITS) % 32)) & ((1u << BITS) - 1u);
    };
    auto cost_grad = [&](int i, int j) {
        if constexpr (FORM == F_DENSE) {
            if (a.grad_dense) a.grad_dense[pr.off_s + int64_t(i) * pr.si + int64_t(j) * pr.sj] += g;
        } else if constexpr (FORM == F_SERIES) {
#pragma unroll
            for (int d = 0; d < SD; ++d) {
                const float diff = a.xs[(pr.off_s + i) * SD + d] - a.xt[(pr.off_t + j) * SD + d];
                const float gd = g * (2.0f * diff);
                if (a.grad_xs) a.grad_xs[(pr.off_s + i) * SD + d] += gd;
                if (a.grad_xt) a.grad_xt[(pr.off_t + j) * SD + d] -= gd;
            }
        }
    };
    if (u != NO_SEED) {
        while (u >= 0 && v >= 0) {
            const uint32_t src = cell(u, v) & 3u;
            if constexpr (GAP == GAP_DTW) cost_grad(u, v);
            if (src == 0) break;  // local: the path starts at a zero cell
            if (src == 1) {
                if constexpr (GAP != GAP_DTW) cost_grad(u, v);
                --u;
                --v;
            } else if (src == 2) {
                if constexpr (GAP == GAP_AFF) {
                    int kk = u;
                    while (kk > 0 && (cell(kk, v) & 4u)) --kk;
                    opens += 1.0f;
                    extends += float(u - kk);
                    u = kk - 1;
                } else {
                    opens += 1.0f;  // linear: one gap symbol
                    --u;
                }
            } else {
                if constexpr (GAP == GAP_AFF) {
                    int kk = v;
                    while (kk > 0 && (cell(u, kk) & 8u)) --kk;
                    opens += 1.0f;
                    extends += float(v - kk);
                    v = kk - 1;
                } else {
                    opens += 1.0f;
                    --v;
                }
            }
        }
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
        a.grad_gap[2 * out] += a.gsign * g * opens;
        a.grad_gap[2 * out + 1] += a.gsign * g * extends;
    }
}

hipError_t launch_backward(int gap, int form, int dim, const BwArgs &args, hipStream_t st) {
    if (args.n == 0) return hipSuccess;
    const dim3 grid(unsigned((args.n + 255) / 256)), block(256);
#define PA_BW(G, F, D) hipLaunchKernelGGL((pa_backward_kernel<G, F, D>), grid, block, 0, st, args)
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

__global__ void pa_zero_u32_kernel(uint32_t *p, int64_t n) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0u;
}

hipError_t launch_zero_u32(uint32_t *p, int64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pa_zero_u32_kernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

}  // namespace pa
