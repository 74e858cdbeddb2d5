// realize.hip -- discrete realizations of the binary population and the
// single-source / background split of h_c^2.
//
// <- holodeck/cyutils.pyx: _loudest_hc_from_sorted (1266-1344),
//    _loudest_hc_and_par_from_sorted_redz (1615-1767), _ss_bg_hc (935-1014),
//    _ss_bg_hc_and_par (1068-1178), _sam_poisson_gwb (862-897).
//
// Per (frequency f, realization r) the reference draws a count n_b for every
// bin b, fills L loud slots from bins in a fixed priority order (with
// multiplicity) and sums n_b*h_b over the rest.  This file takes the counts as
// given -- an injected (B,F,R) array, the exact parity path -- and holds the
// merge shared with the drawn-count kernels of realize_gap.hip:
//   * a workgroup handles one frequency and one chunk of bins, one thread per
//     realization; bin data (h, priority key, parameters) are staged through
//     LDS and shared by all realizations;
//   * each thread keeps the L best instances of its chunk in registers
//     (priority = order rank for the sorted walk, (h desc, bin asc) for argmax)
//     and the FP64 totals sum(n h), sum(n h p_k);
//   * a merge kernel combines the chunk lists per (f, r); the background is
//     total - sum(selected h), which is exactly the reference's accounting
//     (sum(n h) - max for ss_bg_hc; sum((n - taken) h) for the sorted walk,
//     including its treatment of non-integer Normal draws).
#include <type_traits>
#include "common.hpp"
#include "rng.hpp"
#include "../../include/holohip.h"

namespace hd {

constexpr int LMAX = 16;
constexpr int NPAR_MAX = 8;
constexpr int RZ_CHUNK = 16384;   // bins per workgroup
constexpr int RZ_SUB = 256;       // bins staged in LDS at a time
constexpr uint64_t KEY_NONE = ~0ull;

struct RealizeArgs {
    const double* counts;    // (B,F,R) injected counts
    const double* h2fdf;     // (B,F)
    const int32_t* rank;     // (B) position in order (SORTED) or null
    const double* par;       // (npar,B,F) or null
    int64_t B;
    int F, R, L, mode, flags, npar, nchunks;
    const int64_t* order;    // (B) bin at each position of the sorted walk (final merge, SORTED)
    // partials.  The sorted walk keeps only 32-bit keys (its order ranks: unique, so no bin
    // tie-break, and the bin is order[rank]) in p_key viewed as uint32_t; argmax keeps 64-bit
    // keys and the bins.
    double* p_total;      // (F, nchunks, R)
    double* p_ptot;       // (npar, F, nchunks, R)
    uint64_t* p_key;      // (F, nchunks, L, R): realization fastest
    int32_t* p_bin;       // (F, nchunks, L, R), argmax only
};

constexpr uint32_t KEY32_NONE = 0xffffffffu;

// store / load one slot of a partial list in the mode's layout
__device__ __forceinline__ void part_store(const RealizeArgs& A, int64_t o, uint64_t k, int32_t b) {
    if (A.mode == HD_SEL_SORTED) {
        ((uint32_t*)A.p_key)[o] = (k == KEY_NONE) ? KEY32_NONE : (uint32_t)k;
    } else {
        A.p_key[o] = k;
        A.p_bin[o] = b;
    }
}

// ascending image of a double, so that larger h -> smaller key when inverted
__device__ __forceinline__ uint64_t h_priority(double h) {
    uint64_t u = (uint64_t)__double_as_longlong(h);
    u = (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
    return ~u;
}

__device__ __forceinline__ bool better(uint64_t ka, int32_t ba, uint64_t kb, int32_t bb) {
    return (ka < kb) || (ka == kb && ba < bb);
}

// insert one instance into a sorted register list of LMAX slots
__device__ __forceinline__ void list_insert(uint64_t (&lk)[LMAX], int32_t (&lb)[LMAX], uint64_t k, int32_t b) {
#pragma unroll
    for (int i = 0; i < LMAX; i++) {
        bool sw = better(k, b, lk[i], lb[i]);
        uint64_t tk = lk[i];
        int32_t tb = lb[i];
        lk[i] = sw ? k : tk;
        lb[i] = sw ? b : tb;
        k = sw ? tk : k;
        b = sw ? tb : b;
    }
}

__global__ void __launch_bounds__(512) k_realize_chunk(RealizeArgs A) {
    __shared__ double s_h[RZ_SUB];
    __shared__ uint64_t s_key[RZ_SUB];
    __shared__ double s_par[NPAR_MAX][RZ_SUB];

    const int f = blockIdx.y;
    const int chunk = blockIdx.x;
    const int64_t b0 = (int64_t)chunk * RZ_CHUNK;
    const int64_t b1 = min(b0 + (int64_t)RZ_CHUNK, A.B);
    const int F = A.F, R = A.R, L = A.L;
    const bool use_list = (A.mode != HD_SEL_NONE);
    const bool sorted_mode = (A.mode == HD_SEL_SORTED);
    const bool skip_zero_h = (A.flags & HD_FLAG_SKIP_ZERO_H) != 0;

    for (int r0 = 0; r0 < R; r0 += blockDim.x) {
        const int r = r0 + threadIdx.x;
        const bool active = r < R;
        uint64_t lk[LMAX];
        int32_t lb[LMAX];
#pragma unroll
        for (int i = 0; i < LMAX; i++) {
            lk[i] = KEY_NONE;
            lb[i] = 0x7fffffff;
        }
        double total = 0.0;
        double ptot[NPAR_MAX];
#pragma unroll
        for (int k = 0; k < NPAR_MAX; k++) ptot[k] = 0.0;

        for (int64_t s0 = b0; s0 < b1; s0 += RZ_SUB) {
            const int ns = (int)min((int64_t)RZ_SUB, b1 - s0);
            __syncthreads();
            for (int c = threadIdx.x; c < ns; c += blockDim.x) {
                int64_t b = s0 + c;
                double h = A.h2fdf[b * F + f];
                s_h[c] = h;
                s_key[c] = sorted_mode ? (uint64_t)A.rank[b] : h_priority(h);
                for (int k = 0; k < A.npar; k++) s_par[k][c] = A.par[((int64_t)k * A.B + b) * F + f];
            }
            __syncthreads();
            if (!active) continue;
            for (int c = 0; c < ns; c++) {
                const int64_t b = s0 + c;
                const double n = A.counts[(b * F + f) * R + r];
                const double h = s_h[c];
                if (sorted_mode) {
                    if (n < 1) continue;
                    if (skip_zero_h && h == 0) continue;
                } else {
                    if (n == 0) continue;
                }
                const double nh = n * h;
                total += nh;
                for (int k = 0; k < A.npar; k++) ptot[k] += nh * s_par[k][c];
                if (use_list && n > 0) {
                    if (!sorted_mode && !(h > 0)) continue;   // argmax: strict cur > max >= 0
                    const uint64_t key = s_key[c];
                    const int32_t bb = (int32_t)b;
                    // the same (key, bin) for every instance of this bin
                    int ninst = (int)min(ceil(n), (double)L);
                    for (int i = 0; i < ninst; i++) {
                        // compare against slot L-1 (dynamic index avoided: scan)
                        uint64_t wk = KEY_NONE;
                        int32_t wb = 0x7fffffff;
#pragma unroll
                        for (int j = 0; j < LMAX; j++)
                            if (j == L - 1) { wk = lk[j]; wb = lb[j]; }
                        if (!better(key, bb, wk, wb)) break;
                        list_insert(lk, lb, key, bb);
                    }
                }
            }
        }
        if (active) {
            int64_t pr = ((int64_t)f * A.nchunks + chunk) * R + r;
            A.p_total[pr] = total;
            for (int k = 0; k < A.npar; k++) A.p_ptot[(int64_t)k * F * A.nchunks * R + pr] = ptot[k];
            if (use_list) {
#pragma unroll
                for (int j = 0; j < LMAX; j++)
                    if (j < L) {
                        const int64_t o = ((pr / R) * L + j) * R + (pr % R);
                        part_store(A, o, lk[j], lb[j]);
                    }
            }
        }
    }
}

// Running list of the merges: LM slots (LM = L for the common L = 1, 5, 10; 16 slots holding
// any L <= 16 otherwise).  S32: the sorted walk's 32-bit ranks (unique, no bins); else 64-bit
// keys with the bin as tie-break (argmax).
template <int LM, bool S32>
struct MergeList {
    using Key = typename std::conditional<S32, uint32_t, uint64_t>::type;
    static constexpr Key NONE = S32 ? (Key)KEY32_NONE : (Key)KEY_NONE;
    Key k[LM];
    int32_t b[LM];
    __device__ __forceinline__ void init() {
#pragma unroll
        for (int i = 0; i < LM; i++) {
            k[i] = NONE;
            b[i] = 0x7fffffff;
        }
    }
    __device__ __forceinline__ static bool better(Key ka, int32_t ba, Key kb, int32_t bb) {
        return S32 ? (ka < kb) : ((ka < kb) || (ka == kb && ba < bb));
    }
    // (key, bin) of slot L - 1, the worst one kept
    __device__ __forceinline__ void worst(int L, Key& wk, int32_t& wb) const {
        if (LM != 16) {
            wk = k[LM - 1];
            wb = b[LM - 1];
            return;
        }
        wk = NONE;
        wb = 0x7fffffff;
#pragma unroll
        for (int q = 0; q < LM; q++)
            if (q == L - 1) {
                wk = k[q];
                wb = b[q];
            }
    }
    __device__ __forceinline__ void insert(Key key, int32_t bin) {
#pragma unroll
        for (int i = 0; i < LM; i++) {
            const bool sw = better(key, bin, k[i], b[i]);
            const Key tk = k[i];
            const int32_t tb = b[i];
            k[i] = sw ? key : tk;
            b[i] = sw ? bin : tb;
            key = sw ? tk : key;
            bin = sw ? tb : bin;
        }
    }
};

// Merge of chunk partials [c0, c1) of (f, r) into a running list, total and parameter totals.
template <int LM, bool S32>
__device__ __forceinline__ void merge_chunks(const RealizeArgs& A, int f, int r, int c0, int c1,
                                             MergeList<LM, S32>& list, double& total, double (&ptot)[NPAR_MAX]) {
    using Key = typename MergeList<LM, S32>::Key;
    const int F = A.F, R = A.R, NC = A.nchunks;
    const int L = (LM == 16) ? A.L : LM;
    const bool use_list = (A.mode != HD_SEL_NONE);
    for (int c = c0; c < c1; c++) {
        int64_t pr = ((int64_t)f * NC + c) * R + r;
        total += A.p_total[pr];
        for (int k = 0; k < A.npar; k++) ptot[k] += A.p_ptot[(int64_t)k * F * NC * R + pr];
        if (use_list) {
            // the chunk's whole list first: every load issued before any is used (indices clamped
            // to slot L-1, no branch), so a chunk costs one load latency instead of one per slot
            const int64_t o0 = ((int64_t)f * NC + c) * L * (int64_t)R + r;
            Key ck[LM];
            int32_t cb[LM];
#pragma unroll
            for (int j = 0; j < LM; j++) {
                const int64_t o = o0 + (int64_t)min(j, L - 1) * R;
                if (S32) {
                    ck[j] = (Key)((const uint32_t*)A.p_key)[o];
                    cb[j] = 0;   // ranks are unique: no tie-break
                } else {
                    ck[j] = (Key)A.p_key[o];
                    cb[j] = A.p_bin[o];
                }
            }
#pragma unroll
            for (int j = 0; j < LM; j++) {
                if (j >= L || ck[j] == MergeList<LM, S32>::NONE) break;
                Key wk;
                int32_t wb;
                list.worst(L, wk, wb);
                if (!MergeList<LM, S32>::better(ck[j], cb[j], wk, wb)) break;
                list.insert(ck[j], cb[j]);
            }
        }
    }
}

constexpr int MERGE_GROUPS = 8;

// First merge stage: thread per (f, group, r) merges the group's chunks and writes a partial of
// the same layout with G "chunks", so that the final merge reads G partials instead of NC (the
// final merge alone has one thread per (f, r): too few waves to hide its load latency).
template <int LM, bool S32>
__global__ void __launch_bounds__(256) k_realize_group(RealizeArgs A, int G, double* __restrict__ q_total,
                                                       double* __restrict__ q_ptot, uint64_t* __restrict__ q_key,
                                                       int32_t* __restrict__ q_bin) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int F = A.F, R = A.R, NC = A.nchunks;
    const int L = (LM == 16) ? A.L : LM;
    if (t >= (int64_t)F * G * R) return;
    const int r = (int)(t % R), g = (int)((t / R) % G), f = (int)(t / ((int64_t)G * R));
    const int CG = (NC + G - 1) / G;
    const int c0 = min(g * CG, NC), c1 = min(c0 + CG, NC);
    MergeList<LM, S32> list;
    list.init();
    double total = 0.0;
    double ptot[NPAR_MAX];
#pragma unroll
    for (int k = 0; k < NPAR_MAX; k++) ptot[k] = 0.0;
    merge_chunks<LM, S32>(A, f, r, c0, c1, list, total, ptot);
    const int64_t pr = ((int64_t)f * G + g) * R + r;
    q_total[pr] = total;
    for (int k = 0; k < A.npar; k++) q_ptot[(int64_t)k * F * G * R + pr] = ptot[k];
    if (A.mode != HD_SEL_NONE) {
#pragma unroll
        for (int j = 0; j < LM; j++)
            if (j < L) {
                const int64_t o = (((int64_t)f * G + g) * L + j) * R + r;
                if (S32) {
                    ((uint32_t*)q_key)[o] = (uint32_t)list.k[j];
                } else {
                    q_key[o] = (uint64_t)list.k[j];
                    q_bin[o] = list.b[j];
                }
            }
    }
}

template <int LM, bool S32>
__global__ void __launch_bounds__(256) k_realize_merge(RealizeArgs A, double* __restrict__ hc2ss,
                                                       int64_t* __restrict__ ssidx, double* __restrict__ hc2bg,
                                                       double* __restrict__ bgpar) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int F = A.F, R = A.R, NC = A.nchunks;
    const int L = (LM == 16) ? A.L : LM;
    if (t >= (int64_t)F * R) return;
    const int f = (int)(t / R), r = (int)(t % R);
    const bool use_list = (A.mode != HD_SEL_NONE);
    MergeList<LM, S32> list;
    list.init();
    double total = 0.0;
    double ptot[NPAR_MAX];
#pragma unroll
    for (int k = 0; k < NPAR_MAX; k++) ptot[k] = 0.0;
    merge_chunks<LM, S32>(A, f, r, 0, NC, list, total, ptot);
    double sel = 0.0;
    double psel[NPAR_MAX];
#pragma unroll
    for (int k = 0; k < NPAR_MAX; k++) psel[k] = 0.0;
    if (use_list) {
#pragma unroll
        for (int j = 0; j < LM; j++) {
            if (j >= L) continue;
            int64_t o = ((int64_t)f * R + r) * L + j;
            if (list.k[j] == MergeList<LM, S32>::NONE) {
                hc2ss[o] = 0.0;
                ssidx[o] = -1;
            } else {
                const int64_t b = S32 ? A.order[list.k[j]] : (int64_t)list.b[j];
                double h = A.h2fdf[b * F + f];
                hc2ss[o] = h;
                ssidx[o] = b;
                sel += h;
                for (int k = 0; k < A.npar; k++) psel[k] += h * A.par[((int64_t)k * A.B + b) * F + f];
            }
        }
    }
    double bg = total - sel;
    hc2bg[(int64_t)f * R + r] = bg;
    if (bgpar) {
        bool none = use_list && (list.k[0] == MergeList<LM, S32>::NONE) && (A.mode == HD_SEL_ARGMAX);
        for (int k = 0; k < A.npar; k++) {
            double v = none ? __builtin_nan("") : (ptot[k] - psel[k]) / bg;
            bgpar[((int64_t)k * F + f) * R + r] = v;
        }
    }
}

__global__ void k_rank_from_order(const int64_t* __restrict__ order, int64_t B, int32_t* __restrict__ rank) {
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= B) return;
    rank[order[p]] = (int32_t)p;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t realize_group_ws(int F, int R, int L, int npar);
int realize_merge_launch(const double* h2fdf, const double* par, const int64_t* order, int64_t B, int F, int R, int L,
                         int mode, int npar, int nchunks, const double* p_total, const double* p_ptot, const uint64_t* p_key,
                         const int32_t* p_bin, double* hc2ss, int64_t* ssidx, double* hc2bg, double* bgpar,
                         void* group_ws, hipStream_t st);

static size_t realize_ws(int64_t B, int F, int R, int L, int npar) {
    int64_t nch = cdiv(B, RZ_CHUNK);
    size_t pr = (size_t)F * nch * R;
    size_t s = 0;
    s += align256((size_t)B * 4);               // rank
    s += align256(pr * 8);                      // totals
    s += align256(pr * 8 * (size_t)npar);       // param totals
    s += align256(pr * 8 * (size_t)L);          // keys
    s += align256(pr * 4 * (size_t)L);          // bins
    s += align256(realize_group_ws(F, R, L, npar));   // first merge stage
    return s + 256;
}

static int realize_counts_impl(const double* counts, const double* h2fdf, const int64_t* order, int64_t B, int F,
                               int R, int L, int mode, int flags, const double* par, int npar, double* hc2ss,
                               int64_t* ssidx, double* hc2bg, double* bgpar, void* ws, size_t ws_bytes,
                               hipStream_t st) {
    if (B < 1 || F < 1 || R < 1 || L < 0 || !h2fdf || !hc2bg) return set_error(HD_ERR_ARG, "realize: bad sizes/pointers");
    if (B > 0x7fffffffll) return set_error(HD_ERR_SIZE, "realize: B=%lld exceeds 2^31-1", (long long)B);
    if (F > 65535) return set_error(HD_ERR_SIZE, "realize: F=%d > 65535", F);
    if (mode != HD_SEL_SORTED && mode != HD_SEL_ARGMAX && mode != HD_SEL_NONE)
        return set_error(HD_ERR_ARG, "realize: unknown mode %d", mode);
    if (mode == HD_SEL_NONE) L = 0;
    else if (L < 1) return set_error(HD_ERR_ARG, "realize: L must be >= 1 for loud-source modes");
    if (L > LMAX) return set_error(HD_ERR_SIZE, "realize: L=%d > %d", L, LMAX);
    if (npar < 0 || npar > NPAR_MAX || (npar > 0 && !par)) return set_error(HD_ERR_ARG, "realize: npar=%d invalid", npar);
    if (mode == HD_SEL_SORTED && !order) return set_error(HD_ERR_ARG, "realize: sorted mode needs an order");
    if (mode != HD_SEL_NONE && (!hc2ss || !ssidx)) return set_error(HD_ERR_ARG, "realize: hc2ss/ssidx required");
    if (!counts) return set_error(HD_ERR_ARG, "realize: counts required");
    size_t need = realize_ws(B, F, R, L, npar);
    if (!ws || ws_bytes < need) return set_error(HD_ERR_SIZE, "realize: workspace too small (%zu < %zu)", ws_bytes, need);

    RealizeArgs A;
    A.counts = counts;
    A.h2fdf = h2fdf;
    A.par = par;
    A.B = B;
    A.F = F;
    A.R = R;
    A.L = L;
    A.mode = mode;
    A.flags = flags;
    A.npar = npar;
    A.nchunks = (int)cdiv(B, RZ_CHUNK);
    size_t pr = (size_t)F * A.nchunks * R;
    char* p = (char*)(((uintptr_t)ws + 255) & ~(uintptr_t)255);
    int32_t* rank = (int32_t*)p;  p += align256((size_t)B * 4);
    A.p_total = (double*)p;       p += align256(pr * 8);
    A.p_ptot = (double*)p;        p += align256(pr * 8 * (size_t)npar);
    A.p_key = (uint64_t*)p;       p += align256(pr * 8 * (size_t)L);
    A.p_bin = (int32_t*)p;        p += align256(pr * 4 * (size_t)L);
    void* group_ws = p;
    A.rank = (mode == HD_SEL_SORTED) ? rank : nullptr;
    A.order = order;

    if (mode == HD_SEL_SORTED) {
        hipLaunchKernelGGL(k_rank_from_order, dim3((unsigned)cdiv(B, 256)), dim3(256), 0, st, order, B, rank);
        int rc = check_launch("k_rank_from_order");
        if (rc) return rc;
    }
    int threads = (int)min((int64_t)512, cdiv(R, 64) * 64);
    dim3 grid(A.nchunks, F);
    hipLaunchKernelGGL(k_realize_chunk, grid, dim3(threads), 0, st, A);
    int rc = check_launch("k_realize_chunk");
    if (rc) return rc;
    // the same two-stage merge as the drawn-count kernels (groups of chunks first when there are many)
    return realize_merge_launch(h2fdf, par, order, B, F, R, L, mode, npar, A.nchunks, A.p_total, A.p_ptot, A.p_key,
                                A.p_bin, hc2ss, ssidx, hc2bg, (npar > 0) ? bgpar : nullptr, group_ws, st);
}

// Merge of per-chunk partials, shared with the drawn-count kernel (realize_gap.hip).
size_t realize_group_ws(int F, int R, int L, int npar) {
    return (size_t)F * MERGE_GROUPS * R * (sizeof(double) * (1 + (size_t)npar) + (sizeof(uint64_t) + sizeof(int32_t)) * (size_t)L) +
           4 * 256;
}

// the two merge stages for list capacity LM and key width (S32: the sorted walk's ranks)
template <int LM, bool S32>
static int merge_launch_k(RealizeArgs A, void* group_ws, double* hc2ss, int64_t* ssidx, double* hc2bg, double* bgpar,
                          hipStream_t st) {
    const int F = A.F, R = A.R, L = A.L, npar = A.npar;
    if (group_ws && A.nchunks > 2 * MERGE_GROUPS) {
        const int G = MERGE_GROUPS;
        const size_t n = (size_t)F * G * R;
        char* w = (char*)(((uintptr_t)group_ws + 255) & ~(uintptr_t)255);
        double* q_total = (double*)w;         w += n * sizeof(double);
        double* q_ptot = (double*)w;          w += n * sizeof(double) * (size_t)npar;
        uint64_t* q_key = (uint64_t*)w;       w += n * sizeof(uint64_t) * (size_t)L;
        int32_t* q_bin = (int32_t*)w;
        hipLaunchKernelGGL((k_realize_group<LM, S32>), dim3((unsigned)cdiv((int64_t)n, 256)), dim3(256), 0, st, A, G,
                           q_total, q_ptot, q_key, q_bin);
        int rc = check_launch("k_realize_group");
        if (rc) return rc;
        A.nchunks = G;
        A.p_total = q_total;
        A.p_ptot = q_ptot;
        A.p_key = q_key;
        A.p_bin = q_bin;
    }
    hipLaunchKernelGGL((k_realize_merge<LM, S32>), dim3((unsigned)cdiv((int64_t)F * R, 256)), dim3(256), 0, st, A, hc2ss,
                       ssidx, hc2bg, bgpar);
    return check_launch("k_realize_merge");
}

int realize_merge_launch(const double* h2fdf, const double* par, const int64_t* order, int64_t B, int F, int R, int L,
                         int mode, int npar, int nchunks, const double* p_total, const double* p_ptot, const uint64_t* p_key,
                         const int32_t* p_bin, double* hc2ss, int64_t* ssidx, double* hc2bg, double* bgpar,
                         void* group_ws, hipStream_t st) {
    RealizeArgs A;
    A.counts = nullptr;
    A.h2fdf = h2fdf;
    A.rank = nullptr;
    A.order = order;
    A.par = par;
    A.B = B;
    A.F = F;
    A.R = R;
    A.L = L;
    A.mode = mode;
    A.flags = 0;
    A.npar = npar;
    A.nchunks = nchunks;
    A.p_total = const_cast<double*>(p_total);
    A.p_ptot = const_cast<double*>(p_ptot);
    A.p_key = const_cast<uint64_t*>(p_key);
    A.p_bin = const_cast<int32_t*>(p_bin);
    if (mode == HD_SEL_SORTED) {
        if (L == 1) return merge_launch_k<1, true>(A, group_ws, hc2ss, ssidx, hc2bg, bgpar, st);
        if (L == 5) return merge_launch_k<5, true>(A, group_ws, hc2ss, ssidx, hc2bg, bgpar, st);
        if (L == 10) return merge_launch_k<10, true>(A, group_ws, hc2ss, ssidx, hc2bg, bgpar, st);
        return merge_launch_k<16, true>(A, group_ws, hc2ss, ssidx, hc2bg, bgpar, st);
    }
    if (mode == HD_SEL_ARGMAX) {
        if (L == 1) return merge_launch_k<1, false>(A, group_ws, hc2ss, ssidx, hc2bg, bgpar, st);
        if (L == 5) return merge_launch_k<5, false>(A, group_ws, hc2ss, ssidx, hc2bg, bgpar, st);
        if (L == 10) return merge_launch_k<10, false>(A, group_ws, hc2ss, ssidx, hc2bg, bgpar, st);
        return merge_launch_k<16, false>(A, group_ws, hc2ss, ssidx, hc2bg, bgpar, st);
    }
    return merge_launch_k<1, false>(A, group_ws, hc2ss, ssidx, hc2bg, bgpar, st);   // no lists
}

int rank_from_order_launch(const int64_t* order, int64_t B, int32_t* rank, hipStream_t st) {
    hipLaunchKernelGGL(k_rank_from_order, dim3((unsigned)cdiv(B, 256)), dim3(256), 0, st, order, B, rank);
    return check_launch("k_rank_from_order");
}

}  // namespace hd

using namespace hd;

extern "C" size_t hd_realize_counts_workspace_size(int64_t B, int F, int R, int L, int npar) {
    return realize_ws(B, F, R, L, npar);
}

extern "C" int hd_realize_counts(const double* counts, const double* h2fdf, const int64_t* order, int64_t B, int F, int R,
                                 int L, int mode, int flags, const double* par, int npar, double* hc2ss,
                                 int64_t* ssidx, double* hc2bg, double* bgpar, void* workspace, size_t ws_bytes,
                                 void* stream) {
    return realize_counts_impl(counts, h2fdf, order, B, F, R, L, mode, flags, par, npar, hc2ss, ssidx, hc2bg, bgpar,
                               workspace, ws_bytes, (hipStream_t)stream);
}

extern "C" void hd_philox4x32_10(const uint32_t* ctr, const uint32_t* key, uint32_t* out) {
    U4 c = {ctr[0], ctr[1], ctr[2], ctr[3]};
    U4 o = philox4x32_10(c, key[0], key[1]);
    out[0] = o.x;
    out[1] = o.y;
    out[2] = o.z;
    out[3] = o.w;
}
